"""CPU checks of the boundary: libedgpu.so loads and exports every entry
point include/edgpu.h declares, the Python mirror declares them all, and
host-side argument checking rejects bad input without touching a GPU."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "edgpu.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(edg_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = header_functions()
    for n in ("edg_hx", "edg_lanczos_iteration", "edg_ground_state", "edg_cluster_green",
              "edg_apply_c", "edg_observables", "edg_spin_table", "edg_create", "edg_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(edgpu):
    out = subprocess.check_output(["nm", "-D", "--defined-only", edgpu.LIB_PATH], text=True)
    exported = set(re.findall(r"\bT (edg_\w+)", out))
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing
    L = edgpu.load_library()
    for n in header_functions():
        assert hasattr(L, n)


def test_python_mirror_covers_header(edgpu):
    assert sorted(edgpu.EXPORTS) == header_functions()


def test_no_device_means_loud_failure(edgpu):
    if edgpu.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(edgpu.EdgpuError):
        edgpu.Solver(norb=1, nbath=3)


def test_create_rejects_arpack_and_bad_sizes(edgpu):
    import ctypes as C
    L = edgpu.load_library()
    for kw in (dict(diag_solver=2), dict(nbath=15), dict(nspin=3), dict(maxnstep=0)):
        p = edgpu._Params()
        L.edg_default_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        h = C.c_void_p()
        assert L.edg_create(C.byref(p), 0, C.byref(h)) == 3
