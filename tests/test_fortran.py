"""The Fortran side of the boundary: the iso_c_binding module compiles with
amdflang and the example caller (built by build()) solves config 1 on the
GPU through libedgpu.so with the golden fixture's numbers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "dmft-ed-zt_amd", "fortran")
FLANG = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")
def test_interface_module_compiles(tmp_path):
    subprocess.check_call([FLANG, "-c", os.path.join(FDIR, "edgpu_c.f90"), "-module-dir",
                           str(tmp_path), "-o", str(tmp_path / "m.o")])
    assert (tmp_path / "edgpu_c.mod").exists()


@pytest.mark.gpu
def test_fortran_caller_solves_config1():
    exe = os.path.join(FDIR, "example_solve")
    if not os.path.exists(exe):
        pytest.fail("example_solve not built: run __graft_entry__.build()")
    g = np.load(os.path.join(ROOT, "tests", "golden", "config1_ns6.npz"))
    inp = " ".join(f"{v:.17e}" for v in g["ek"][:, 0]) + "\n" + \
          " ".join(f"{v:.17e}" for v in g["vmk"][0, :, 0]) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {l.split()[0]: [float(t) for t in l.split()[1:]] for l in out.stdout.splitlines() if l.strip()}
    assert abs(vals["E0"][0] - float(g["E0"])) < 1e-10
    assert abs(vals["n_up"][0] - float(g["n_up"])) < 1e-10
    assert abs(vals["docc"][0] - float(g["docc"])) < 1e-10
    assert abs(complex(*vals["G0"]) - g["G"][0]) < 1e-9
