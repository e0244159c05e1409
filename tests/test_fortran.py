"""The Fortran side of the boundary: the iso_c_binding module compiles with
amdflang and the example caller (built by build()) solves the Ns=6 problem on the
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
def test_fortran_caller_solves_ns6():
    exe = os.path.join(FDIR, "example_solve")
    if not os.path.exists(exe):
        pytest.fail("example_solve not built: run __graft_entry__.build()")
    g = np.load(os.path.join(ROOT, "tests", "golden", "bethe_ns6.npz"))
    inp = " ".join(f"{v:.17e}" for v in g["ek"][:, 0]) + "\n" + \
          " ".join(f"{v:.17e}" for v in g["vmk"][0, :, 0]) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {l.split()[0]: [float(t) for t in l.split()[1:]] for l in out.stdout.splitlines() if l.strip()}
    assert abs(vals["E0"][0] - float(g["E0"])) < 1e-10
    assert abs(vals["n_up"][0] - float(g["n_up"])) < 1e-10
    assert abs(vals["docc"][0] - float(g["docc"])) < 1e-10
    assert abs(complex(*vals["G0"]) - g["G"][0]) < 1e-9


@pytest.mark.gpu
def test_fortran_batch_caller_vs_oracle():
    """Three problems solved in one batch from Fortran (edg_batch_* through
    iso_c_binding): E0 and G(i w_0) against the oracle, FP64 vectors."""
    from oracle import oracle
    oracle.build()
    exe = os.path.join(FDIR, "example_batch")
    if not os.path.exists(exe):
        pytest.fail("example_batch not built: run __graft_entry__.build()")
    rng = np.random.default_rng(12)
    probs, inp = [], ""
    for k in range(3):
        U = 1.0 + k
        e = rng.uniform(-1, 1, 5)
        v = rng.uniform(0.2, 0.6, 5)
        probs.append((U, e, v))
        inp += f"{U:.17e}\n" + " ".join(f"{x:.17e}" for x in e) + "\n" + " ".join(f"{x:.17e}" for x in v) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, timeout=120, check=True).stdout
    rows = [line.split() for line in out.splitlines() if line.startswith("P ")]
    assert len(rows) == 3
    om = oracle.matsubara(64, 100.0)
    for row in rows:
        k = int(row[1]) - 1
        U, e, v = probs[k]
        em = np.zeros((1, 2))
        ek = np.stack([e, e], axis=1)
        vmk = np.stack([v, v], axis=1)[None, :, :]
        m = oracle.Model(1, 5, U=U, mu=U / 2)
        gs = oracle.ground_state(m, em, ek, vmk, [(3, 3)], 80, seed=21 + k)
        G, _ = oracle.cluster_green_ftn(m, em, ek, vmk, gs, 80, om)
        assert abs(float(row[2]) - gs.val) < 1e-9
        assert abs(complex(float(row[3]), float(row[4])) - G[0, 0, 0]) < 1e-8
