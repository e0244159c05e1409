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


DROPIN = os.path.join(ROOT, "tests", "fortran")


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")
def test_dropin_compiles_against_reference_interfaces(tmp_path):
    """edgpu_dropin.f90 (the reference's argument lists) compiles against
    stand-ins of the reference modules it uses (tests/fortran/refmods.f90)."""
    for src in (os.path.join(DROPIN, "refmods.f90"), os.path.join(FDIR, "edgpu_c.f90"),
                os.path.join(FDIR, "edgpu_dropin.f90")):
        subprocess.check_call([FLANG, "-c", src, "-module-dir", str(tmp_path), "-o",
                               str(tmp_path / (os.path.basename(src) + ".o"))], cwd=tmp_path)
    assert (tmp_path / "edgpu_dropin.mod").exists()


@pytest.mark.gpu
def test_fortran_dropin_vs_golden_and_oracle(tmp_path, oracle):
    """The drop-ins called from Fortran with the reference's argument lists
    (ground_state, cluster_green_ftn, hx, lanczos_iteration) on the Ns=6
    golden problem, FP64: E0, n_up, docc from gs%vec within 1e-10 and G(i w)
    within 1e-9 of the dense-ED fixture; hx within 1e-12 and the first eight
    Lanczos coefficients within 1e-10 of the oracle; g_coeffs.dat in the
    reference layout with the oracle's leading coefficients (1e-10); and a
    G computed from a gs argument that is not the ground state held on the
    device (another bath solved in between) equal to the first (1e-12); with
    edgpu_gs_to_host = .false. an empty gs%vec, the device's occupancies
    (1e-10 of the fixture), G from the held vector equal to the first
    (1e-12), and Sigma and Z against the host formula (1e-10)."""
    exe = os.path.join(DROPIN, "dropin_driver")
    if not os.path.exists(exe):
        pytest.fail("dropin_driver not built: run __graft_entry__.build()")
    g = np.load(os.path.join(ROOT, "tests", "golden", "bethe_ns6.npz"))
    rng = np.random.default_rng(66)
    ek2, v2 = rng.uniform(-1, 1, 5), rng.uniform(0.2, 0.6, 5)
    f = lambda a: " ".join(f"{v:.17e}" for v in a)
    inp = (f"{float(g['U']):.17e} {float(g['mu']):.17e} 100.0 2048\n" + f(g["ek"][:, 0]) + "\n" +
           f(g["vmk"][0, :, 0]) + "\n" + f(ek2) + "\n" + f(v2) + "\n")
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.strip()]
    val = lambda k: [[float(t) for t in r[1:]] for r in rows if r[0] == k]
    assert abs(val("E0")[0][1] - float(g["E0"])) < 1e-10
    assert abs(val("n_up")[0][0] - float(g["n_up"])) < 1e-10
    assert abs(val("docc")[0][0] - float(g["docc"])) < 1e-10
    for i, re, im in val("G"):
        assert abs(complex(re, im) - g["G"][int(i) - 1]) < 1e-9, i
    m = oracle.Model(1, 5, U=float(g["U"]), mu=float(g["mu"]))
    b = oracle.Basis(6, 3, 3)
    x = np.cos(0.1 * np.arange(1, b.ntot + 1))
    hx = oracle.hx(m, g["em"], g["ek"], g["vmk"], b, x)
    for r, y in val("HX"):
        assert abs(y - hx[int(r) - 1]) < 1e-12 * np.abs(hx).max()
    ns, a, bb = oracle.lanczos_iteration(m, g["em"], g["ek"], g["vmk"], b, x, 200)
    assert int(val("NSTEP")[0][0]) == ns
    for i, ai, bi in val("AB"):
        assert abs(ai - a[int(i) - 1]) < 1e-10 and abs(bi - bb[int(i) - 1]) < 1e-10
    assert val("GDIFF")[0][0] < 1e-12
    # edgpu_gs_to_host = .false.: gs%vec empty, the vector stays on the
    # device; occupancies and Sigma/Z from the device (post_process.f90:
    # 40-58,63-112,121-185) against the fixture and the host formula on the
    # driver's own G
    size, e0c = val("E0C")[0]
    assert int(size) == 0 and abs(e0c - float(g["E0"])) < 1e-10
    n_up, n_dn, dd = val("OBS")[0]
    assert abs(n_up - float(g["n_up"])) < 1e-10 and abs(n_dn - float(g["n_up"])) < 1e-10
    assert abs(dd - float(g["docc"])) < 1e-10
    assert val("GDIFFC")[0][0] < 1e-12
    om = np.pi * (2 * np.arange(2048) + 1) / 100.0
    Gd = {int(i): complex(re, im) for i, re, im in val("G")}
    ek1, v1 = g["ek"][:, 0], g["vmk"][0, :, 0]
    for i, re, im in val("SIG"):
        w = 1j * om[int(i) - 1]
        hyb = np.sum(v1 ** 2 / (w - ek1))
        want = w + float(g["mu"]) - float(g["em"][0, 0]) - hyb - 1.0 / Gd[int(i)]
        assert abs(complex(re, im) - want) < 1e-10 * max(1.0, abs(want)), i
    s1 = complex(*val("SIG")[0][1:])
    assert abs(val("Z")[0][0] - 1.0 / (1.0 - s1.imag / om[0])) < 1e-10
    # g_coeffs.dat (dmft_green.f90:66-69,89-109)
    lines = (tmp_path / "g_coeffs.dat").read_text().split("\n")
    assert [int(t) for t in lines[0].split()] == [200, 1]
    blocks = {}
    k = 1
    for spin in (1, 2):
        head = lines[k].split()
        assert (int(head[0]), int(head[1])) == (1, spin)
        assert abs(float(head[2]) - float(g["E0"])) < 1e-10
        blocks[spin] = np.array([[float(t) for t in lines[k + 1 + i].split()] for i in range(200)])
        k += 201
    assert np.array_equal(blocks[1], blocks[2])  # even: spin-1 coefficients repeated
    gs = oracle.ground_state(m, g["em"], g["ek"], g["vmk"], [(3, 3)], 200)
    _, co = oracle.cluster_green_ftn(m, g["em"], g["ek"], g["vmk"], gs, 200, oracle.matsubara(16, 100.0))
    want = np.stack(co[(1, 1)], axis=1)  # (200, 4): ap bp am bm
    assert np.abs(blocks[1][:6] - want[:6]).max() < 1e-10
    assert abs(blocks[1][0, 1] ** 2 + blocks[1][0, 3] ** 2 - 1.0) < 1e-12
