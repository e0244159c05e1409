"""FP64 reorthogonalization of the Green's-function chains
(edg_set_reorthogonalization; the reference runs the plain three-term
recurrence, lanczos_solver.f90:147-170, from dmft_green.f90:218,248).

Checks that do not retype the kernels' formulas:
  * ghost Ritz values: a plain Lanczos chain run past the loss of
    orthogonality repeats converged Ritz values of its tridiagonal matrix
    (Paige's theory); a chain kept orthogonal (full) or semi-orthogonal
    (partial, Simon's omega recurrence) holds each once.  The reference
    eigenvalues come from a Lanczos run with twice-repeated classical
    Gram-Schmidt in numpy on the oracle's H.v, whose basis is orthogonal to
    1e-15 (the textbook remedy).
  * FP64 coefficients: with reorthogonalization the first K coefficients
    follow that numpy run, while the plain chain departs from it by O(1).
    The coefficients are sensitive to rounding in any implementation: a
    1e-15 change of the start vector moves them by up to 1e-7 within these
    K steps, so the tolerance is 1e-6.
  * G(i w) against the oracle's FP64 solve of the same ground state: 1e-9
    (FP64 vectors), 1e-5 (FP32 vectors).
  * the step counts of edg_reorth_stats.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def orthogonal_lanczos(oracle, m, em, ek, vmk, basis, v0, nstep):
    """Lanczos with full reorthogonalization (classical Gram-Schmidt, twice)
    in FP64 numpy on the oracle's H.v -> alpha(nstep), beta(nstep + 1)."""
    Q = np.zeros((nstep + 1, v0.size))
    a = np.zeros(nstep)
    b = np.zeros(nstep + 1)
    b[0] = np.linalg.norm(v0)
    Q[0] = v0 / b[0]
    for j in range(nstep):
        w = oracle.hx(m, em, ek, vmk, basis, Q[j])
        if j:
            w = w - b[j] * Q[j - 1]
        a[j] = Q[j] @ w
        w = w - a[j] * Q[j]
        for _ in range(2):
            w = w - Q[: j + 1].T @ (Q[: j + 1] @ w)
        b[j + 1] = np.linalg.norm(w)
        Q[j + 1] = w / b[j + 1]
    return a, b


def ritz(a, b, n):
    from scipy.linalg import eigh_tridiagonal
    return eigh_tridiagonal(a[:n], b[1:n])[0]


CASES = {}


def case(oracle, nbath):
    """Single-orbital impurity, sector (Ns/2, Ns/2), the oracle's ground state
    and the orthogonal reference runs of its c+ and c chains (150 steps)."""
    if nbath in CASES:
        return CASES[nbath]
    from edgpu import workloads as W
    em, ek, vmk, mu = W.single_orbital(nbath, 2.5, seed=nbath)
    m = oracle.Model(1, nbath, U=2.5, mu=mu)
    ns = nbath + 1
    nu = nd = ns // 2
    gs = oracle.ground_state(m, em, ek, vmk, [(nu, nd)], 200, seed=3)
    ref = {}
    for pm, nu2 in ((1, nu + 1), (2, nu - 1)):
        bo = oracle.Basis(ns, nu2, nd)
        v0 = oracle.apply_c(gs.basis, gs.vec, pm, 1, 1, bo)
        ref[pm] = orthogonal_lanczos(oracle, m, em, ek, vmk, bo, v0, 150)
    CASES[nbath] = (em, ek, vmk, mu, m, gs, ref)
    return CASES[nbath]


def run(edgpu, nbath, mu, em, ek, vmk, gs, precision, mode, om):
    s = edgpu.Solver(norb=1, nbath=nbath, U=2.5, mu=mu, maxnstep=200, maxnstep_green=150,
                     precision=precision)
    s.set_reorthogonalization(mode)
    s.set_gs(gs.basis.ne_up, gs.basis.ne_dn, gs.val, gs.vec)
    G, co = s.cluster_green_ftn(em, ek, vmk, om)
    st = s.reorth_stats()
    s.close()
    return G, co[(1, 1)], st


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("nbath", [5, 7])
def test_ghost_ritz_values_and_green(edgpu, oracle, nbath, mode, precision):
    print(f"reorth nbath={nbath} mode={mode} {precision}", flush=True)
    em, ek, vmk, mu, m, gs, ref = case(oracle, nbath)
    om = oracle.matsubara(512, 100.0)
    G, (ap, bp, am, bm), st = run(edgpu, nbath, mu, em, ek, vmk, gs, precision, mode, om)
    ghosts = 0
    for pm, (a, b) in ((1, (ap, bp)), (2, (am, bm))):
        lo = ritz(*ref[pm], 150)
        th = ritz(a, b, 150)
        copies = [int((np.abs(th - lo[i]) < 1e-4).sum()) for i in (0, 1, 2, -1)]
        if mode:
            assert copies == [1, 1, 1, 1], (pm, copies)
        ghosts += max(copies) > 1
    if mode == 0:
        assert ghosts > 0  # the plain recurrence, for contrast
        assert st == dict(events=0, steps=0)
    elif mode == 2:
        assert st["events"] == st["steps"] == 2 * 149
    else:
        assert 0 < st["events"] < st["steps"] == 2 * 149
    G2, _ = oracle.cluster_green_ftn(m, em, ek, vmk, gs, 150, om)
    assert np.abs(G - G2).max() < (1e-9 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("nbath,K", [(5, 60), (7, 100)])
def test_fp64_coefficients_follow_orthogonal_lanczos(edgpu, oracle, nbath, K, mode):
    print(f"reorth coefficients nbath={nbath} mode={mode}", flush=True)
    em, ek, vmk, mu, m, gs, ref = case(oracle, nbath)
    om = oracle.matsubara(64, 100.0)
    _, (ap, bp, am, bm), _ = run(edgpu, nbath, mu, em, ek, vmk, gs, "fp64", mode, om)
    _, (ap0, bp0, _, _), _ = run(edgpu, nbath, mu, em, ek, vmk, gs, "fp64", 0, om)
    for (a, b), (ra, rb) in (((ap, bp), ref[1]), ((am, bm), ref[2])):
        assert abs(b[0] - rb[0]) < 1e-12  # |c+-|gs>|
        assert np.abs(a[:K] - ra[:K]).max() < 1e-6
        assert np.abs(b[1:K] - rb[1:K]).max() < 1e-6
    ra, rb = ref[1]
    assert max(np.abs(ap0[:K] - ra[:K]).max(), np.abs(bp0[1:K] - rb[1:K]).max()) > 1e-2


def test_reorth_off_by_default_and_rejects_bad_mode(edgpu):
    s = edgpu.Solver(norb=1, nbath=5, U=2.0, mu=1.0, maxnstep=60)
    with pytest.raises(edgpu.EdgpuError):
        s.set_reorthogonalization(3)
    assert s.reorth_stats() == dict(events=0, steps=0)
    s.close()


def test_partial_fp32_full_size(edgpu):
    """Ns=14 single-orbital impurity (U = 2, sector (7,7), chains of 150
    steps, FP32 vectors): partial reorthogonalization against the FP64
    oracle's solve stored in tests/golden/ns14_oracle.npz (G within 1e-5),
    and the extra time it costs over the plain chains."""
    import os
    import time
    from edgpu import workloads as W
    ref = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ns14_oracle.npz")))
    em, ek, vmk, mu = W.single_orbital(13, 2.0, seed=1)
    om = ref["omega"]
    out = {}
    for mode in (0, 1):
        print(f"reorth Ns=14 mode={mode}", flush=True)
        s = edgpu.Solver(norb=1, nbath=13, U=2.0, mu=mu, maxnstep=200, maxnstep_green=150,
                         precision="fp32", both_spins=True, seed=1)
        s.set_reorthogonalization(mode)
        s.ground_state(em, ek, vmk, [(7, 7)])
        s.cluster_green_ftn(em, ek, vmk, om)  # tables and buffers
        t = time.perf_counter()
        G, _ = s.cluster_green_ftn(em, ek, vmk, om)
        out[mode] = (G, time.perf_counter() - t, s.reorth_stats())
        s.close()
    G1, t1, st = out[1]
    print(f"plain {out[0][1] * 1e3:.1f} ms, partial {t1 * 1e3:.1f} ms, {st}", flush=True)
    assert np.abs(G1[:, 0, 0] - ref["u2_G"][:, 0]).max() < 1e-5
    assert 0 < st["events"] < st["steps"] == 4 * 149
