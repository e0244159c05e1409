"""Pins the FP64 CPU oracle (oracle/ed_oracle.c) before it is trusted as the
checker of the GPU path.  Each check is chosen so that a plausible mistake
(dropped term, wrong sign or index, transposed operand) fails it:

  basis ranking     vs itertools enumeration and the combinatorial number
                    system closed form; vs the SHA-256 fixtures
  hx                vs an explicit Jordan-Wigner dense Hamiltonian (no sgn
                    shortcuts), norb = 1, 2, 3, asymmetric spins, Kanamori
  diag_lanczos      vs numpy eigh; residual; start-vector independence
  tridiagonal       vs scipy eigh_tridiagonal (LAPACK)
  apply_c           vs Jordan-Wigner c+/c, both spins
  green_diag / CF   vs the Lehmann sum; U=0 closed form 1/(iw+mu-em-Delta);
                    sum rule bp1^2+bm1^2 = 1; two-level continued fraction
  observables       vs occupancies from the dense eigenvector; n = bm1^2
  golden fixtures   config-1 E0, n, docc, G(iw) from dense ED
"""
import hashlib
import itertools
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def rand_params(rng, norb, nbath, spin_sym=False):
    em = rng.uniform(-1, 1, (norb, 2))
    ek = rng.uniform(-1, 1, (nbath, 2))
    vmk = rng.uniform(0.2, 0.6, (norb, nbath, 2))
    if spin_sym:
        em[:, 1] = em[:, 0]
        ek[:, 1] = ek[:, 0]
        vmk[:, :, 1] = vmk[:, :, 0]
    return em, ek, vmk


def oracle_dense(O, m, em, ek, vmk, b):
    H = np.zeros((b.ntot, b.ntot))
    for j in range(b.ntot):
        e = np.zeros(b.ntot)
        e[j] = 1.0
        H[:, j] = O.hx(m, em, ek, vmk, b, e)
    return H


@pytest.mark.parametrize("ns,n", [(4, 2), (6, 3), (8, 4), (8, 1), (8, 7), (14, 7), (15, 7)])
def test_basis_ranking_closed_form(oracle, ns, n):
    b = oracle.Basis(ns, n, n)
    pats = sorted(sum(1 << p for p in c) for c in itertools.combinations(range(ns), n))
    assert list(b.up) == pats and list(b.dn) == pats
    from math import comb
    for r, p in enumerate(pats):
        bits = [i for i in range(ns) if (p >> i) & 1]
        assert sum(comb(bp, k + 1) for k, bp in enumerate(bits)) == r
        assert b.idx_up[p] == r + 1
    assert b.ntot == comb(ns, n) ** 2


def test_basis_fixture_hashes(oracle):
    with open(os.path.join(HERE, "golden", "tables.json")) as f:
        ref = json.load(f)
    for ns, n in ((6, 3), (8, 4), (14, 7)):
        b = oracle.Basis(ns, n, n)
        h = hashlib.sha256(b.up.astype("<i4").tobytes()).hexdigest()
        assert h == ref[f"ns{ns}_n{n}_patterns"]
        assert b.nup == ref[f"ns{ns}_n{n}_count"]


def test_row_order_up_fastest(oracle):
    b = oracle.Basis(5, 2, 3)
    st = b.states()
    assert st[1] == b.up[1] + (b.dn[0] << 5)
    assert st[b.nup] == b.up[0] + (b.dn[1] << 5)


@pytest.mark.parametrize("norb,nbath,nu,nd,sym", [
    (1, 3, 2, 2, False), (1, 5, 3, 3, True), (1, 4, 1, 4, False), (2, 2, 2, 2, False),
    (2, 3, 3, 2, False), (2, 2, 1, 3, True), (3, 1, 2, 2, False), (3, 2, 3, 2, False),
    (4, 2, 3, 2, False), (5, 2, 4, 3, False), (1, 5, 0, 3, False), (1, 5, 6, 2, False)])
def test_hx_matches_jordan_wigner(oracle, dense, norb, nbath, nu, nd, sym):
    rng = np.random.default_rng(10 * norb + nbath + nu)
    em, ek, vmk = rand_params(rng, norb, nbath, sym)
    m = oracle.Model(norb, nbath, U=2.3, J=0.4, mu=0.7)
    b = oracle.Basis(norb + nbath, nu, nd)
    H = oracle_dense(oracle, m, em, ek, vmk, b)
    Hd, st = dense.dense_h(norb, nbath, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, nu, nd)
    assert np.array_equal(st, b.states())
    assert np.abs(H - Hd).max() < 1e-13
    assert np.abs(H - H.T).max() == 0.0


@pytest.mark.parametrize("norb,nbath,nu,nd", [(1, 5, 3, 3), (2, 3, 2, 3), (2, 4, 3, 3), (3, 2, 3, 2)])
def test_hx_negative_hybridization_matches_jordan_wigner(oracle, dense, norb, nbath, nu, nd):
    """V of either sign (one bath site coupled +V / -V to two orbitals): the
    reference's sgn/sgn2 shortcuts must give parity * V, never parity * |V|."""
    rng = np.random.default_rng(40 + 10 * norb + nbath)
    em, ek, vmk = rand_params(rng, norb, nbath)
    vmk *= rng.choice([-1.0, 1.0], vmk.shape)
    vmk[0, 0, :] = -np.abs(vmk[0, 0, :])
    if norb >= 2:
        vmk[1, 0, :] = np.abs(vmk[0, 0, :])
    m = oracle.Model(norb, nbath, U=2.3, J=0.4, mu=0.7)
    b = oracle.Basis(norb + nbath, nu, nd)
    H = oracle_dense(oracle, m, em, ek, vmk, b)
    Hd, _ = dense.dense_h(norb, nbath, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, nu, nd)
    assert np.abs(H - Hd).max() < 1e-13
    Ha = oracle_dense(oracle, m, em, ek, np.abs(vmk), b)
    assert np.abs(H - Ha).max() > 0.1  # the signs are visible element-wise


def test_hx_kanamori_fixture(oracle):
    g = np.load(os.path.join(HERE, "golden", "kanamori_ns4.npz"))
    m = oracle.Model(2, 2, U=float(g["U"]), J=float(g["J"]), Up=float(g["Up"]), Jp=float(g["Jp"]),
                     mu=float(g["mu"]))
    b = oracle.Basis(4, 2, 2)
    H = oracle_dense(oracle, m, g["em"], g["ek"], g["vmk"], b)
    assert np.abs(H - g["H"]).max() < 1e-13
    # the Kanamori terms must be present: spin-flip / pair-hop elements equal +-Jp
    off = np.abs(g["H"][np.abs(g["H"]) > 0])
    assert np.any(np.isclose(off, float(g["Jp"])))


def test_hx_accumulate(oracle):
    rng = np.random.default_rng(3)
    em, ek, vmk = rand_params(rng, 1, 5)
    m = oracle.Model(1, 5, U=2.0, mu=1.0)
    b = oracle.Basis(6, 3, 3)
    x, y0 = rng.normal(size=b.ntot), rng.normal(size=b.ntot)
    assert np.allclose(oracle.hx(m, em, ek, vmk, b, x, y0, py=True), y0 + oracle.hx(m, em, ek, vmk, b, x),
                       atol=1e-14)


def test_tridiagonal_vs_lapack(oracle):
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 60, 200):
        d = rng.normal(size=n)
        e = rng.uniform(0.05, 1.0, max(n - 1, 0))
        w, z = oracle.tridiag_eig(d, e)
        if n == 1:
            assert w[0] == d[0]
            continue
        w2, z2 = sl.eigh_tridiagonal(d, e)
        assert np.abs(w - w2).max() < 1e-12
        assert abs(abs(z[:, 0] @ z2[:, 0]) - 1) < 1e-10


def test_diag_lanczos_vs_eigh(oracle, dense):
    rng = np.random.default_rng(7)
    em, ek, vmk = rand_params(rng, 1, 5, True)
    m = oracle.Model(1, 5, U=3.0, mu=1.5)
    b = oracle.Basis(6, 3, 3)
    Hd, _ = dense.dense_h(1, 5, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, 3, 3)
    w, v = np.linalg.eigh(Hd)
    for seed in (1, 2):
        e0, vec, nstep, res = oracle.diag_lanczos(m, em, ek, vmk, b, oracle.start_vector(seed, b.ntot), 100)
        assert abs(e0 - w[0]) < 1e-11
        assert abs(abs(vec @ v[:, 0]) - 1) < 1e-10
        assert res < 1e-18
        assert nstep == 100


def test_lanczos_zero_padding_and_b1(oracle):
    rng = np.random.default_rng(8)
    em, ek, vmk = rand_params(rng, 1, 3)
    m = oracle.Model(1, 3, U=1.0, mu=0.5)
    b = oracle.Basis(4, 2, 2)
    n, a, bb = oracle.lanczos_iteration(m, em, ek, vmk, b, oracle.start_vector(3, b.ntot), 12)
    assert bb[0] == 0.0 and n == 12


def test_lanczos_breakdown_on_eigenvector(oracle, dense):
    """An eigenvector start gives b(2) ~ 0: the reference exits with nstep=2
    after dividing by b(2) first (lanczos_solver.f90:152-163)."""
    rng = np.random.default_rng(9)
    em, ek, vmk = rand_params(rng, 1, 3)
    m = oracle.Model(1, 3, U=1.0, mu=0.5)
    b = oracle.Basis(4, 2, 2)
    Hd, _ = dense.dense_h(1, 3, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, 2, 2)
    w, v = np.linalg.eigh(Hd)
    n, a, bb = oracle.lanczos_iteration(m, em, ek, vmk, b, v[:, 0], 10)
    assert abs(a[0] - w[0]) < 1e-12
    assert bb[1] < 1e-13


def test_ground_state_sector_sweep(oracle, dense):
    rng = np.random.default_rng(11)
    em, ek, vmk = rand_params(rng, 1, 5, True)
    m = oracle.Model(1, 5, U=2.0, mu=1.0)
    secs = [(2, 2), (3, 3), (3, 2), (4, 4)]
    gs = oracle.ground_state(m, em, ek, vmk, secs, maxnstep=120)
    e = [np.linalg.eigvalsh(dense.dense_h(1, 5, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, *s)[0])[0]
         for s in secs]
    assert gs.isector == int(np.argmin(e)) + 1
    assert abs(gs.val - min(e)) < 1e-10


@pytest.mark.parametrize("pm,ispin,norb", [(1, 1, 1), (2, 1, 1), (1, 2, 1), (2, 2, 1), (1, 2, 2), (2, 1, 2)])
def test_apply_c_vs_jordan_wigner(oracle, dense, pm, ispin, norb):
    rng = np.random.default_rng(pm + 2 * ispin)
    ns = norb + 3
    nu, nd = 2, 2
    b = oracle.Basis(ns, nu, nd)
    nu2 = nu + (0 if ispin == 2 else (1 if pm == 1 else -1))
    nd2 = nd + (0 if ispin == 1 else (1 if pm == 1 else -1))
    bo = oracle.Basis(ns, nu2, nd2)
    vec = rng.normal(size=b.ntot)
    iorb = norb
    out = oracle.apply_c(b, vec, pm, iorb, ispin, bo)
    mode = (ispin - 1) * ns + iorb - 1
    ref = dense.apply_op(b.states(), vec, 1 if pm == 1 else 0, mode, bo.states())
    assert np.array_equal(out, ref)


def test_continued_fraction_two_level(oracle):
    # f(z) = b1^2/(z - a1 - b2^2/(z - a2)) has poles at the eigenvalues of [[a1,b2],[b2,a2]]
    a = np.array([0.3, -0.7])
    b = np.array([1.3, 0.5])
    T = np.array([[a[0], b[1]], [b[1], a[1]]])
    w, v = np.linalg.eigh(T)
    for z in (0.2 + 1j, -1.0 + 0.1j, 3.0 + 2.0j):
        ref = b[0] ** 2 * np.sum(v[0, :] ** 2 / (z - w))
        assert abs(oracle.continued_fraction(z, 2, a, b) - ref) < 1e-14
        refm = b[0] ** 2 * np.sum(v[0, :] ** 2 / (z + w))
        assert abs(oracle.continued_fraction(z, 2, a, b, minus=True) - refm) < 1e-14


def test_continued_fraction_n1_quirk(oracle):
    """numeric_utils.f90:18-22 applies the last level after the loop even for n=1."""
    z = 0.5 + 1j
    f1 = 4.0 / (z - 0.25)
    assert abs(oracle.continued_fraction(z, 1, [0.25], [2.0]) - 4.0 / (z - 0.25 - f1)) < 1e-14


def test_green_vs_lehmann_and_sum_rule(oracle, dense):
    rng = np.random.default_rng(12)
    em, ek, vmk = rand_params(rng, 1, 5, True)
    m = oracle.Model(1, 5, U=2.5, mu=1.25)
    om = oracle.matsubara(256, 50.0)
    gs = oracle.ground_state(m, em, ek, vmk, [(3, 3)], maxnstep=150)
    G, co = oracle.cluster_green_ftn(m, em, ek, vmk, gs, 150, om)
    Gl, E0, wp, wm = dense.lehmann_green(1, 5, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, 3, 3, 1, 1, om)
    assert np.abs(G[:, 0, 0] - Gl).max() < 1e-10
    ap, bp, am, bm = co[(1, 1)]
    assert abs(bp[0] ** 2 + bm[0] ** 2 - 1.0) < 1e-12
    # weights from the Lanczos eigenvector: accurate to its residual (~1e-11)
    assert abs(bp[0] ** 2 - wp) < 1e-10 and abs(bm[0] ** 2 - wm) < 1e-10
    occ, docc = oracle.observables(m, gs.basis, gs.vec)
    assert abs(occ[0, 0] - bm[0] ** 2) < 1e-12


def test_green_noninteracting_closed_form(oracle, dense):
    rng = np.random.default_rng(13)
    e = rng.uniform(-1, 1, 5)
    v = rng.uniform(0.2, 0.6, 5)
    em = np.zeros((1, 2))
    ek = np.stack([e, e], 1)
    vmk = np.stack([v, v], 1)[None]
    m = oracle.Model(1, 5, U=0.0, mu=0.0)
    om = oracle.matsubara(512, 100.0)
    gs = oracle.ground_state(m, em, ek, vmk, [(3, 3)], maxnstep=150)
    G, _ = oracle.cluster_green_ftn(m, em, ek, vmk, gs, 150, om)
    Gc, h = dense.noninteracting(0.0, e, v, 0.0, om)
    assert np.abs(G[:, 0, 0] - Gc).max() < 1e-10
    assert abs(gs.val - 2 * np.linalg.eigvalsh(h)[:3].sum()) < 1e-10


def test_green_spin_asymmetric_average(oracle, dense):
    """nspin=1 with ne_up != ne_dn averages the two spins (dmft_green.f90:113-114)."""
    rng = np.random.default_rng(14)
    em, ek, vmk = rand_params(rng, 1, 3, True)
    m = oracle.Model(1, 3, U=1.5, mu=0.75)
    om = oracle.matsubara(64, 30.0)
    gs = oracle.ground_state(m, em, ek, vmk, [(2, 1)], maxnstep=60)
    G, co = oracle.cluster_green_ftn(m, em, ek, vmk, gs, 60, om)
    g1 = dense.lehmann_green(1, 3, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, 2, 1, 1, 1, om)[0]
    g2 = dense.lehmann_green(1, 3, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, 2, 1, 1, 2, om)[0]
    assert np.abs(G[:, 0, 0] - 0.5 * (g1 + g2)).max() < 1e-10


def test_golden_bethe_ns6(oracle):
    g = np.load(os.path.join(HERE, "golden", "bethe_ns6.npz"))
    m = oracle.Model(1, 5, U=float(g["U"]), J=float(g["J"]), mu=float(g["mu"]))
    gs = oracle.ground_state(m, g["em"], g["ek"], g["vmk"], [(3, 3)], maxnstep=200, seed=9)
    assert abs(gs.val - float(g["E0"])) < 1e-11
    occ, docc = oracle.observables(m, gs.basis, gs.vec)
    assert abs(occ[0, 0] - float(g["n_up"])) < 1e-10 and abs(occ[0, 1] - float(g["n_dn"])) < 1e-10
    assert abs(docc[0] - float(g["docc"])) < 1e-10
    G, _ = oracle.cluster_green_ftn(m, g["em"], g["ek"], g["vmk"], gs, 200, g["omega"])
    assert np.abs(G[:, 0, 0] - g["G"]).max() < 1e-10
