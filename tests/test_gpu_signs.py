"""Hybridizations of either sign through every H.v path, against the oracle.

The hop records carry the Jordan-Wigner parity of a hop and an index into
the signed hybridization table, so a negative V_lm must come out as
parity * V_lm, never parity * |V_lm|.  For one orbital a bath-site gauge
change (b_l -> -b_l) leaves the spectrum unchanged but not H.x, so the
element-wise hx checks below catch a lost sign even where E0 would not;
for two orbitals a bath site coupled with opposite signs to the two
orbitals changes E0 and G itself.

Tolerances (written per test): hx FP64 1e-13, FP32 2e-6 of max|y|; E0
FP64 1e-10, FP32 1e-6; G FP64 1e-9; batched E0 FP64 1e-9, G FP64 1e-8
of max|G|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed_params(rng, norb, nbath):
    em = rng.uniform(-1, 1, (norb, 2))
    ek = rng.uniform(-1, 1, (nbath, 2))
    vmk = rng.uniform(0.2, 0.6, (norb, nbath, 2)) * rng.choice([-1.0, 1.0], (norb, nbath, 2))
    vmk[0, 0, :] = -np.abs(vmk[0, 0, :])  # at least one negative per spin
    if norb >= 2:  # one bath site coupled with opposite signs to orbitals 1 and 2
        vmk[1, 0, :] = np.abs(vmk[0, 0, :])
    return em, ek, vmk


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
@pytest.mark.parametrize("norb,nbath,nu,nd", [(1, 5, 3, 3), (1, 9, 5, 4), (1, 13, 7, 7), (2, 4, 3, 3),
                                              (2, 6, 4, 3), (3, 3, 3, 3)])
def test_hx_negative_hybridization(edgpu, oracle, precision, tol, norb, nbath, nu, nd):
    print(f"hx signs {precision} norb={norb} nbath={nbath} sector=({nu},{nd})", flush=True)
    rng = np.random.default_rng(700 + 10 * norb + nbath)
    em, ek, vmk = mixed_params(rng, norb, nbath)
    s = edgpu.Solver(norb=norb, nbath=nbath, U=2.1, J=0.4, mu=0.8, maxnstep=40, precision=precision)
    m = oracle.Model(norb, nbath, U=2.1, J=0.4, mu=0.8)
    b = oracle.Basis(norb + nbath, nu, nd)
    x = rng.normal(size=b.ntot)
    if precision == "fp32":
        x = x.astype(np.float32).astype(np.float64)
    ref = oracle.hx(m, em, ek, vmk, b, x)
    sched = (1, 2, 4, 8, 16, 32, 64) if b.ntot < 2_000_000 else (1, 2, 16)
    for cols in sched:
        s.set_tuning(cols, False)
        y = s.hx(em, ek, vmk, nu, nd, x)
        assert np.abs(y - ref).max() <= tol * max(1.0, np.abs(ref).max()), cols
    if norb >= 2:  # FP64 row sums as well
        s.set_tuning(2, True)
        y = s.hx(em, ek, vmk, nu, nd, x)
        assert np.abs(y - ref).max() <= tol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("precision,etol,gtol", [("fp64", 1e-10, 1e-9), ("fp32", 1e-6, 1e-4)])
def test_kanamori_opposite_sign_bath_vs_oracle(edgpu, oracle, dense, precision, etol, gtol):
    """Two-orbital Kanamori with a bath site coupled +V / -V to the two
    orbitals: E0 and G(i w) of the GPU solve against the oracle, and the
    oracle's E0 against the dense Jordan-Wigner ground state."""
    rng = np.random.default_rng(808)
    norb, nbath = 2, 4
    em, ek, vmk = mixed_params(rng, norb, nbath)
    U, J, mu = 3.0, 0.75, 2.0
    kw = dict(norb=norb, nbath=nbath, U=U, J=J, mu=mu, maxnstep=120)
    s = edgpu.Solver(precision=precision, both_spins=True, **kw)
    m = oracle.Model(norb, nbath, U=U, J=J, mu=mu)
    gs = s.ground_state(em, ek, vmk, [(3, 3)])
    ogs = oracle.ground_state(m, em, ek, vmk, [(3, 3)], 120)
    H, _ = dense.dense_h(norb, nbath, em, ek, vmk, m.U, m.Up, m.J, m.Jp, m.mu, 3, 3)
    assert abs(ogs.val - np.linalg.eigvalsh(H)[0]) < 1e-10
    assert abs(gs.val - ogs.val) < etol
    om = oracle.matsubara(128, 100.0)
    G, _ = s.cluster_green_ftn(em, ek, vmk, om)
    G2, _ = oracle.cluster_green_ftn(m, em, ek, vmk, ogs, 120, om, both_spins=True)
    assert np.abs(G - G2).max() < gtol
    # the sign matters here: with |V| the ground state moves
    gabs = oracle.ground_state(m, em, ek, np.abs(vmk), [(3, 3)], 120)
    assert abs(gabs.val - ogs.val) > 1e-3


def test_batched_negative_hybridization_vs_oracle(edgpu, oracle):
    """Batched single-orbital problems with V of either sign (short- and
    long-column tiles), FP64, against the oracle's ground_state +
    cluster_green_ftn."""
    omega = oracle.matsubara(128, 100.0)
    probs = []
    for i, nb in enumerate((5, 7, 9, 11)):
        rng = np.random.Generator(np.random.PCG64(900 + i))
        e = rng.uniform(-1, 1, nb)
        vv = rng.uniform(0.2, 0.6, nb) * rng.choice([-1.0, 1.0], nb)
        vv[0] = -abs(vv[0])
        em = np.zeros((1, 2))
        ek = np.stack([e, e], axis=1)
        vmk = np.stack([vv, vv], axis=1)[None, :, :]
        ns = nb + 1
        probs.append(edgpu.Problem(nb, 2.0, 1.0, em, ek, vmk, ns // 2, ns // 2, maxnstep=120,
                                   seed=31 + i))
    b = edgpu.Batch(omega, precision="fp64")
    ids = [b.submit(p) for p in probs]
    while b.step(16):
        pass
    res = {r["id"]: r for r in b.poll()}
    b.close()
    for pr, i in zip(probs, ids):
        print(f"batched signs nbath={pr.nbath}", flush=True)
        m = oracle.Model(1, pr.nbath, U=pr.U, mu=pr.mu)
        em = pr.em.reshape(1, 2, order="F")
        ek = pr.ek.reshape(pr.nbath, 2, order="F")
        vmk = pr.vmk.reshape(1, pr.nbath, 2, order="F")
        gs = oracle.ground_state(m, em, ek, vmk, [(pr.nup, pr.ndn)], pr.maxnstep, seed=pr.seed)
        G, _ = oracle.cluster_green_ftn(m, em, ek, vmk, gs, pr.maxnstep, omega)
        assert abs(res[i]["e0"] - gs.val) < 1e-9
        # relative to max|G| (about 10 at the lowest Matsubara point, beta = 100)
        assert np.abs(res[i]["G"][:, 0] - G[:, 0, 0]).max() < 1e-8 * np.abs(G).max()
