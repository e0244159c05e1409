"""Batched solver with two-orbital (Kanamori) problems (edg_problem.norb = 2):
the c+/c chains of both orbitals and spins (dmft_green.f90:71-118) in the
batched launches, against the FP64 oracle's solve of the same problem, next
to single-orbital problems in the same batch.

Tolerances (written per test):
  E0 FP64 / FP32 vs oracle          1e-9 / 1e-6 (absolute)
  G_cl FP64 / FP32 vs oracle        1e-8 / 1e-5 (max abs over the grid;
                                    FP32 1e-3 at the first two Matsubara
                                    points: FP32 pole errors grow as
                                    1/w_n^2 there, and the problem with the
                                    sign-flipped coupling has weight near
                                    zero energy: 3.2e-4 at w_0 through the
                                    single-problem path and the batch alike)
  leading chain coefficients FP64   1e-9 (the first 6 of each chain, for
                                    the half-filled problems: their
                                    120-step ground state moves those
                                    coefficients by < 1e-9 between 120 and
                                    300 oracle steps; in sector (3,2) they
                                    move by 2.3e-9, and the chains amplify
                                    the start vector's difference from the
                                    oracle's to 4e-8 by the fourth step)
  alone vs in a mix                 bit-identical
  Ns = 14 (FP32) vs the stored FP64 oracle solve: E0 1e-6, G 1e-5
The problems have a ground-state gap of 0.16 or more in their sector: the
oracle's own G then moves by less than 1e-11 between 120 and 300
ground-state steps.  (With U = 4, seed 3, sector (3,2) the gap is 0.018 and
the oracle's G at the lowest Matsubara point moves by 1.8e-7 between 120
and 300 steps, so no 1e-8 comparison is meaningful there.)
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def kan_problem(edgpu, W, nb_per_orb, U, seed, nup=None, ndn=None, maxnstep=120, maxnstep_green=80,
                both_spins=True, flip=False):
    em, ek, vmk, mu, J, Up = W.two_orbital_kanamori(nb_per_orb, U, seed=seed)
    em = em + np.array([[-0.1, -0.1], [0.15, 0.15]])  # split the orbitals
    if flip:  # one bath site coupled with opposite signs to the two orbitals
        vmk = vmk.copy()
        vmk[1, 0, :] = -0.3
        vmk[0, 0, :] = 0.25
    nb = 2 * nb_per_orb
    ns = nb + 2
    nup = ns // 2 if nup is None else nup
    ndn = ns // 2 if ndn is None else ndn
    return edgpu.Problem(nb, U, mu, em, ek, vmk, nup, ndn, maxnstep=maxnstep, maxnstep_green=maxnstep_green,
                         both_spins=both_spins, seed=seed, norb=2, J=J, Up=Up, Jp=J)


def oracle_kan(oracle, pr, omega, nspin=1):
    m = oracle.Model(2, pr.nbath, U=pr.U, J=pr.J, Up=pr.Up, Jp=pr.Jp, mu=pr.mu)
    em = pr.em.reshape(2, 2, order="F")
    ek = pr.ek.reshape(pr.nbath, 2, order="F")
    vmk = pr.vmk.reshape(2, pr.nbath, 2, order="F")
    gs = oracle.ground_state(m, em, ek, vmk, [(pr.nup, pr.ndn)], pr.maxnstep, seed=pr.seed)
    G, co = oracle.cluster_green_ftn(m, em, ek, vmk, gs, pr.maxnstep_green or pr.maxnstep, omega,
                                     nspin=nspin, both_spins=pr.both_spins)
    return gs.val, G, co


def run(edgpu, probs, omega, precision, nspin=1, graphs=True):
    b = edgpu.Batch(omega, precision=precision, nspin=nspin, graphs=graphs)
    ids = [b.submit(p) for p in probs]
    b.run()
    res = {r["id"]: r for r in b.poll()}
    b.close()
    return [res[i] for i in ids]


@pytest.mark.parametrize("precision,etol,gtol", [("fp64", 1e-9, 1e-8), ("fp32", 1e-6, 1e-5)])
def test_kanamori_batch_vs_oracle(edgpu, oracle, precision, etol, gtol):
    """Ns = 6 and 8 Kanamori problems (one with a sign-flipped coupling, one
    off half filling) in one batch with single-orbital problems."""
    from edgpu import workloads as W
    print(f"batched kanamori {precision}", flush=True)
    om = W.matsubara(256, 100.0)
    kan = [kan_problem(edgpu, W, 2, 3.0, 1), kan_problem(edgpu, W, 3, 2.5, 2, flip=True),
           kan_problem(edgpu, W, 2, 3.0, 6, nup=3, ndn=2), kan_problem(edgpu, W, 3, 2.0, 4, both_spins=False)]
    em1, ek1, vmk1, mu1 = W.single_orbital(7, 2.0, seed=5)
    single = edgpu.Problem(7, 2.0, mu1, em1, ek1, vmk1, 4, 4, maxnstep=120, maxnstep_green=80, seed=5)
    res = run(edgpu, kan + [single], om, precision)
    for pr, r in zip(kan, res):
        e0, G, co = oracle_kan(oracle, pr, om)
        assert abs(r["e0"] - e0) < etol
        assert r["G"].shape == (256, 2, 1)
        tol = np.full(om.size, gtol)
        if precision == "fp32":
            tol[:2] = 1e-3
        assert (np.abs(r["G"] - G).max(axis=(1, 2)) < tol).all()
        if precision == "fp64" and pr.nup == pr.ndn:
            for key, (ap, bp, am, bm) in r["coeffs"].items():
                rap, rbp, ram, rbm = co[key]
                for x, y in ((ap, rap), (bp, rbp), (am, ram), (bm, rbm)):
                    assert np.abs(x[:6] - y[:6]).max() < 1e-9, key
    m = oracle.Model(1, 7, U=2.0, mu=mu1)
    gs = oracle.ground_state(m, em1, ek1, vmk1, [(4, 4)], 120, seed=5)
    G1, _ = oracle.cluster_green_ftn(m, em1, ek1, vmk1, gs, 80, om)
    assert abs(res[-1]["e0"] - gs.val) < etol
    assert np.abs(res[-1]["G"][:, 0] - G1[:, 0, 0]).max() < gtol


def test_kanamori_nspin2_and_independence(edgpu, oracle):
    """nspin = 2 (both spins of G_cl) with a spin-polarized bath, and the
    same problem alone and in a mix gives bit-identical results."""
    from edgpu import workloads as W
    om = W.matsubara(128, 50.0)
    pr = kan_problem(edgpu, W, 2, 3.0, 7, nup=4, ndn=3)
    pr.ek = pr.ek.copy()
    pr.ek[pr.nbath:] += 0.05  # spin-down bath levels shifted
    e0, G, _ = oracle_kan(oracle, pr, om, nspin=2)
    alone = run(edgpu, [pr], om, "fp64", nspin=2)[0]
    others = [kan_problem(edgpu, W, 3, 2.0, 8), kan_problem(edgpu, W, 2, 2.0, 9)]
    mixed = run(edgpu, others + [pr], om, "fp64", nspin=2)[-1]
    assert abs(alone["e0"] - e0) < 1e-9
    assert alone["G"].shape == (128, 2, 2)
    assert np.abs(alone["G"] - G).max() < 1e-8
    assert alone["e0"] == mixed["e0"]
    assert np.array_equal(alone["G"], mixed["G"])


def test_kanamori_batch_matches_single_problem_path(edgpu, oracle):
    """The batched Kanamori solve against edg_ground_state +
    edg_cluster_green on the same problem (FP64): E0 1e-10, G 1e-9."""
    from edgpu import workloads as W
    om = W.matsubara(128, 100.0)
    pr = kan_problem(edgpu, W, 3, 3.0, 11)
    r = run(edgpu, [pr], om, "fp64")[0]
    s = edgpu.Solver(norb=2, nbath=pr.nbath, U=pr.U, J=pr.J, Up=pr.Up, Jp=pr.Jp, mu=pr.mu,
                     maxnstep=pr.maxnstep, maxnstep_green=pr.maxnstep_green, both_spins=True, seed=pr.seed)
    em = pr.em.reshape(2, 2, order="F")
    ek = pr.ek.reshape(pr.nbath, 2, order="F")
    vmk = pr.vmk.reshape(2, pr.nbath, 2, order="F")
    gs = s.ground_state(em, ek, vmk, [(pr.nup, pr.ndn)])
    G, _ = s.cluster_green_ftn(em, ek, vmk, om)
    s.close()
    assert abs(r["e0"] - gs.val) < 1e-10
    assert np.abs(r["G"] - G).max() < 1e-9


def test_kanamori_batch_full_size(edgpu):
    """Ns = 14 Kanamori (6 bath sites per orbital, sector (7,7), 200 + 8 x 150
    steps, FP32 vectors) against the FP64 oracle's solve stored in
    tests/golden/ns14_oracle.npz, two copies in one batch."""
    from edgpu import workloads as W
    ref = dict(np.load(os.path.join(HERE, "golden", "ns14_oracle.npz")))
    em, ek, vmk, mu, J, Up = W.two_orbital_kanamori(6, 3.0, seed=1)
    pr = edgpu.Problem(12, 3.0, mu, em, ek, vmk, 7, 7, maxnstep=200, maxnstep_green=150, both_spins=True,
                       seed=1, norb=2, J=J, Up=Up, Jp=J)
    print("batched kanamori Ns=14", flush=True)
    res = run(edgpu, [pr, pr], ref["omega"], "fp32")
    for r in res:
        assert abs(r["e0"] - float(ref["kan_E0"])) < 1e-6
        assert np.abs(r["G"][:, :, 0] - ref["kan_G"]).max() < 1e-5
