"""Self-energy and quasiparticle weight on the device (edg_self_energy,
post_process.f90:40-58,121-185) against the host formula of the same
routine and against the non-interacting limit.

Tolerances: against the host formula on the same G, 1e-12 relative; at
U = 0 the self-energy of the FP64 solve vanishes to 1e-8 (G is exact to
about 1e-12, and 1/G magnifies its error by |G|^-2 <= 50 on this grid).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_sigma_of_golden_g_matches_host_formula(edgpu):
    from edgpu import dmft
    g = np.load(os.path.join(HERE, "golden", "bethe_ns6.npz"))
    om = g["omega"]
    G = g["G"].reshape(-1, 1, 1)
    s = edgpu.Solver(norb=1, nbath=5, U=float(g["U"]), mu=float(g["mu"]), maxnstep=20)
    S, Z = s.self_energy(g["em"], g["ek"], g["vmk"], om, G)
    D = dmft.cluster_hybridization_ftn(g["ek"], g["vmk"], om, 1)
    want = dmft.self_energy(g["em"], D, G, om, float(g["mu"]))
    assert np.abs(S - want).max() <= 1e-12 * np.abs(want).max()
    zw = dmft.quasiparticle_weight(want, om)
    assert np.abs(Z - zw).max() < 1e-12
    assert 0.0 < Z[0, 0] < 1.0  # U = 2: a correlated impurity


def test_sigma_vanishes_at_u0(edgpu):
    from edgpu import workloads as W
    em, ek, vmk, _ = W.single_orbital(5, 0.0, seed=4)
    mu = 0.3
    om = W.matsubara(256, 100.0)
    s = edgpu.Solver(norb=1, nbath=5, U=0.0, mu=mu, maxnstep=200, precision="fp64", both_spins=True)
    s.ground_state(em, ek, vmk, [(3, 3)])
    G, _ = s.cluster_green_ftn(em, ek, vmk, om)
    S, Z = s.self_energy(em, ek, vmk, om, G)
    assert np.abs(S).max() < 1e-8
    assert abs(Z[0, 0] - 1.0) < 1e-8


def test_sigma_two_orbitals_two_spins(edgpu):
    """norb = 2, nspin = 2: every (orbital, spin) against the host formula
    (the reference keeps only the last orbital's Sigma, dump_sigma :161)."""
    from edgpu import dmft
    rng = np.random.default_rng(61)
    em = rng.uniform(-0.5, 0.5, (2, 2))
    ek = rng.uniform(-1, 1, (4, 2))
    vmk = rng.uniform(0.2, 0.5, (2, 4, 2))
    om = (2 * np.arange(64) + 1) * np.pi / 50.0
    s = edgpu.Solver(norb=2, nbath=4, U=3.0, J=0.75, mu=2.0, maxnstep=120, nspin=2)
    s.ground_state(em, ek, vmk, [(3, 3)])
    G, _ = s.cluster_green_ftn(em, ek, vmk, om)
    S, Z = s.self_energy(em, ek, vmk, om, G)
    iw = 1j * om
    D = np.zeros_like(G)
    for sp in range(2):
        for m in range(2):
            D[:, m, sp] = np.sum(vmk[m, :, sp][None, :] ** 2 / (iw[:, None] - ek[:, sp][None, :]), axis=1)
    want = dmft.self_energy(em, D, G, om, 2.0)
    assert np.abs(S - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(Z - dmft.quasiparticle_weight(want, om)).max() < 1e-12
    assert np.all(np.abs(S[0]) > 1e-3)
