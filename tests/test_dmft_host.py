"""Host-side DMFT pieces (edgpu.dmft): each pinned by an independent
construction rather than by restating its formula.

  cluster_hybridization_ftn  through the U = 0 impurity: 1/(i w + mu - em -
                             Delta) against the dense Jordan-Wigner Lehmann G
  local_green_ftn            against the closed form of the semicircular
                             density, 2 (zeta - sqrt(zeta^2 - 1)), to the
                             rectangle rule's error (2e-6 for |Im zeta| >= 0.05)
  g_coeffs.dat               layout of dmft_green.f90:66-109 (line count,
                             header, record fields) and a write/read round trip
  spectral_function          against the Lorentzian-broadened Lehmann sum of
                             the dense construction (exact chains, Ns = 4)
  self_energy, Z             Sigma = 0 and Z = 1 for the U = 0 impurity (dense
                             Lehmann G); Sigma -> U n_(-s) at high frequency
                             for U > 0 (the Hartree limit); sigma.dat layout
"""
import os

import numpy as np
import pytest

from edgpu import dmft


def bath(seed, nbath, norb=1):
    rng = np.random.default_rng(seed)
    em = rng.uniform(-0.3, 0.3, (norb, 2))
    ek = rng.uniform(-1, 1, (nbath, 2))
    vmk = rng.uniform(0.2, 0.6, (norb, nbath, 2))
    return em, ek, vmk


def test_hybridization_gives_noninteracting_impurity(oracle, dense):
    """At U = 0 the ground state over all sectors is the Fermi sea and its
    impurity G is 1/(i w + mu - em - Delta(i w)) exactly."""
    em, ek, vmk = bath(1, 4)
    om = oracle.matsubara(64, 20.0)
    mu = 0.4
    D = dmft.cluster_hybridization_ftn(ek, vmk, om, nspin=2)
    assert D.shape == (64, 1, 2)
    best = None
    for nu in range(6):
        for nd in range(6):
            H0, _ = dense.dense_h(1, 4, em, ek, vmk, 0.0, 0.0, 0.0, 0.0, mu, nu, nd)
            e = np.linalg.eigvalsh(H0)[0]
            if best is None or e < best[0] - 1e-12:
                best = (e, nu, nd)
    _, nu, nd = best
    for s in (1, 2):
        Gx = dense.lehmann_green(1, 4, em, ek, vmk, 0.0, 0.0, 0.0, 0.0, mu, nu, nd, 1, s, om)[0]
        G0 = 1.0 / (1j * om + mu - em[0, s - 1] - D[:, 0, s - 1])
        assert np.abs(G0 - Gx).max() < 1e-10, s


def test_local_green_bethe_closed_form():
    om = (2 * np.arange(32) + 1) * np.pi / 10.0
    rng = np.random.default_rng(3)
    nw, norb, nspin = om.size, 2, 2
    em = rng.uniform(-0.5, 0.5, (norb, nspin))
    D = rng.uniform(-0.3, 0.3, (nw, norb, nspin)) - 1j * rng.uniform(0.05, 0.3, (nw, norb, nspin))
    G = 1.0 / (1j * om[:, None, None] + rng.uniform(-0.5, 0.5, (1, norb, nspin)) - D)
    h = 0.07
    Gl = dmft.local_green_ftn(em, D, G, hfield=h)
    for s, hs in ((0, h), (1, -h)):
        zeta = em[None, :, s] + hs + D[:, :, s] + 1.0 / G[:, :, s]
        assert np.all(np.abs(zeta.imag) >= 0.05)
        sq = np.sqrt(zeta * zeta - 1.0)
        sq = np.where((sq.imag * zeta.imag) < 0, -sq, sq)
        exact = 2.0 * (zeta - sq)
        assert np.abs(Gl[:, :, s] - exact).max() < 2e-6


def test_g_coeffs_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    norb, ng = 2, 5
    coeffs = {(i, s): tuple(rng.normal(size=ng) for _ in range(4)) for i in (1, 2) for s in (1, 2)}
    p = str(tmp_path / "g_coeffs.dat")
    dmft.write_g_coeffs(p, -3.25, coeffs, norb, ng)
    lines = open(p).read().splitlines()
    assert len(lines) == 1 + 2 * norb * (1 + ng)
    assert [int(t) for t in lines[0].split()] == [ng, norb]
    # record order: orbital outer, spin inner (dmft_green.f90:85-109)
    heads = [lines[1 + k * (1 + ng)].split() for k in range(2 * norb)]
    assert [(int(a), int(b)) for a, b, _ in heads] == [(1, 1), (1, 2), (2, 1), (2, 2)]
    assert all(abs(float(c) + 3.25) < 1e-15 for _, _, c in heads)
    e0, back, n2, g2 = dmft.read_g_coeffs(p)
    assert (n2, g2, e0) == (norb, ng, -3.25)
    for k in coeffs:
        for q in range(4):
            assert np.abs(back[k][q] - coeffs[k][q]).max() < 1e-15


def test_spectral_function_vs_lehmann(oracle, dense):
    em, ek, vmk = bath(5, 3)
    em[:] = 0.0
    ek[:, 1] = ek[:, 0]
    vmk[:, :, 1] = vmk[:, :, 0]
    U, mu = 2.0, 1.0
    m = oracle.Model(1, 3, U=U, mu=mu)
    gs = oracle.ground_state(m, em, ek, vmk, [(2, 2)], maxnstep=40)
    om = oracle.matsubara(8, 10.0)
    _, ap, bp, am, bm, _ = oracle.green_diag(m, em, ek, vmk, gs, 1, 1, 24, om)
    w = np.linspace(-6, 6, 801)
    eta = 0.05
    A = dmft.spectral_function(gs.val, (ap, bp, am, bm), w, eta)
    # Lehmann: weights and poles of the dense construction
    H0, st0 = dense.dense_h(1, 3, em, ek, vmk, U, 0.0, 0.0, 0.0, mu, 2, 2)
    w0, v0 = np.linalg.eigh(H0)
    E0, g = w0[0], v0[:, 0]
    ref = np.zeros_like(w)
    for dag, (nu, nd) in ((1, (3, 2)), (0, (1, 2))):
        H1, st1 = dense.dense_h(1, 3, em, ek, vmk, U, 0.0, 0.0, 0.0, mu, nu, nd)
        e1, v1 = np.linalg.eigh(H1)
        amp = v1.T @ dense.apply_op(st0, g, dag, 0, st1)
        for en, a2 in zip(e1, amp ** 2):
            x = (en - E0) if dag else -(en - E0)
            ref += a2 * eta / np.pi / ((w - x) ** 2 + eta ** 2)
    assert abs(gs.val - E0) < 1e-10
    assert np.abs(A - ref).max() < 1e-7
    assert abs(np.trapezoid(A, w) - (bp[0] ** 2 + bm[0] ** 2)) < 2e-2


def test_self_energy_limits(oracle, dense, tmp_path):
    em, ek, vmk = bath(7, 3)
    em[:] = 0.0
    ek[:, 1] = ek[:, 0]
    vmk[:, :, 1] = vmk[:, :, 0]
    om = oracle.matsubara(400, 2.0)
    mu = 0.3
    D = dmft.cluster_hybridization_ftn(ek, vmk, om, nspin=1)
    # U = 0: Fermi sea over all sectors -> Sigma = 0, Z = 1
    best = min(((np.linalg.eigvalsh(dense.dense_h(1, 3, em, ek, vmk, 0, 0, 0, 0, mu, nu, nd)[0])[0], nu, nd)
                for nu in range(5) for nd in range(5)), key=lambda t: t[0])
    G = dense.lehmann_green(1, 3, em, ek, vmk, 0, 0, 0, 0, mu, best[1], best[2], 1, 1, om)[0]
    S = dmft.self_energy(em, D, G[:, None, None], om, mu)
    assert np.abs(S).max() < 1e-9
    assert abs(dmft.quasiparticle_weight(S, om)[0, 0] - 1.0) < 1e-9
    # U > 0, sector (2,2): Sigma(i w) -> U <n_dn> as w -> infinity
    U = 2.0
    H0, st0 = dense.dense_h(1, 3, em, ek, vmk, U, 0, 0, 0, U / 2, 2, 2)
    w0, v0 = np.linalg.eigh(H0)
    n_dn = sum(abs(v0[i, 0]) ** 2 for i, s in enumerate(st0) if (s >> 4) & 1)
    G = dense.lehmann_green(1, 3, em, ek, vmk, U, 0, 0, 0, U / 2, 2, 2, 1, 1, om)[0]
    S = dmft.self_energy(em, D, G[:, None, None], om, U / 2)
    assert abs(S[-1, 0, 0].real - U * n_dn) < 2e-3
    assert 0.0 < dmft.quasiparticle_weight(S, om)[0, 0] < 1.0
    paths = dmft.write_sigma(om[:3], S[:3], str(tmp_path))
    lines = open(paths[0]).read().splitlines()
    assert paths[0].endswith("sigma.dat.1.1")
    assert [int(t) for t in lines[0].split()] == [3, 1, 1] and len(lines) == 4
    assert len(lines[1]) == 36 and abs(float(lines[1][12:24]) - S[0, 0, 0].real) < 1e-8
