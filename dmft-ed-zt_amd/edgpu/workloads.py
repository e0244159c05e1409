"""Seeded synthetic impurity problems of the benchmark configurations.

Bath parameters come from numpy's PCG64 generator: for each bath site the
level eps_l ~ U[eps_lo, eps_hi] is drawn first, then the hybridization
V_l ~ U[v_lo, v_hi] (all levels, then all hybridizations).  Spin-symmetric
baths copy spin 1 into spin 2.  Half filling of the single-orbital model
uses em = 0 and mu = U/2 (mu acts on the impurity only,
impurity_hamiltonian.f90:119-120).
"""
import numpy as np


def single_orbital(nbath, U, seed, eps=(-1.0, 1.0), v=(0.2, 0.6)):
    """(em(1,2), ek(nbath,2), vmk(1,nbath,2), mu) for one orbital."""
    rng = np.random.Generator(np.random.PCG64(seed))
    e = rng.uniform(eps[0], eps[1], nbath)
    vv = rng.uniform(v[0], v[1], nbath)
    em = np.zeros((1, 2))
    ek = np.stack([e, e], axis=1)
    vmk = np.stack([vv, vv], axis=1)[None, :, :]
    return em, ek, vmk, U / 2.0


def two_orbital_kanamori(nbath_per_orb, U, seed, eps=(-1.0, 1.0), v=(0.2, 0.5)):
    """Two orbitals, each with its own bath block (vmk block-diagonal)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = 2 * nbath_per_orb
    ek1 = np.zeros(nb)
    vmk = np.zeros((2, nb, 2))
    for m in range(2):
        e = rng.uniform(eps[0], eps[1], nbath_per_orb)
        vv = rng.uniform(v[0], v[1], nbath_per_orb)
        sl = slice(m * nbath_per_orb, (m + 1) * nbath_per_orb)
        ek1[sl] = e
        vmk[m, sl, 0] = vv
        vmk[m, sl, 1] = vv
    em = np.zeros((2, 2))
    ek = np.stack([ek1, ek1], axis=1)
    J = U / 4.0
    Up = U - 2 * J
    # half filling of the Kanamori atom: mu = (U + 2U' - J)/2 for two orbitals
    mu = 0.5 * (U + 2 * Up - J)
    return em, ek, vmk, mu, J, Up


def matsubara(nw, beta):
    """omega_n = (2n+1) pi / beta, n = 0..nw-1 (dmft_grid.f90:42-53)."""
    return (2.0 * np.arange(nw) + 1.0) * np.pi / beta
