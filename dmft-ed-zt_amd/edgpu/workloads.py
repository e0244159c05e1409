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


def batched_problems(edgpu, n=512, seed=7, ns_choices=(8, 10, 12), U=(0.0, 6.0), filling=(0.8, 1.2),
                     eps=(-1.5, 1.5), v=(0.1, 0.7), maxnstep=200, maxnstep_green=150):
    """Independent single-orbital problems of the batched workload: Ns
    uniform over ns_choices, U ~ U[U], filling n ~ U[filling] with
    N = round(n Ns), N_up = ceil(N/2), N_dn = floor(N/2) (the reference takes
    sectors from its input), bath eps ~ U[eps], V ~ U[v], spin-symmetric."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        ns = int(rng.choice(ns_choices))
        u = float(rng.uniform(*U))
        fill = float(rng.uniform(*filling))
        nb = ns - 1
        e = rng.uniform(eps[0], eps[1], nb)
        vv = rng.uniform(v[0], v[1], nb)
        ntot = int(round(fill * ns))
        nup, ndn = (ntot + 1) // 2, ntot // 2
        em = np.zeros((1, 2))
        ek = np.stack([e, e], axis=1)
        vmk = np.stack([vv, vv], axis=1)[None, :, :]
        out.append(edgpu.Problem(nb, u, u / 2.0, em, ek, vmk, nup, ndn, maxnstep=maxnstep,
                                 maxnstep_green=maxnstep_green, seed=seed * 100003 + i))
    return out


def served_requests(edgpu, rate=200.0, duration=60.0, seed=11, ns_weights=((10, 0.5), (12, 0.35), (14, 0.15)),
                    U=(0.0, 8.0), maxnstep=200, maxnstep_green=150, both_spins=True):
    """Poisson arrivals at `rate` per second over `duration` seconds of
    half-filled single-orbital problems: Ns drawn with ns_weights, U ~ U[U],
    seeded baths eps ~ U[-1,1], V ~ U[0.2,0.6]; each solved with the
    full-size step counts, 200 ground-state steps and four c+/c chains
    (both spins) of 150.  Returns [(t_arrival, Problem)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns_vals = np.array([w[0] for w in ns_weights])
    p = np.array([w[1] for w in ns_weights], dtype=float)
    p /= p.sum()
    t, out, i = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        ns = int(rng.choice(ns_vals, p=p))
        u = float(rng.uniform(*U))
        nb = ns - 1
        e = rng.uniform(-1.0, 1.0, nb)
        vv = rng.uniform(0.2, 0.6, nb)
        em = np.zeros((1, 2))
        ek = np.stack([e, e], axis=1)
        vmk = np.stack([vv, vv], axis=1)[None, :, :]
        out.append((t, edgpu.Problem(nb, u, u / 2.0, em, ek, vmk, ns // 2, ns // 2, maxnstep=maxnstep,
                                     maxnstep_green=maxnstep_green, both_spins=both_spins,
                                     seed=seed * 1000003 + i)))
        i += 1
