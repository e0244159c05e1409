"""Writes tests/golden/ns14_oracle.npz: the FP64 CPU oracle's whole solve at
full size, the reference answer of the GPU solves at Ns=14.

Run on the CPU (about 15 minutes on 8 cores):
    python tests/golden/make_ns14.py

Problems (the seeded baths of edgpu.workloads, seed 1), each in sector (7,7)
with D = 3432^2 = 11,778,624 states:
  u1, u2, u4  single-orbital Bethe impurity, Nbath=13, U = 1, 2, 4,
              half filling (em = 0, mu = U/2)
  kan         two-orbital Kanamori impurity, 6 bath sites per orbital
              (Nbath=12), U=3, J=U/4, U'=U-2J, Jp=J, mu = (U+2U'-J)/2

For each: 200 Lanczos steps for the ground state from the counter-based
start vector of seed 1 (ground_state.f90:28-73, lanczos_solver.f90:20-90,
213-300), the occupancies and double occupancy (post_process.f90:63-112),
and G(i w_n) on 2048 Matsubara points at beta=100 from c+/c chains of 150
steps (dmft_green.f90:47-126,177-260), with the chains' ap/bp/am/bm.

Stored per problem prefix P: P_E0, P_occ (norb,2), P_docc (norb),
P_G (2048, norb) complex, P_coef (norb, 4, 150) = ap, bp, am, bm of spin 1
(the spin-symmetric bath and N_up = N_dn make spin 2 equal: `even`,
dmft_green.f90:76-80), P_nstep_gs.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dmft-ed-zt_amd"))

from oracle import oracle as O  # noqa: E402
from edgpu import workloads as W  # noqa: E402

NW, BETA, GS_STEPS, CHAIN_STEPS, SEED = 2048, 100.0, 200, 150, 1


def problems():
    out = []
    for U in (1.0, 2.0, 4.0):
        em, ek, vmk, mu = W.single_orbital(13, U, seed=SEED)
        out.append((f"u{int(U)}", O.Model(1, 13, U=U, mu=mu), em, ek, vmk))
    em, ek, vmk, mu, J, Up = W.two_orbital_kanamori(6, 3.0, seed=SEED)
    out.append(("kan", O.Model(2, 12, U=3.0, J=J, Up=Up, Jp=J, mu=mu), em, ek, vmk))
    return out


def solve(m, em, ek, vmk):
    om = O.matsubara(NW, BETA)
    b = O.Basis(m.nsite, 7, 7)
    e0, vec, nstep, res = O.diag_lanczos(m, em, ek, vmk, b, O.start_vector(SEED, b.ntot), GS_STEPS)
    gs = O.EigPair(1, e0, vec, b)
    occ, docc = O.observables(m, b, vec)
    G, coeffs = O.cluster_green_ftn(m, em, ek, vmk, gs, CHAIN_STEPS, om, nspin=1)
    coef = np.array([np.stack(coeffs[(io, 1)]) for io in range(1, m.norb + 1)])
    return dict(E0=e0, occ=occ, docc=docc, G=G[:, :, 0], coef=coef, nstep_gs=nstep, residual=res)


def main():
    O.build()
    O.set_threads(os.cpu_count() or 8)
    out = {"omega": O.matsubara(NW, BETA)}
    for name, m, em, ek, vmk in problems():
        t = time.time()
        r = solve(m, em, ek, vmk)
        for k, v in r.items():
            out[f"{name}_{k}"] = v
        print(f"{name}: E0={r['E0']:.12f} occ={r['occ'].ravel()} docc={r['docc']} "
              f"residual={r['residual']:.2e} {time.time() - t:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "ns14_oracle.npz"), **out)


if __name__ == "__main__":
    main()
