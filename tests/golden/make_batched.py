"""Writes tests/golden/batched_oracle.npz: the FP64 CPU oracle's solves of
problems of the batched workload (edgpu.workloads.batched_problems, n=512,
seed 7: the problems the bench's `batched` figure solves), at the Ns = 10
and 12 size classes, with the full step counts.

Run on the CPU (about a minute on 8 cores):
    python tests/golden/make_batched.py

For each chosen problem: the ground state in its sector from 200 Lanczos
steps of the counter-based start vector of the problem's seed
(ground_state.f90:28-73, lanczos_solver.f90:20-90, 213-300) and G(i w_n) on
2048 Matsubara points at beta = 100 from c+/c chains of 150 steps
(dmft_green.f90:47-126, 177-260; spin-averaged when N_up != N_dn,
dmft_green.f90:76-118).

Stored: index (positions in the 512-problem list), E0 (per problem),
G (problems, 2048) complex, omega.
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
import edgpu  # noqa: E402
from edgpu import workloads as W  # noqa: E402

NW, BETA = 2048, 100.0
PER_CLASS = {10: 2, 12: 2}


def chosen(probs):
    """The first PER_CLASS[ns] problems of each class, and for each class
    one with N_up != N_dn when the first ones are all spin-balanced."""
    out = []
    for ns, k in PER_CLASS.items():
        cls = [i for i, p in enumerate(probs) if p.nbath + 1 == ns]
        pick = cls[:k]
        if all(probs[i].nup == probs[i].ndn for i in pick):
            odd = [i for i in cls if probs[i].nup != probs[i].ndn]
            if odd:
                pick[-1] = odd[0]
        out += pick
    return sorted(out)


def solve(pr, omega):
    m = O.Model(1, pr.nbath, U=pr.U, mu=pr.mu)
    em = pr.em.reshape(1, 2, order="F")
    ek = pr.ek.reshape(pr.nbath, 2, order="F")
    vmk = pr.vmk.reshape(1, pr.nbath, 2, order="F")
    gs = O.ground_state(m, em, ek, vmk, [(pr.nup, pr.ndn)], pr.maxnstep, seed=pr.seed)
    G, _ = O.cluster_green_ftn(m, em, ek, vmk, gs, pr.maxnstep_green or pr.maxnstep, omega, nspin=1,
                               both_spins=pr.both_spins)
    return gs.val, G[:, 0, 0]


def main():
    O.build()
    O.set_threads(os.cpu_count() or 8)
    omega = O.matsubara(NW, BETA)
    probs = W.batched_problems(edgpu, n=512)
    idx = chosen(probs)
    e0s, gs = [], []
    for i in idx:
        t = time.time()
        e0, G = solve(probs[i], omega)
        e0s.append(e0)
        gs.append(G)
        p = probs[i]
        print(f"problem {i}: Ns={p.nbath + 1} sector ({p.nup},{p.ndn}) U={p.U:.3f} E0={e0:.12f} "
              f"{time.time() - t:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "batched_oracle.npz"), index=np.array(idx), E0=np.array(e0s),
                        G=np.array(gs), omega=omega)


if __name__ == "__main__":
    main()
