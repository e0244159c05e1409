"""Writes the golden fixtures of tests/golden/ from dense exact diagonalization.

The reference ships no golden vectors, known-answer tests or sample outputs,
so these fixtures are computed here WITHOUT the C oracle and without the HIP
library: explicit Jordan-Wigner fermion operators (oracle/dense.py), numpy
eigh and the Lehmann sum.  They pin both the oracle and the GPU path.

  bethe_ns6.npz     single-orbital Bethe bath, Nbath=5 (Ns=6), U=2, half
                    filling sector (3,3), D=400, seed 2024: E0, gs occupancy
                    and double occupancy, G(i w_n) on the 2048-point beta=100
                    grid from the Lehmann sum, and the sum-rule weights.
  kanamori_ns4.npz  two orbitals + 2 bath sites (Ns=4), U=3, J=U/4, sector
                    (2,2): the dense Hamiltonian matrix.
  tables.json       SHA-256 of the int32 up-spin rank tables of sectors
                    N=3 of Ns=6, N=4 of Ns=8 and N=7 of Ns=14 (patterns in
                    increasing integer order), from itertools.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dmft-ed-zt_amd"))

from oracle import dense as Dn  # noqa: E402
from edgpu import workloads as W  # noqa: E402


def bethe_ns6():
    em, ek, vmk, mu = W.single_orbital(5, 2.0, seed=2024)
    U, J = 2.0, 0.3
    Up, Jp = U - 2 * J, J
    om = W.matsubara(2048, 100.0)
    H, st = Dn.dense_h(1, 5, em, ek, vmk, U, Up, J, Jp, mu, 3, 3)
    w, v = np.linalg.eigh(H)
    gs = v[:, 0]
    up = st & ((1 << 6) - 1)
    dn = st >> 6
    n_up = float(np.sum(gs ** 2 * (up & 1)))
    n_dn = float(np.sum(gs ** 2 * (dn & 1)))
    docc = float(np.sum(gs ** 2 * (up & 1) * (dn & 1)))
    G, E0, wp, wm = Dn.lehmann_green(1, 5, em, ek, vmk, U, Up, J, Jp, mu, 3, 3, 1, 1, om)
    np.savez(os.path.join(HERE, "bethe_ns6.npz"), em=em, ek=ek, vmk=vmk, mu=mu, U=U, J=J,
             omega=om, E0=w[0], E1=w[1], n_up=n_up, n_dn=n_dn, docc=docc, G=G, w_plus=wp,
             w_minus=wm)


def kanamori():
    rng = np.random.Generator(np.random.PCG64(77))
    em = rng.uniform(-0.5, 0.5, (2, 2))
    ek = rng.uniform(-1, 1, (2, 2))
    vmk = rng.uniform(0.2, 0.5, (2, 2, 2))
    U = 3.0
    J = U / 4
    Up, Jp = U - 2 * J, J
    mu = 1.1
    H, st = Dn.dense_h(2, 2, em, ek, vmk, U, Up, J, Jp, mu, 2, 2)
    np.savez(os.path.join(HERE, "kanamori_ns4.npz"), em=em, ek=ek, vmk=vmk, U=U, J=J, Up=Up,
             Jp=Jp, mu=mu, H=H, states=st)


def tables():
    out = {}
    for ns, n in ((6, 3), (8, 4), (14, 7)):
        pats = sorted(sum(1 << p for p in c) for c in itertools.combinations(range(ns), n))
        arr = np.array(pats, dtype="<i4")
        out[f"ns{ns}_n{n}_patterns"] = hashlib.sha256(arr.tobytes()).hexdigest()
        out[f"ns{ns}_n{n}_count"] = len(pats)
    with open(os.path.join(HERE, "tables.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    bethe_ns6()
    kanamori()
    tables()
    print("fixtures written to", HERE)
