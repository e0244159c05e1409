"""Independent dense constructions that pin the oracle (test infrastructure).

Nothing here reuses the reference's sign shortcuts (sgn / sgn2): every
operator is applied as an explicit product of Jordan-Wigner fermion
operators, c+_p |s> = (-1)^{sum_{q<p} n_q} |s + e_p>, with the mode order of
the reference (all up modes, then all down modes; mode = (spin-1)*nsite +
site-1).  The Hamiltonian is written in its textbook second-quantized
Kanamori form:

  H = sum_{m,s} (em_{ms} - mu) n_{ms} + U sum_m n_{m,up} n_{m,dn}
    + sum_{m<m'} [(U'-J)(n_{m up} n_{m' up} + n_{m dn} n_{m' dn})
                  + U'(n_{m up} n_{m' dn} + n_{m dn} n_{m' up})]
    + sum_{m,l,s} V_{mls} (c+_{ms} b_{ls} + b+_{ls} c_{ms})
    + sum_{l,s} eps_{ls} n_{ls}
    - Jp sum_{m != m'} c+_{m up} c_{m dn} c+_{m' dn} c_{m' up}
    + Jp sum_{m != m'} c+_{m up} c+_{m dn} c_{m' dn} c_{m' up}

so a mismatch with the oracle's literal restatement of hx exposes a sign,
index or term error in either.
"""
import itertools

import numpy as np


def sector_states(nsite, nu, nd):
    """Fock states of sector (nu, nd) in the row order 'up rank fastest,
    ranks in increasing integer order' built with itertools, not with the
    reference's range scan."""
    ups = sorted(sum(1 << p for p in c) for c in itertools.combinations(range(nsite), nu))
    dns = sorted(sum(1 << p for p in c) for c in itertools.combinations(range(nsite), nd))
    return np.array([u + (d << nsite) for d in dns for u in ups], dtype=np.int64)


def _cdag(p, s):
    if (s >> p) & 1:
        return None, 0
    sign = -1 if bin(s & ((1 << p) - 1)).count("1") & 1 else 1
    return s | (1 << p), sign


def _c(p, s):
    if not (s >> p) & 1:
        return None, 0
    sign = -1 if bin(s & ((1 << p) - 1)).count("1") & 1 else 1
    return s & ~(1 << p), sign


def apply_string(ops, s):
    """Apply an operator string [(dag, mode), ...] written left-to-right as in
    the formula (rightmost acts first).  Returns (state, sign) or (None, 0)."""
    sign = 1
    for dag, p in reversed(ops):
        s, sg = (_cdag if dag else _c)(p, s)
        if s is None:
            return None, 0
        sign *= sg
    return s, sign


def hamiltonian_terms(norb, nbath, em, ek, vmk, U, Up, J, Jp, mu):
    """List of (coefficient, operator string) for H above."""
    nsite = norb + nbath
    em, ek, vmk = (np.asarray(a, dtype=np.float64) for a in (em, ek, vmk))

    def mode(site, spin):  # site, spin 1-based
        return (spin - 1) * nsite + site - 1

    T = []
    for m in range(1, norb + 1):
        for s in (1, 2):
            T.append((em[m - 1, s - 1] - mu, [(1, mode(m, s)), (0, mode(m, s))]))
        T.append((U, [(1, mode(m, 1)), (0, mode(m, 1)), (1, mode(m, 2)), (0, mode(m, 2))]))
        for m2 in range(m + 1, norb + 1):
            for s1, s2, coef in ((1, 1, Up - J), (2, 2, Up - J), (1, 2, Up), (2, 1, Up)):
                T.append((coef, [(1, mode(m, s1)), (0, mode(m, s1)),
                                 (1, mode(m2, s2)), (0, mode(m2, s2))]))
        for l in range(1, nbath + 1):
            for s in (1, 2):
                v = vmk[m - 1, l - 1, s - 1]
                if v == 0.0:
                    continue
                T.append((v, [(1, mode(m, s)), (0, mode(norb + l, s))]))
                T.append((v, [(1, mode(norb + l, s)), (0, mode(m, s))]))
        for m2 in range(1, norb + 1):
            if m2 == m:
                continue
            T.append((-Jp, [(1, mode(m, 1)), (0, mode(m, 2)), (1, mode(m2, 2)), (0, mode(m2, 1))]))
            T.append((Jp, [(1, mode(m, 1)), (1, mode(m, 2)), (0, mode(m2, 2)), (0, mode(m2, 1))]))
    for l in range(1, nbath + 1):
        for s in (1, 2):
            T.append((ek[l - 1, s - 1], [(1, mode(norb + l, s)), (0, mode(norb + l, s))]))
    return T


def dense_h(norb, nbath, em, ek, vmk, U, Up, J, Jp, mu, nu, nd):
    """Dense H in sector (nu, nd), rows/cols in sector_states order."""
    nsite = norb + nbath
    states = sector_states(nsite, nu, nd)
    index = {int(s): i for i, s in enumerate(states)}
    H = np.zeros((len(states), len(states)))
    terms = hamiltonian_terms(norb, nbath, em, ek, vmk, U, Up, J, Jp, mu)
    for j, s in enumerate(states):
        for coef, ops in terms:
            t, sg = apply_string(ops, int(s))
            if t is not None:
                H[index[t], j] += coef * sg
    return H, states


def apply_op(states_in, vec, dag, mode, states_out):
    """Dense c+_mode (dag=1) or c_mode (dag=0) from one sector to another."""
    index = {int(s): i for i, s in enumerate(states_out)}
    out = np.zeros(len(states_out))
    for j, s in enumerate(states_in):
        t, sg = (_cdag if dag else _c)(mode, int(s))
        if t is not None:
            out[index[t]] += sg * vec[j]
    return out


def lehmann_green(norb, nbath, em, ek, vmk, U, Up, J, Jp, mu, nu, nd, iorb, ispin, omega):
    """Exact T=0 G(i w) of orbital iorb, spin ispin from the lowest state of
    sector (nu, nd): sum_n |<n|c+|0>|^2/(iw+E0-En) + |<n|c|0>|^2/(iw-E0+En).
    Returns (G, E0, weight_plus, weight_minus)."""
    nsite = norb + nbath
    args = (norb, nbath, em, ek, vmk, U, Up, J, Jp, mu)
    H0, st0 = dense_h(*args, nu, nd)
    w0, v0 = np.linalg.eigh(H0)
    E0, gs = w0[0], v0[:, 0]
    mode = (ispin - 1) * nsite + iorb - 1
    iw = 1j * np.asarray(omega)
    G = np.zeros(len(omega), dtype=np.complex128)
    wts = []
    for dag, dn in ((1, 1), (0, -1)):
        nu2, nd2 = (nu + dn, nd) if ispin == 1 else (nu, nd + dn)
        if not (0 <= nu2 <= nsite and 0 <= nd2 <= nsite):
            wts.append(0.0)
            continue
        H1, st1 = dense_h(*args, nu2, nd2)
        w1, v1 = np.linalg.eigh(H1)
        phi = apply_op(st0, gs, dag, mode, st1)
        amp = v1.T @ phi
        wts.append(float(amp @ amp))
        if dag:
            G += ((amp ** 2)[None, :] / (iw[:, None] + E0 - w1[None, :])).sum(1)
        else:
            G += ((amp ** 2)[None, :] / (iw[:, None] - E0 + w1[None, :])).sum(1)
    return G, E0, wts[0], wts[1]


def noninteracting(norb_is_one_em, ek, v, mu, omega):
    """U=0 single-orbital closed forms: G(iw) = 1/(iw + mu - em - Delta(iw)),
    Delta = sum_l V_l^2/(iw - eps_l) (dmft_green.f90:24-45), and the
    one-body matrix h of the star geometry (impurity level em - mu)."""
    ek = np.asarray(ek, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    iw = 1j * np.asarray(omega)
    delta = (v[None, :] ** 2 / (iw[:, None] - ek[None, :])).sum(1)
    G = 1.0 / (iw + mu - norb_is_one_em - delta)
    n = 1 + len(ek)
    h = np.zeros((n, n))
    h[0, 0] = norb_is_one_em - mu
    h[0, 1:] = h[1:, 0] = v
    h[np.arange(1, n), np.arange(1, n)] = ek
    return G, h
