"""Host-side pieces of the DMFT step on either side of the impurity solve,
with the reference's names and array layouts (norb, nspin and nw as in
dmft_green.f90), for callers that drive the self-consistency from Python:

  cluster_hybridization_ftn  Delta(i w) of the bath (dmft_green.f90:24-45)
  local_green_ftn            G_loc of the Bethe lattice by the reference's
                             10000-point rectangle rule (dmft_green.f90:128-170)
  write_g_coeffs             g_coeffs.dat as written by cluster_green_ftn
                             (dmft_green.f90:66-69, 89-109); read_g_coeffs reads it
  spectral_function          A(w) on the real axis from the chain coefficients,
                             taking the c+ chain from (ap, bp) and the c chain
                             from (am, bm)
  self_energy, write_sigma   Sigma(i w) of dump_sigma and its sigma.dat.<m>.<s>
                             files (post_process.f90:128-185)
  quasiparticle_weight       Z = 1 / (1 - Im Sigma(i w_1) / w_1) (post_process.f90:40-56)
  ground_state_sweep         the sector sweep of ground_state.f90:30-73 with
                             every candidate sector in one batched solve

These run on the host (the hybridization, local Green's function and file
I/O are O(nw * nbath) or O(nw * 10000) work per DMFT iteration) except the
sweep, which runs on the device through edgpu.Batch.
"""
import numpy as np


def cluster_hybridization_ftn(ek, vmk, omega, nspin=1):
    """D_cl(nw, norb, nspin) = sum_b vmk(iorb,b,ispin)^2 / (i w - ek(b,ispin))."""
    ek = np.asarray(ek, dtype=np.float64)
    vmk = np.asarray(vmk, dtype=np.float64)
    iw = 1j * np.asarray(omega, dtype=np.float64)
    norb = vmk.shape[0]
    D = np.zeros((iw.size, norb, nspin), dtype=np.complex128)
    for s in range(nspin):
        for m in range(norb):
            D[:, m, s] = ((vmk[m, :, s] ** 2)[None, :] / (iw[:, None] - ek[None, :, s])).sum(axis=1)
    return D


def local_green_ftn(em, D_cl, G_cl, hfield=0.0, ne=10000):
    """G_loc(nw, norb, nspin) of the Bethe lattice (half bandwidth 1):
    (2/(ne-1)) sum_x (2/pi) sqrt(1-x^2) / (em + h_s + D_cl + 1/G_cl - x) on
    the ne-point grid x = -1 .. 1, h = (+hfield, -hfield) by spin -- the
    reference's rule, including its sign conventions."""
    em = np.asarray(em, dtype=np.float64)
    D_cl = np.asarray(D_cl, dtype=np.complex128)
    G_cl = np.asarray(G_cl, dtype=np.complex128)
    nw, norb, nspin = G_cl.shape
    x = -1.0 + np.arange(ne) * 2.0 / (ne - 1)
    w = 2.0 / np.pi * np.sqrt(np.maximum(1.0 - x * x, 0.0))
    hs = (hfield, -hfield)
    G = np.zeros_like(G_cl)
    for s in range(nspin):
        for m in range(norb):
            zeta = em[m, s] + hs[s] + D_cl[:, m, s] + 1.0 / G_cl[:, m, s]
            G[:, m, s] = (w[None, :] / (zeta[:, None] - x[None, :])).sum(axis=1) / (ne - 1) * 2.0
    return G


def write_g_coeffs(path, e0, coeffs, norb, maxnstep):
    """g_coeffs.dat: `maxnstep norb`, then per orbital (outer) and spin
    (inner) the line `iorb ispin E0` and maxnstep lines `ap bp am bm`.
    coeffs[(iorb, ispin)] = (ap, bp, am, bm) as returned by
    Solver.cluster_green_ftn (spin 2 repeats spin 1 when the reference's
    spin symmetry skipped it)."""
    with open(path, "w") as f:
        f.write(f" {maxnstep:11d} {norb:11d}\n")
        for iorb in range(1, norb + 1):
            for ispin in (1, 2):
                f.write(f"{iorb:4d}{ispin:4d} {e0:25.16f}\n")
                ap, bp, am, bm = (np.asarray(c, dtype=np.float64) for c in coeffs[(iorb, ispin)])
                for i in range(maxnstep):
                    f.write(f"{ap[i]:25.16f} {bp[i]:25.16f} {am[i]:25.16f} {bm[i]:25.16f} \n")


def read_g_coeffs(path):
    """(e0, coeffs, norb, maxnstep) from a g_coeffs.dat in the layout above."""
    with open(path) as f:
        maxnstep, norb = (int(t) for t in f.readline().split())
        coeffs, e0 = {}, None
        for _ in range(2 * norb):
            iorb, ispin, ev = f.readline().split()
            e0 = float(ev)
            rows = np.array([[float(t) for t in f.readline().split()] for _ in range(maxnstep)])
            coeffs[(int(iorb), int(ispin))] = tuple(rows[:, q].copy() for q in range(4))
    return e0, coeffs, norb, maxnstep


def _cfrac(z, a, b, sign):
    """b1^2/(z + sign*a1 - b2^2/(z + sign*a2 - ...)) over the nonzero levels,
    vectorized over z (levels past the last nonzero b are dropped)."""
    n = int(np.max(np.nonzero(np.asarray(b))[0]) + 1) if np.any(b) else 0
    if n == 0:
        return np.zeros_like(z)
    f = np.zeros_like(z)
    for i in range(n - 1, -1, -1):
        f = b[i] * b[i] / (z + sign * a[i] - f)
    return f


def spectral_function(e0, coeffs, w, eta=0.02):
    """A(w) = -Im[G+(w + E0 + i eta) + G-(w - E0 + i eta)] / pi, with G+ from
    the c+ chain (ap, bp) and G- from the c chain (am, bm).  (The reference's
    utils/spectral_ftn.py:42 evaluates the second term with ap, bp.)"""
    ap, bp, am, bm = (np.asarray(c, dtype=np.float64) for c in coeffs)
    w = np.asarray(w, dtype=np.float64)
    gp = _cfrac(w + e0 + 1j * eta, ap, bp, -1.0)
    gm = _cfrac(w - e0 + 1j * eta, am, bm, +1.0)
    return -(gp + gm).imag / np.pi


def ground_state_sweep(edgpu, omega, nbath, U, mu, em, ek, vmk, sectors, maxnstep=200, seed=1,
                       precision="fp32", batch=None):
    """ground_state.f90:30-73 over `sectors` [(ne_up, ne_dn), ...] with all
    candidate sectors solved in one batch (ground state only).  Sector i
    starts from seed + i - 1 as in edg_ground_state.  Returns (isector,
    e0, e0 per sector) with the strict minimum, first lowest winning
    (:62); isector is 1-based."""
    own = batch is None
    b = edgpu.Batch(omega, precision=precision) if own else batch
    ids = []
    for i, (nu, nd) in enumerate(sectors):
        p = edgpu.Problem(nbath, U, mu, em, ek, vmk, nu, nd, maxnstep=maxnstep, seed=seed + i,
                          gs_only=True)
        ids.append(b.submit(p))
    pending = len(ids)
    res = {}
    while pending:
        pending = b.step(64)
        for r in b.poll():
            res[r["id"]] = r
    for r in b.poll():
        res[r["id"]] = r
    if own:
        b.close()
    e = [res[i]["e0"] for i in ids]
    best, bi = np.inf, 0
    for i, v in enumerate(e):
        if v < best:
            best, bi = v, i + 1
    return bi, best, e


def self_energy(em, D_cl, G_cl, omega, mu):
    """Sigma(nw, norb, nspin) = i w + mu - em - D_cl - 1/G_cl (dump_sigma)."""
    em = np.asarray(em, dtype=np.float64)
    iw = 1j * np.asarray(omega, dtype=np.float64)
    G_cl = np.asarray(G_cl, dtype=np.complex128)
    nw, norb, nspin = G_cl.shape
    S = np.zeros_like(G_cl)
    for s in range(nspin):
        for m in range(norb):
            S[:, m, s] = iw + mu - em[m, s] - D_cl[:, m, s] - 1.0 / G_cl[:, m, s]
    return S


def quasiparticle_weight(Sigma, omega):
    """Z(norb, nspin) from the first Matsubara point."""
    zq = np.asarray(Sigma)[0].imag / float(np.asarray(omega)[0])
    return 1.0 / (1.0 - zq)


def write_sigma(omega, Sigma, directory="."):
    """sigma.dat.<iorb>.<ispin>: header `nw iorb ispin`, then per frequency
    w, Re Sigma, Im Sigma in F12.8 fields."""
    import os
    nw, norb, nspin = Sigma.shape
    paths = []
    for s in range(nspin):
        for m in range(norb):
            path = os.path.join(directory, f"sigma.dat.{m + 1}.{s + 1}")
            with open(path, "w") as f:
                f.write(f" {nw:11d} {m + 1:11d} {s + 1:11d}\n")
                for iw in range(nw):
                    f.write(f"{omega[iw]:12.8f}{Sigma[iw, m, s].real:12.8f}{Sigma[iw, m, s].imag:12.8f}\n")
            paths.append(path)
    return paths
