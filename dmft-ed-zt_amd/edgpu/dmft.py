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
  fit_bath                   the bath fit of fit_h_imp_params
                             (ed_projection.f90:23-550), conjugate gradients
  run_dmft                   the self-consistency loop of dmft.f90:60-137
                             around any impurity solver

These run on the host (the hybridization, local Green's function, bath fit
and file I/O are O(nw * nbath) or O(nw * 10000) work per DMFT iteration)
except the sweep, which runs on the device through edgpu.Batch, and the
impurity solve that run_dmft is given.
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


# --------------------------------------------------------------- bath fit
#
# fit_h_imp_params (ed_projection.f90:23-115): new (em, ek, vmk) of one spin
# minimize
#   f(x) = 1/(norb nw) sum_{m, n} |em_m - em_m^old + D_m(iw_n; ek, vmk) - D_m^old(iw_n)
#                                  + 1/G_loc^old - 1/G_cl^old|^2 / w_n
# (func :182-213, weight 1/w_n :215-219) over x = (em(:), ek(:), vmk(:,:)) in
# the order of emk_to_x (:117-135), by Polak-Ribiere conjugate gradients
# with the analytic gradient (dfunc :221-280) and a bracketed Brent line
# search (frprmn :313-357, linmin :359-378, mnbrak / brent :380-550): the
# published Numerical Recipes method, written here for numpy.

class _FitProblem:
    """Objective and gradient of one spin's fit in real arithmetic:
    1/(i w - e) = (-e - i w) q with q = 1/(w^2 + e^2), so D, diff and the
    derivatives are a few real matrix products over (nw, nbath) arrays."""

    def __init__(self, em, ek, vmk, D_cl, G_cl, G_loc, omega, ispin):
        self.norb, self.nbath = vmk.shape[0], vmk.shape[1]
        self.s = ispin
        self.w = np.asarray(omega, dtype=np.float64)
        self.w2 = self.w * self.w
        self.rw = 1.0 / self.w                      # weight 1/w_n (:215-219)
        self.em_old = np.array(em[:, ispin], dtype=np.float64)
        # the part of diff that does not depend on x
        c = (-self.em_old[None, :] - D_cl[:, :, ispin] + 1.0 / G_loc[:, :, ispin]
             - 1.0 / G_cl[:, :, ispin])
        self.cr, self.ci = np.ascontiguousarray(c.real), np.ascontiguousarray(c.imag)
        self.nw = self.w.size
        self.scale = 1.0 / (self.norb * self.nw)

    def split(self, x):
        no, nb = self.norb, self.nbath
        em = x[:no]
        ek = x[no:no + nb]
        v = x[no + nb:].reshape(nb, no).T  # vmk(iorb, ibath): iorb fastest in x
        return em, ek, v

    def _diff(self, x):
        em, ek, v = self.split(x)
        q = 1.0 / (self.w2[:, None] + (ek * ek)[None, :])          # (nw, nbath)
        v2 = (v * v).T                                             # (nbath, norb)
        dr = em[None, :] - (q * ek[None, :]) @ v2 + self.cr        # Re diff (nw, norb)
        di = -self.w[:, None] * (q @ v2) + self.ci                 # Im diff
        return em, ek, v, q, dr, di

    def f(self, x):
        _, _, _, _, dr, di = self._diff(x)
        return float((self.rw @ (dr * dr + di * di)).sum() * self.scale)

    def df(self, x):
        em, ek, v, q, dr, di = self._diff(x)
        drw, diw = dr * self.rw[:, None], di * self.rw[:, None]    # d / w_n
        g_em = 2.0 * drw.sum(axis=0)
        # dD_m/de_b = v^2 (e^2 - w^2 + 2 i e w) q^2 ; dD_m/dv_mb = 2 v (-e - i w) q
        q2 = q * q
        re_e = (ek * ek)[None, :] * q2 - self.w2[:, None] * q2    # (nw, nbath)
        im_e = 2.0 * self.w[:, None] * ek[None, :] * q2
        A = drw.T @ re_e + diw.T @ im_e                            # (norb, nbath)
        g_ek = 2.0 * (A * (v * v)).sum(axis=0)
        B = drw.T @ (q * ek[None, :]) + diw.T @ (q * self.w[:, None])
        g_v = -4.0 * v * B
        return np.concatenate([g_em, g_ek, g_v.T.reshape(-1)]) * self.scale


def _mnbrak(f1, ax, bx, gold=1.618034, glimit=100.0, tiny=1e-20):
    fa, fb = f1(ax), f1(bx)
    if fb > fa:
        ax, bx, fa, fb = bx, ax, fb, fa
    cx = bx + gold * (bx - ax)
    fc = f1(cx)
    while fb >= fc:
        r = (bx - ax) * (fb - fc)
        q = (bx - cx) * (fb - fa)
        u = bx - ((bx - cx) * q - (bx - ax) * r) / (2.0 * np.copysign(max(abs(q - r), tiny), q - r))
        ulim = bx + glimit * (cx - bx)
        if (bx - u) * (u - cx) > 0.0:
            fu = f1(u)
            if fu < fc:
                return bx, u, cx
            if fu > fb:
                return ax, bx, u
            u = cx + gold * (cx - bx)
            fu = f1(u)
        elif (cx - u) * (u - ulim) > 0.0:
            fu = f1(u)
            if fu < fc:
                bx, cx, u = cx, u, u + gold * (u - cx)
                fb, fc, fu = fc, fu, f1(u)
        elif (u - ulim) * (ulim - cx) >= 0.0:
            u = ulim
            fu = f1(u)
        else:
            u = cx + gold * (cx - bx)
            fu = f1(u)
        ax, bx, cx = bx, cx, u
        fa, fb, fc = fb, fc, fu
    return ax, bx, cx


def _brent(f1, ax, bx, cx, tol=1e-8, itmax=10000, zeps=1e-10):
    cgold = 0.3819660
    a, b = min(ax, cx), max(ax, cx)
    x = w = v = bx
    fx = fw = fv = f1(x)
    d = e = 0.0
    for _ in range(itmax):
        xm = 0.5 * (a + b)
        tol1 = tol * abs(x) + zeps
        tol2 = 2.0 * tol1
        if abs(x - xm) <= tol2 - 0.5 * (b - a):
            return x, fx
        golden = True
        if abs(e) > tol1:
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2.0 * (q - r)
            if q > 0.0:
                p = -p
            q = abs(q)
            etemp, e = e, d
            if not (abs(p) >= abs(0.5 * q * etemp) or p <= q * (a - x) or p >= q * (b - x)):
                d = p / q
                u = x + d
                if u - a < tol2 or b - u < tol2:
                    d = np.copysign(tol1, xm - x)
                golden = False
        if golden:
            e = (a - x) if x >= xm else (b - x)
            d = cgold * e
        u = x + d if abs(d) >= tol1 else x + np.copysign(tol1, d)
        fu = f1(u)
        if fu <= fx:
            if u >= x:
                a = x
            else:
                b = x
            v, fv, w, fw, x, fx = w, fw, x, fx, u, fu
        else:
            if u < x:
                a = u
            else:
                b = u
            if fu <= fw or w == x:
                v, fv, w, fw = w, fw, u, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    raise RuntimeError("brent: maximum iterations exceeded")


def conjugate_gradient(f, df, p, ftol=1e-6, itmax=20000, eps=1e-10):
    """Polak-Ribiere minimization of f from p (frprmn): returns (p, fret,
    iterations).  Stops when 2|f_new - f_old| <= ftol (|f_new| + |f_old| + eps)
    after a line minimization, or when the gradient vanishes."""
    p = np.array(p, dtype=np.float64)
    fp = f(p)
    xi = df(p)
    g = -xi
    h = g.copy()
    xi = h.copy()
    for it in range(1, itmax + 1):
        f1 = lambda t: f(p + t * xi)
        ax, bx, cx = _mnbrak(f1, 0.0, 1.0)
        xmin, fret = _brent(f1, ax, bx, cx)
        p = p + xmin * xi
        if 2.0 * abs(fret - fp) <= ftol * (abs(fret) + abs(fp) + eps):
            return p, fret, it
        fp = fret
        xi = df(p)
        gg = float(g @ g)
        if gg == 0.0:
            return p, fret, it
        gam = float((xi + g) @ xi) / gg
        g = -xi
        h = g + gam * h
        xi = h.copy()
    raise RuntimeError("conjugate_gradient: maximum iterations exceeded")


def fit_bath(em, ek, vmk, D_cl, G_cl, G_loc, omega, nspin=1, tol=1e-6):
    """New (em, ek, vmk) from the previous iteration's D_cl, G_cl and G_loc
    (fit_h_imp_params, ed_projection.f90:23-115).  Arrays in the reference
    layouts em(norb,2), ek(nbath,2), vmk(norb,nbath,2), D_cl/G_cl/G_loc
    (nw,norb,nspin).  Returns (em, ek, vmk, [(iterations, f) per spin]); for
    nspin = 1 spin 2 copies spin 1 (:95-100)."""
    em = np.array(em, dtype=np.float64)
    ek = np.array(ek, dtype=np.float64)
    vmk = np.array(vmk, dtype=np.float64)
    D_cl, G_cl, G_loc = (np.asarray(a, dtype=np.complex128) for a in (D_cl, G_cl, G_loc))
    info = []
    for s in range(nspin):
        prob = _FitProblem(em, ek, vmk, D_cl, G_cl, G_loc, omega, s)
        x0 = np.concatenate([em[:, s], ek[:, s], vmk[:, :, s].T.reshape(-1)])
        x, fret, it = conjugate_gradient(prob.f, prob.df, x0, ftol=tol)
        e, k, v = prob.split(x)
        em[:, s], ek[:, s], vmk[:, :, s] = e, k, v
        info.append((it, fret))
    if nspin == 1:
        em[:, 1], ek[:, 1], vmk[:, :, 1] = em[:, 0], ek[:, 0], vmk[:, :, 0]
    return em, ek, vmk, info


def scf_difference(G_cl, G_loc, nbath):
    """test_convergence (dmft.f90:111-137): sum |G_cl - G_loc|^2 / (norb nbath nw)."""
    nw, norb, nspin = np.shape(G_cl)
    return float((np.abs(np.asarray(G_cl) - np.asarray(G_loc)) ** 2).sum() / (norb * nbath * nw))


def run_dmft(solve, em, ek, vmk, omega, nspin=1, nloop=100, scf_tol=1e-6, log=None, trace=False):
    """The self-consistency loop of dmft.f90:60-104 on the Bethe lattice
    (half bandwidth 1): per iteration, from the second on, fit the bath to
    the previous iteration, then D_cl, the impurity solve, G_loc, and the
    convergence test.  solve(em, ek, vmk) -> G_cl(nw, norb, nspin) is the
    impurity solver (ground state + cluster Green's function).  Returns a
    dict with the final em/ek/vmk, D_cl, G_cl, G_loc and the history of
    the convergence measure (with trace=True also every iteration's
    (em, ek, vmk, G_cl))."""
    em, ek, vmk = (np.array(a, dtype=np.float64) for a in (em, ek, vmk))
    hist, steps = [], []
    D_cl = G_cl = G_loc = None
    converged = False
    for it in range(1, nloop + 1):
        if it > 1:
            em, ek, vmk, _ = fit_bath(em, ek, vmk, D_cl, G_cl, G_loc, omega, nspin=nspin)
        D_cl = cluster_hybridization_ftn(ek, vmk, omega, nspin=nspin)
        G_cl = solve(em, ek, vmk)
        G_loc = local_green_ftn(em, D_cl, G_cl)
        d = scf_difference(G_cl, G_loc, ek.shape[0])
        hist.append(d)
        if trace:
            steps.append((em.copy(), ek.copy(), vmk.copy(), np.array(G_cl)))
        if log:
            log(f"DMFT iteration {it}: |G_cl - G_loc|^2 = {d:.3e}")
        if d < scf_tol:
            converged = True
            break
    return dict(em=em, ek=ek, vmk=vmk, D_cl=D_cl, G_cl=G_cl, G_loc=G_loc, history=hist,
                converged=converged, iterations=len(hist), trace=steps)
