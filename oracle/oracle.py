"""Python face of the FP64 CPU oracle (ed_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
(libedgpu.so) never calls into it.

Each wrapper names the reference routine it restates (paths relative to the
reference tree).  Arrays follow the reference's Fortran shapes:
em(norb,2), ek(nbath,2), vmk(norb,nbath,2), column-major FP64.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build(force=False):
    """Compile ed_oracle.c with gcc (-O2 -fopenmp) into oracle/liboracle.so."""
    src = os.path.join(HERE, "ed_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-C", HERE, "-s"])
    return LIB_PATH


_lib = None


class _Model(C.Structure):
    _fields_ = [("norb", C.c_int), ("nbath", C.c_int), ("nsite", C.c_int),
                ("U", C.c_double), ("Up", C.c_double), ("Jex", C.c_double),
                ("Jp", C.c_double), ("mu", C.c_double)]


class _Basis(C.Structure):
    _fields_ = [("nsite", C.c_int), ("ne_up", C.c_int), ("ne_dn", C.c_int),
                ("nup", C.c_int), ("ndn", C.c_int), ("ntot", C.c_long),
                ("up", C.POINTER(C.c_int)), ("dn", C.POINTER(C.c_int)),
                ("idx_up", C.POINTER(C.c_int)), ("idx_dn", C.POINTER(C.c_int))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_generate_basis.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(_Basis)]
        L.orc_free_basis.argtypes = [C.POINTER(_Basis)]
        L.orc_hx.argtypes = [C.POINTER(_Model), dp, dp, dp, C.POINTER(_Basis), dp, dp, C.c_int]
        L.orc_lanczos_iteration.argtypes = [C.POINTER(_Model), dp, dp, dp, C.POINTER(_Basis),
                                            dp, C.c_int, C.POINTER(C.c_int), dp, dp]
        L.orc_tridiag_eig.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_tridiag_eig.restype = C.c_int
        L.orc_lanczos_ground_state.argtypes = [C.POINTER(_Model), dp, dp, dp, C.POINTER(_Basis),
                                               dp, C.c_int, dp, dp, dp, C.c_double, dp,
                                               C.POINTER(C.c_double)]
        L.orc_diag_lanczos.argtypes = [C.POINTER(_Model), dp, dp, dp, C.POINTER(_Basis), dp,
                                       C.c_int, C.POINTER(C.c_double), dp, C.POINTER(C.c_int),
                                       C.POINTER(C.c_double)]
        L.orc_diag_lanczos.restype = C.c_int
        L.orc_start_vector.argtypes = [C.c_uint64, C.c_long, dp]
        L.orc_apply_c.argtypes = [C.POINTER(_Basis), dp, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(_Basis), dp]
        L.orc_continued_fraction.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp, C.c_int,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_green_diag.argtypes = [C.POINTER(_Model), dp, dp, dp, C.POINTER(_Basis), dp,
                                     C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp,
                                     dp, dp, dp, np.ctypeslib.ndpointer(np.int32)]
        L.orc_observables.argtypes = [C.POINTER(_Model), C.POINTER(_Basis), dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_icom.argtypes = [C.c_int, C.c_int]
        L.orc_icom.restype = C.c_long
        _lib = L
    return _lib


def set_threads(n):
    return lib().orc_set_threads(int(n))


class Model:
    """Module globals of impurity_hamiltonian.f90:4-6 / dmft_params.f90:78-98.

    Up < 0 -> U - 2J and Jp < 0 -> J (dmft_params.f90:90-96)."""

    def __init__(self, norb=1, nbath=5, U=1.0, J=0.3, Up=-1.0, Jp=-1.0, mu=0.5):
        self.norb, self.nbath = int(norb), int(nbath)
        self.nsite = self.norb + self.nbath
        self.U, self.J, self.mu = float(U), float(J), float(mu)
        self.Up = float(Up) if Up >= 0 else self.U - 2 * self.J
        self.Jp = float(Jp) if Jp >= 0 else self.J

    def c(self):
        return _Model(self.norb, self.nbath, self.nsite, self.U, self.Up, self.J, self.Jp, self.mu)


class Basis:
    """ed_basis.f90:42-114 generate_basis (single rank)."""

    def __init__(self, nsite, ne_up, ne_dn):
        self._b = _Basis()
        rc = lib().orc_generate_basis(nsite, ne_up, ne_dn, C.byref(self._b))
        if rc != 0:
            raise ValueError(f"bad sector ({ne_up},{ne_dn}) for nsite={nsite}")
        b = self._b
        self.nsite, self.ne_up, self.ne_dn = nsite, ne_up, ne_dn
        self.nup, self.ndn, self.ntot = b.nup, b.ndn, b.ntot
        self.up = np.ctypeslib.as_array(b.up, (b.nup,)).astype(np.int64)
        self.dn = np.ctypeslib.as_array(b.dn, (b.ndn,)).astype(np.int64)
        self.idx_up = np.ctypeslib.as_array(b.idx_up, (1 << nsite,)).astype(np.int64)
        self.idx_dn = np.ctypeslib.as_array(b.idx_dn, (1 << nsite,)).astype(np.int64)

    def states(self):
        """Fock state of every row, up index fastest (ed_basis.f90:122-136)."""
        return (self.up[None, :] + (self.dn[:, None] << self.nsite)).reshape(-1)

    def ptr(self):
        return C.byref(self._b)

    def __del__(self):
        try:
            lib().orc_free_basis(C.byref(self._b))
        except Exception:
            pass


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))


def hx(model, em, ek, vmk, basis, x, y=None, py=False):
    """impurity_hamiltonian.f90:77-247 hx."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(basis.ntot) if y is None else np.array(y, dtype=np.float64, copy=True)
    m = model.c()
    lib().orc_hx(C.byref(m), _f(em), _f(ek), _f(vmk), basis.ptr(), x, y, int(bool(py)))
    return y


def lanczos_iteration(model, em, ek, vmk, basis, v_init, maxnstep):
    """lanczos_solver.f90:92-173 lanczos_iteration -> (nstep, a, b)."""
    a = np.zeros(maxnstep)
    b = np.zeros(maxnstep)
    n = C.c_int(0)
    m = model.c()
    lib().orc_lanczos_iteration(C.byref(m), _f(em), _f(ek), _f(vmk), basis.ptr(),
                                np.ascontiguousarray(v_init, dtype=np.float64), maxnstep,
                                C.byref(n), a, b)
    return n.value, a, b


def tridiag_eig(d, e):
    """dstevx('V','A') stand-in (lanczos_solver.f90:175-211): ascending w, z[:, j]."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    n = d.size
    e = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.float64)[: n - 1], [0.0]]))
    w = np.zeros(n)
    z = np.zeros(n * n)
    info = lib().orc_tridiag_eig(n, d, e, w, z)
    if info:
        raise RuntimeError(f"tridiagonal QL failed at {info}")
    return w, z.reshape(n, n).T.copy()


def start_vector(seed, n):
    """Counter-based U[0,1) start vector (stands in for lanczos_solver.f90:42-46)."""
    v = np.zeros(int(n))
    lib().orc_start_vector(C.c_uint64(seed), int(n), v)
    return v


def diag_lanczos(model, em, ek, vmk, basis, v_init, maxnstep):
    """lanczos_solver.f90:20-90 diag_lanczos -> (e0, eigvec, nstep, residual)."""
    vec = np.zeros(basis.ntot)
    e0 = C.c_double(0)
    res = C.c_double(0)
    n = C.c_int(0)
    m = model.c()
    rc = lib().orc_diag_lanczos(C.byref(m), _f(em), _f(ek), _f(vmk), basis.ptr(),
                                np.ascontiguousarray(v_init, dtype=np.float64), maxnstep,
                                C.byref(e0), vec, C.byref(n), C.byref(res))
    if rc:
        raise RuntimeError("tridiagonal eigensolve failed")
    return e0.value, vec, n.value, res.value


class EigPair:
    """eigpair_t (eigpair.f90:5-9)."""

    def __init__(self, isector, val, vec, basis):
        self.isector, self.val, self.vec, self.basis = isector, val, vec, basis


def ground_state(model, em, ek, vmk, sectors, maxnstep, seed=1):
    """ground_state.f90:1-81: sweep sectors [(ne_up, ne_dn), ...], keep the
    strict minimum (:62).  isector is 1-based.  Sector s uses start vector
    seed + s - 1."""
    gs = EigPair(0, np.inf, None, None)
    for isec, (nu, nd) in enumerate(sectors, start=1):
        b = Basis(model.nsite, nu, nd)
        e0, vec, _, _ = diag_lanczos(model, em, ek, vmk, b, start_vector(seed + isec - 1, b.ntot),
                                     maxnstep)
        if e0 < gs.val:
            gs = EigPair(isec, e0, vec, b)
    return gs


def apply_c(basis_in, vec, pm, iorb, ispin, basis_out):
    """ed_operator.f90:13-57 apply_c (pm=1 create, pm=2 destroy)."""
    out = np.zeros(basis_out.ntot)
    lib().orc_apply_c(basis_in.ptr(), np.ascontiguousarray(vec, dtype=np.float64), pm, iorb,
                      ispin, basis_out.ptr(), out)
    return out


def continued_fraction(z, n, a, b, minus=False):
    """numeric_utils.f90:11-44 continued_fraction_p / _m."""
    fr, fi = C.c_double(0), C.c_double(0)
    lib().orc_continued_fraction(float(np.real(z)), float(np.imag(z)), int(n),
                                 np.ascontiguousarray(a, dtype=np.float64),
                                 np.ascontiguousarray(b, dtype=np.float64), int(bool(minus)),
                                 C.byref(fr), C.byref(fi))
    return complex(fr.value, fi.value)


def green_diag(model, em, ek, vmk, gs, iorb, ispin, maxnstep, omega):
    """dmft_green.f90:177-260 green_diag -> (g(nw) complex, ap, bp, am, bm, nsteps)."""
    nw = len(omega)
    g = np.zeros(2 * nw)
    ap, bp, am, bm = (np.zeros(maxnstep) for _ in range(4))
    ns = np.zeros(2, dtype=np.int32)
    m = model.c()
    lib().orc_green_diag(C.byref(m), _f(em), _f(ek), _f(vmk), gs.basis.ptr(),
                         np.ascontiguousarray(gs.vec, dtype=np.float64), gs.val, iorb, ispin,
                         maxnstep, nw, np.ascontiguousarray(omega, dtype=np.float64), g, ap, bp,
                         am, bm, ns)
    return g[0::2] + 1j * g[1::2], ap, bp, am, bm, (int(ns[0]), int(ns[1]))


def cluster_green_ftn(model, em, ek, vmk, gs, maxnstep, omega, nspin=1, both_spins=False):
    """dmft_green.f90:47-126 cluster_green_ftn -> (G_cl(nw,norb,nspin), coeffs).

    coeffs[(iorb, ispin)] = (ap, bp, am, bm) as written to g_coeffs.dat
    (:89-109).  both_spins=True also computes spin 2 when the reference's
    `even` shortcut (:76-80) would skip it."""
    nw = len(omega)
    G = np.zeros((nw, model.norb, nspin), dtype=np.complex128)
    coeffs = {}
    ne_u, ne_d = gs.basis.ne_up, gs.basis.ne_dn
    even = not (nspin == 2 or ne_u != ne_d)
    for iorb in range(1, model.norb + 1):
        g1, *c1, _ = green_diag(model, em, ek, vmk, gs, iorb, 1, maxnstep, omega)
        coeffs[(iorb, 1)] = tuple(c1)
        g2, c2 = g1, tuple(c1)
        if not even or both_spins:
            g2, *c2l, _ = green_diag(model, em, ek, vmk, gs, iorb, 2, maxnstep, omega)
            c2 = tuple(c2l)
        coeffs[(iorb, 2)] = c2
        if nspin == 1 and even:
            G[:, iorb - 1, 0] = g1
        elif nspin == 1:
            G[:, iorb - 1, 0] = 0.5 * (g1 + g2)
        else:
            G[:, iorb - 1, 0] = g1
            G[:, iorb - 1, 1] = g2
    return G, coeffs


def observables(model, basis, vec):
    """post_process.f90:63-112 -> occ(norb,2), docc(norb)."""
    occ = np.zeros(2 * model.norb)
    docc = np.zeros(model.norb)
    m = model.c()
    lib().orc_observables(C.byref(m), basis.ptr(), np.ascontiguousarray(vec, dtype=np.float64),
                          occ, docc)
    return occ.reshape(model.norb, 2, order="F"), docc


def matsubara(nw, beta):
    """dmft_grid.f90:42-53 on one rank: omega_n = (2n+1) pi / beta."""
    return (2.0 * np.arange(nw) + 1.0) * np.pi / beta


def icom(n, m):
    return lib().orc_icom(n, m)
