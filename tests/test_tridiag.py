"""The product's tridiagonal eigensolver (dmft-ed-zt_amd/csrc/tridiag.cpp,
the stand-in for MKL dstevx('V','A') of lanczos_solver.f90:175-211) built
with g++ on the CPU and checked against LAPACK (scipy.linalg.eigh_tridiagonal).

Cases: random matrices of 1..200 levels; matrices split into blocks by tiny
off-diagonals whose lowest eigenvalues coincide; and the tridiagonals of real
200-step plain Lanczos runs of the oracle (Ns = 8 and 10), whose loss of
orthogonality makes near-degenerate copies ("ghosts") of the lowest Ritz
value.  Tolerances: lowest eigenvalue within 1e-13 * max(1, |T|) of LAPACK's;
residual |T z - lambda z| below 1e-11 * max(1, |T|); z inside the LAPACK
eigenspace of the eigenvalues within 1e-8 of lambda (projection norm
1 - 1e-10), and z the LAPACK eigenvector up to sign when that eigenvalue
is isolated.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "dmft-ed-zt_amd", "csrc", "tridiag.cpp")

SHIM = r"""
namespace edg { int tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z); }
extern "C" int t_lowest(int n, const double *d, const double *e, double *lambda, double *z)
{ return edg::tridiag_lowest(n, d, e, lambda, z); }
"""


@pytest.fixture(scope="module")
def lowest(tmp_path_factory):
    d = tmp_path_factory.mktemp("tridiag")
    shim = d / "shim.cpp"
    shim.write_text(SHIM)
    lib = d / "libtridiag.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", SRC, str(shim), "-o", str(lib)])
    L = C.CDLL(str(lib))
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.t_lowest.argtypes = [C.c_int, dp, dp, C.POINTER(C.c_double), dp]
    L.t_lowest.restype = C.c_int

    def f(dg, e):
        dg = np.ascontiguousarray(dg, dtype=np.float64)
        e = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.float64), [0.0]]))
        z = np.zeros(dg.size)
        lam = C.c_double(0)
        rc = L.t_lowest(dg.size, dg, e, C.byref(lam), z)
        assert rc == 0
        return lam.value, z
    return f


def check(lowest, dg, e):
    lam, z = lowest(dg, e)
    w, V = sl.eigh_tridiagonal(dg, e)
    n = dg.size
    T = np.diag(dg) + (np.diag(e, 1) + np.diag(e, -1) if n > 1 else 0)
    scale = max(1.0, np.abs(T).max())
    assert abs(lam - w[0]) <= 1e-13 * scale
    assert np.linalg.norm(T @ z - lam * z) <= 1e-11 * scale
    assert abs(np.linalg.norm(z) - 1.0) < 1e-12
    cl = np.abs(w - lam) < 1e-8 * scale
    proj = V[:, cl].T @ z
    assert np.linalg.norm(proj) > 1 - 1e-10
    if cl.sum() == 1 and (n == 1 or w[1] - w[0] > 1e-6 * scale):
        assert min(np.abs(z - V[:, 0]).max(), np.abs(z + V[:, 0]).max()) < 1e-9
    return lam, z, int(cl.sum())


@pytest.mark.parametrize("n", [1, 2, 3, 7, 60, 200])
def test_random_tridiagonals(lowest, n):
    rng = np.random.default_rng(n)
    for _ in range(5):
        check(lowest, rng.normal(size=n), rng.normal(size=n - 1))


def test_largest_component_positive(lowest):
    rng = np.random.default_rng(4)
    dg, e = rng.normal(size=30), rng.normal(size=29)
    _, z = lowest(dg, e)
    assert z[np.argmax(np.abs(z))] > 0


def test_coinciding_blocks(lowest):
    """Two identical blocks coupled by off-diagonals of 1e-14 .. 0: the
    lowest eigenvalue is (numerically) double."""
    rng = np.random.default_rng(9)
    d1, e1 = rng.normal(size=12), rng.normal(size=11)
    for eps in (0.0, 1e-14, 1e-10):
        dg = np.concatenate([d1, d1])
        e = np.concatenate([e1, [eps], e1])
        _, _, ncl = check(lowest, dg, e)
        assert ncl == 2


@pytest.mark.parametrize("nbath,sector", [(7, (4, 4)), (9, (5, 5))])
def test_lanczos_ghost_clusters(lowest, oracle, nbath, sector):
    """Tridiagonals of 200 plain Lanczos steps (no reorthogonalization) of
    the oracle: the lowest Ritz value repeats (ghosts), and the solver's
    eigenvalue is still LAPACK's lowest, with z in the cluster's span."""
    rng = np.random.default_rng(nbath)
    em = rng.uniform(-1, 1, (1, 2))
    ek = rng.uniform(-1, 1, (nbath, 2))
    vmk = rng.uniform(0.2, 0.6, (1, nbath, 2))
    m = oracle.Model(1, nbath, U=2.0, mu=1.0)
    b = oracle.Basis(nbath + 1, *sector)
    n, a, bb = oracle.lanczos_iteration(m, em, ek, vmk, b, oracle.start_vector(3, b.ntot), 200)
    assert n == 200
    dg, e = a[:n], bb[1:n]
    w = sl.eigh_tridiagonal(dg, e, eigvals_only=True)
    assert np.sum(np.abs(w - w[0]) < 1e-8) >= 2, "expected a ghost copy of the lowest Ritz value"
    _, _, ncl = check(lowest, dg, e)
    assert ncl >= 2
