"""edg_hx_noalloc (the H.v that allocates nothing) through the C ABI against
edg_hx, the entry point it is a variant of: the same kernels on the same
tables, so the results are bit-identical (and both match the oracle: FP64
1e-13, FP32 2e-6 of max|y|).  Two calls in a row on the device's one
scratch array, the second with a smaller sector, then an accumulate call,
and a sector whose tables exceed the scratch array (refused with a message).
Device buffers of the test itself come from hipMalloc through ctypes.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 16))) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 16))) == 0
        self.bufs.append(p)
        return p

    def download(self, p, n, dtype):
        out = np.zeros(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def fortran(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))


def params(edgpu, norb, nbath, U, J, mu, precision):
    p = edgpu._Params()
    edgpu.load_library().edg_default_params(C.byref(p))
    p.norb, p.nbath, p.U, p.J, p.mu = norb, nbath, U, J, mu
    p.precision = {"fp64": edgpu.EDG_FP64, "fp32": edgpu.EDG_FP32}[precision]
    return p


def run_noalloc(edgpu, hip, p, em, ek, vmk, nu, nd, x, y0=None):
    L = edgpu.load_library()
    dt = np.float32 if p.precision == edgpu.EDG_FP32 else np.float64
    dem, dek, dvm = hip.upload(fortran(em)), hip.upload(fortran(ek)), hip.upload(fortran(vmk))
    dx = hip.upload(x.astype(dt))
    dy = hip.upload(y0.astype(dt)) if y0 is not None else hip.empty(x.size * np.dtype(dt).itemsize)
    rc = L.edg_hx_noalloc(C.byref(p), dem, dek, dvm, nu, nd, dx, dy, int(y0 is not None), None)
    assert rc == 0, L.edg_last_error(None).decode()
    hip.sync()
    return hip.download(dy, x.size, dt).astype(np.float64)


def case(rng, norb, nbath):
    em = rng.uniform(-1, 1, (norb, 2))
    ek = rng.uniform(-1, 1, (nbath, 2))
    vmk = rng.uniform(0.2, 0.6, (norb, nbath, 2)) * rng.choice([-1.0, 1.0], (norb, nbath, 2))
    return em, ek, vmk


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("fp64", 1e-13)])
def test_noalloc_matches_hx(edgpu, oracle, precision, tol):
    hip = Hip()
    try:
        # first call: Ns=14 single orbital, sector (7,7); second, smaller:
        # two-orbital Kanamori Ns=8, sector (4,3); then accumulate on Ns=10
        for k, (norb, nbath, nu, nd, U, J, mu) in enumerate([(1, 13, 7, 7, 2.0, 0.3, 1.0),
                                                               (2, 6, 4, 3, 3.0, 0.75, 2.0),
                                                               (1, 9, 5, 4, 2.5, 0.3, 1.2)]):
            print(f"noalloc {precision} call {k}: norb={norb} nbath={nbath} sector=({nu},{nd})", flush=True)
            rng = np.random.default_rng(1000 + k)
            em, ek, vmk = case(rng, norb, nbath)
            s = edgpu.Solver(norb=norb, nbath=nbath, U=U, J=J, mu=mu, maxnstep=20, precision=precision)
            D = s.dim(nu, nd)
            x = rng.normal(size=D)
            if precision == "fp32":
                x = x.astype(np.float32).astype(np.float64)
            acc = k == 2
            y0 = rng.normal(size=D) if acc else None
            if acc and precision == "fp32":
                y0 = y0.astype(np.float32).astype(np.float64)
            p = params(edgpu, norb, nbath, U, J, mu, precision)
            got = run_noalloc(edgpu, hip, p, em, ek, vmk, nu, nd, x, y0)
            want = s.hx(em, ek, vmk, nu, nd, x, y0, py=acc)
            s.close()
            assert np.array_equal(got, want), k
            m = oracle.Model(norb, nbath, U=U, J=J, mu=mu)
            ref = oracle.hx(m, em, ek, vmk, oracle.Basis(norb + nbath, nu, nd), x, y0, py=acc)
            assert np.abs(got - ref).max() <= tol * max(1.0, np.abs(ref).max()), k
    finally:
        hip.close()


def test_noalloc_refuses_tables_beyond_scratch(edgpu):
    """Five orbitals and ten bath sites, sector (7,7): the tables need about
    17 MB, more than the 16 MiB scratch array."""
    hip = Hip()
    try:
        L = edgpu.load_library()
        p = params(edgpu, 5, 10, 2.0, 0.3, 1.0, "fp32")
        d = hip.empty(1 << 12)
        rc = L.edg_hx_noalloc(C.byref(p), d, d, d, 7, 7, d, hip.empty(1 << 12), 0, None)
        assert rc == 3  # EDG_ERR_ARG
        assert b"scratch holds" in L.edg_last_error(None)
    finally:
        hip.close()
