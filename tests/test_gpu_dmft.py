"""The DMFT self-consistency loop (edgpu.dmft.run_dmft, dmft.f90:60-104)
around the GPU solver against the same loop around the FP64 oracle, config 1
(Ns=6, U=2, half filling, Bethe lattice): iteration by iteration the
convergence measure, and at the end the fitted bath and G_loc.

Tolerances: the solvers agree to ~1e-12 per solve; the bath fit's line
searches stop at 1e-8 relative, so parameters and G_loc are compared at 1e-6
and the convergence history at 1e-6 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dmft_loop_gpu_matches_oracle(edgpu, oracle):
    from edgpu import dmft
    nw = 512
    omega = (2 * np.arange(nw) + 1) * np.pi / 100.0
    rng = np.random.Generator(np.random.PCG64(1))
    e = rng.uniform(-1, 1, 5)
    v = rng.uniform(0.2, 0.6, 5)
    em = np.zeros((1, 2))
    ek = np.stack([e, e], axis=1)
    vmk = np.stack([v, v], axis=1)[None, :, :]
    U = 2.0
    m = oracle.Model(1, 5, U=U, mu=U / 2)

    def solve_oracle(em, ek, vmk):
        gs = oracle.ground_state(m, em, ek, vmk, [(3, 3)], 80, seed=1)
        G, _ = oracle.cluster_green_ftn(m, em, ek, vmk, gs, 80, omega)
        return G

    s = edgpu.Solver(norb=1, nbath=5, U=U, mu=U / 2, maxnstep=80, precision="fp64", seed=1)

    def solve_gpu(em, ek, vmk):
        s.ground_state(em, ek, vmk, [(3, 3)])
        G, _ = s.cluster_green_ftn(em, ek, vmk, omega)
        return G

    ref = dmft.run_dmft(solve_oracle, em, ek, vmk, omega, nloop=6, scf_tol=1e-6)
    got = dmft.run_dmft(solve_gpu, em, ek, vmk, omega, nloop=6, scf_tol=1e-6)
    s.close()
    print("history", ref["history"], got["history"], flush=True)
    assert got["iterations"] == ref["iterations"] and got["converged"]
    assert np.allclose(got["history"], ref["history"], rtol=1e-6, atol=1e-14)
    for k in ("em", "ek", "vmk"):
        assert np.abs(got[k] - ref[k]).max() < 1e-6, k
    assert np.abs(got["G_loc"] - ref["G_loc"]).max() < 1e-6
