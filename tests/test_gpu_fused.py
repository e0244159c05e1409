"""The fused Lanczos step (one H.v pass per step, edg_set_fused_steps; off by
default) against the two-pass step and the oracle: same recurrence,
different rounding.  The cases include chains that start low in the
spectrum (alpha/beta large), where the unshifted form of the fused step
drifted by 5e-8 in G even in FP64.

Tolerances: FP64 vectors, E0 within 1e-10 and G within 1e-9 of the two-pass
solve and of the oracle, the first 25 Lanczos coefficients within 1e-10 of
the oracle's; FP32 vectors, E0 within 1e-6 and G within 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,etol,gtol", [("fp64", 1e-10, 1e-9), ("fp32", 1e-6, 1e-5)])
@pytest.mark.parametrize("nbath,sector", [(9, (5, 5)), (11, (6, 5)), (13, (7, 7))])
def test_fused_vs_two_pass_and_oracle(edgpu, oracle, precision, etol, gtol, nbath, sector):
    from edgpu import workloads as W
    print(f"fused {precision} nbath={nbath} sector={sector}", flush=True)
    em, ek, vmk, mu = W.single_orbital(nbath, 2.5, seed=nbath)
    om = W.matsubara(512, 100.0)
    res = {}
    for fused in (True, False):
        s = edgpu.Solver(norb=1, nbath=nbath, U=2.5, mu=mu, maxnstep=150, maxnstep_green=100,
                         precision=precision, both_spins=True, seed=3)
        s.set_fused_steps(fused)
        gs = s.ground_state(em, ek, vmk, [sector])
        G, co = s.cluster_green_ftn(em, ek, vmk, om)
        res[fused] = (gs.val, G, co)
        s.close()
    assert abs(res[True][0] - res[False][0]) < etol
    assert np.abs(res[True][1] - res[False][1]).max() < gtol
    if nbath <= 11:
        m = oracle.Model(1, nbath, U=2.5, mu=mu)
        ogs = oracle.ground_state(m, em, ek, vmk, [sector], 150, seed=3)
        G2, _ = oracle.cluster_green_ftn(m, em, ek, vmk, ogs, 100, om, both_spins=True)
        assert abs(res[True][0] - ogs.val) < etol
        assert np.abs(res[True][1] - G2).max() < gtol


def test_fused_lanczos_coefficients_vs_oracle(edgpu, oracle):
    """lanczos_iteration (lanczos_solver.f90:92-173) with fused steps, FP64:
    the leading alpha/beta against the oracle's three-term recurrence."""
    rng = np.random.default_rng(44)
    em = rng.uniform(-1, 1, (1, 2))
    ek = rng.uniform(-1, 1, (9, 2))
    vmk = rng.uniform(0.2, 0.6, (1, 9, 2)) * rng.choice([-1.0, 1.0], (1, 9, 2))
    s = edgpu.Solver(norb=1, nbath=9, U=2.0, mu=1.0, maxnstep=80)
    s.set_fused_steps(True)
    m = oracle.Model(1, 9, U=2.0, mu=1.0)
    b = oracle.Basis(10, 5, 4)
    v0 = oracle.start_vector(7, b.ntot)
    n, a, bb = s.lanczos_iteration(em, ek, vmk, 5, 4, v0)
    n2, a2, b2 = oracle.lanczos_iteration(m, em, ek, vmk, b, v0, 80)
    assert n == n2 == 80 and bb[0] == 0.0
    assert np.abs(a[:25] - a2[:25]).max() < 1e-10
    assert np.abs(bb[:25] - b2[:25]).max() < 1e-10
