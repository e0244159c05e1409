"""The request loop (edgpu.serve.serve over the batched engine): requests
arriving on a Poisson clock are admitted while others run and share the
batched launches; every result must equal the same problem solved alone.

Tolerance: bit-identical E0 and G(i w) (a chain's reductions depend only on
its own tiles, so the mix cannot change a result).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graphs", [True, False])
def test_served_results_equal_alone(edgpu, graphs):
    from edgpu import serve as S
    from edgpu import workloads as W
    om = W.matsubara(128, 100.0)
    req = W.served_requests(edgpu, rate=60.0, duration=0.5, seed=5,
                            ns_weights=((6, 0.4), (8, 0.35), (10, 0.25)), maxnstep=60,
                            maxnstep_green=40)
    assert len(req) >= 15
    print(f"serve {len(req)} requests, graphs={graphs}", flush=True)
    got = {}
    r = S.serve(edgpu, req, om, graphs=graphs, results=got, drain_limit=60.0)
    assert r["completed"] == len(req) == len(got)
    assert r["unfinished"] == 0
    assert r["depth"][-1][3] == 0
    # the stream overlapped: some request was admitted while others were in flight
    assert max(d[3] for d in r["depth"]) >= 1 or r["stats"]["chain_steps"] > r["stats"]["steps"]
    for k, (_, p) in enumerate(req):
        alone, _, _ = S.run_batched(edgpu, [p], om, graphs=graphs)
        assert got[k]["e0"] == alone[0]["e0"], k
        assert np.array_equal(got[k]["G"], alone[0]["G"]), k
