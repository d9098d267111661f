"""Serving loop: a small Poisson stream of 2D and 3D requests (FP64 and
mixed) is admitted under an HBM budget with continuous batching; every
request completes at its end time and its result equals the same request run
alone."""
import numpy as np
import pytest

from alpaca_amd import gpu, serving
from alpaca_amd.ensemble import CASE_DONE, Ensemble

pytestmark = pytest.mark.gpu


def _alone(req):
    ens = Ensemble(req.dim, max_blocks=serving.req_blocks(req), max_cases=1,
                   classes=[(gpu.FLUX_HLLC, req.precision)])
    cid = ens.add_case(serving.case_desc(req), serving.initial_state(req), req.t_end,
                       serving.req_cfl(req))
    ens.run_until_done([cid], poll=4)
    out = ens.download(cid)
    ens.close()
    return out


def test_serving_small_stream_matches_alone():
    import torch
    assert torch.cuda.is_available()
    print("[serving small]", flush=True)
    reqs = serving.generate_requests(rate=40.0, window=0.5, seed=3, sizes_2d=(2, 4, 6),
                                     sizes_3d=(2, 3))
    for r in reqs:
        r.t_end = r.t_end * 0.2  # keep the test short
    # A budget that forces some requests to wait for others to finish.
    server = serving.Server(budget_bytes=0.2e9, share_3d=0.7, max_cases=64, steps_per_poll=2)
    stats, done, results = server.serve(reqs, keep_results=True)
    server.close()
    print(f"  {stats}")
    assert stats["completed"] == len(reqs)
    assert all(r["flags"] & CASE_DONE for r in done)
    assert all(abs(r["time"] - next(q.t_end for q in reqs if q.rid == r["rid"])) == 0.0
               for r in done)
    assert stats["hbm_resident_peak_bytes"] <= 0.2e9
    for req in reqs[:: max(1, len(reqs) // 6)]:
        assert np.array_equal(results[req.rid], _alone(req)), f"request {req.rid}"
