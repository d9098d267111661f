"""Serving loop: Poisson streams of 2D and 3D requests (FP64 and mixed)
admitted under one HBM budget with continuous batching.

* a small stream under a tight budget: every request completes at its end
  time and equals the same request run alone;
* the serving mix at its stated sizes (70% 2D 64^2..512^2, 30% 3D blob+shock
  64^3..128^3, FP64/mixed 50/50, end times scaled down): sampled FP64 results
  equal the CPU oracle bit for bit, every mixed result is within L1 density
  error 1e-6 of its FP64 twin, and the 3D (periodic) mixed results conserve
  mass and energy to 1e-13.
"""
import dataclasses

import numpy as np
import pytest

from alpaca_amd import cases, gpu, serving
from alpaca_amd.ensemble import CASE_DONE, Ensemble
import oracle_py

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
    server = serving.Server(budget_bytes=0.2e9, max_cases=64, steps_per_poll=2)
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


def _serve(reqs, budget):
    server = serving.Server(budget_bytes=budget, max_cases=256, steps_per_poll=4)
    stats, done, results = server.serve(reqs, keep_results=True)
    server.close()
    return stats, done, results


def test_serving_mix_at_stated_sizes():
    import torch
    assert torch.cuda.is_available()
    reqs = serving.generate_requests(rate=120.0, window=0.25, seed=21)
    for r in reqs:
        r.t_end = r.t_end * 0.1  # t_end ~ U[0.005, 0.025]
    sizes = sorted({(r.dim, r.nb) for r in reqs})
    print(f"[serving mix] {len(reqs)} requests, sizes {sizes}", flush=True)
    assert {d for d, _ in sizes} == {2, 3}
    stats, done, results = _serve(reqs, 20e9)
    print(f"  {stats}", flush=True)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert all(r["flags"] & CASE_DONE for r in done)
    by_rid = {r.rid: r for r in reqs}
    # FP64 twins of the mixed requests, served as a second stream.
    twins = [dataclasses.replace(r, precision=gpu.FP64) for r in reqs if r.precision == gpu.MIXED]
    _, _, twin_results = _serve(twins, 20e9)
    worst = 0.0
    for t in twins:
        m, f = results[t.rid], twin_results[t.rid]
        l1 = float(np.abs(m[0] - f[0]).sum() / np.abs(f[0]).sum())
        worst = max(worst, l1)
        assert l1 <= 1e-6, f"request {t.rid} ({t.dim}D nb {t.nb}): L1 rho {l1:.3e}"
        if t.dim == 3:
            s0 = np.asarray(serving.initial_state(by_rid[t.rid]).cpu().numpy())
            for e in (0, 1):  # mass, energy: periodic sides conserve them
                drift = abs(m[e].sum() - s0[e].sum()) / abs(s0[e]).sum()
                assert drift <= 1e-13, f"request {t.rid} equation {e}: drift {drift:.3e}"
    print(f"  mixed vs FP64 twins: worst L1 rho {worst:.3e} over {len(twins)} requests",
          flush=True)
    # FP64 requests against the oracle: the smallest 2D and 3D ones and one
    # 2D request of at least 256^2 cells.
    fp64 = [r for r in reqs if r.precision == gpu.FP64]
    picks = []
    for d in (2, 3):
        cand = sorted((r for r in fp64 if r.dim == d), key=lambda r: (r.nb, r.t_end))
        if cand:
            picks.append(cand[0])
    big2 = [r for r in fp64 if r.dim == 2 and r.nb >= 16]
    if big2:
        picks.append(min(big2, key=lambda r: r.t_end))
    assert picks
    for r in picks:
        print(f"  oracle request {r.rid}: {r.dim}D nb {r.nb} t_end {r.t_end:.4f}", flush=True)
        s0 = serving.initial_state(r)
        s0 = s0.cpu().numpy() if hasattr(s0, "cpu") else s0
        nb = [r.nb] * r.dim
        bc = [0, 0] if r.dim == 2 else [1, 1, 1]
        want, _, tf = oracle_py.run(r.dim, nb, bc, 0, serving.req_cfl(r), r.t_end, 100000,
                                    np.ascontiguousarray(s0), gamma=r.gamma, threads=16)
        assert tf == r.t_end
        got = results[r.rid]
        assert np.array_equal(got, want), f"request {r.rid}: max diff {np.abs(got - want).max()}"
