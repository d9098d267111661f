"""Continuous-batching serving loop over device-resident ensembles.

Requests (independent simulations) arrive as a Poisson stream.  Each is
admitted as soon as its blocks fit the HBM budget (a 2D and a 3D block pool,
FIFO with backfill), stays resident until it reaches its end time, then its
state is read back and its blocks are released.  Between polls both pools
advance by a few macro steps, each on its own stream so their kernels overlap
on the device; every running case takes its own CFL step, so a request's
result does not depend on what else is running.

Reported: p50/p99 time-to-solution (arrival to result), sustained
cell-updates/s, peak and mean HBM occupancy of the admitted state.
"""
from __future__ import annotations

import dataclasses
import json
import math
import time
from collections import deque

import numpy as np

from . import cases, gpu
from .ensemble import CASE_STOP, Ensemble


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float
    dim: int
    nb: int
    t_end: float
    precision: int
    seed: int
    gamma: float = 1.4


def generate_requests(rate: float, window: float, seed: int = 0, sizes_2d=(4, 8, 16, 32),
                      sizes_3d=(4, 5, 6, 7, 8)):
    """Poisson arrivals over [0, window): 70% 2D four-quadrant (64^2..512^2
    cells), 30% 3D blob+shock (64^3..128^3), t_end ~ U[0.05, 0.25], FP64 or
    mixed with equal probability."""
    rng = cases.SplitMix64(0x5E7E, seed)
    reqs = []
    t = 0.0
    rid = 0
    while True:
        t += -math.log(1.0 - rng.uniform(0.0, 1.0)) / rate
        if t >= window:
            break
        is2d = rng.uniform(0.0, 1.0) < 0.7
        if is2d:
            nb = sizes_2d[min(int(rng.uniform(0.0, len(sizes_2d))), len(sizes_2d) - 1)]
        else:
            nb = sizes_3d[min(int(rng.uniform(0.0, len(sizes_3d))), len(sizes_3d) - 1)]
        t_end = rng.uniform(0.05, 0.25)
        prec = gpu.FP64 if rng.uniform(0.0, 1.0) < 0.5 else gpu.MIXED
        reqs.append(Request(rid, t, 2 if is2d else 3, nb, t_end, prec, seed=rid))
        rid += 1
    return reqs


def initial_state(req: Request):
    if req.dim == 2:
        return cases.four_quadrant_2d(req.nb, req.nb, seed=1000 + req.seed, gamma=req.gamma)
    return cases.blobs_shock_3d_torch(req.nb, seed=1000 + req.seed, gamma=req.gamma)


def case_desc(req: Request):
    if req.dim == 2:
        return gpu.make_desc(2, [req.nb, req.nb], [0, 0], flux=gpu.FLUX_HLLC,
                             precision=req.precision, gamma=req.gamma)
    return gpu.make_desc(3, [req.nb] * 3, [1, 1, 1], flux=gpu.FLUX_HLLC, precision=req.precision,
                         gamma=req.gamma)


def req_blocks(req: Request) -> int:
    return req.nb ** req.dim


def req_cfl(req: Request) -> float:
    return 0.6 if req.dim == 2 else 0.5


class Server:
    """Two resident pools (2D, 3D) sized from an HBM budget."""

    def __init__(self, budget_bytes: float, share_3d: float = 0.8, max_cases: int = 4096,
                 steps_per_poll: int = 4, tuning=None):
        import torch

        self.torch = torch
        classes = [(gpu.FLUX_HLLC, gpu.FP64), (gpu.FLUX_HLLC, gpu.MIXED)]
        self.block_bytes = {d: 3 * gpu.block_doubles(d) * 8 for d in (2, 3)}
        budget = {3: budget_bytes * share_3d, 2: budget_bytes * (1.0 - share_3d)}
        self.pools = {}
        for d in (2, 3):
            max_blocks = int(budget[d] // self.block_bytes[d])
            self.pools[d] = Ensemble(d, max_blocks=max_blocks, max_cases=max_cases,
                                     classes=classes)
            if tuning and d in tuning:
                self.pools[d].set_tuning(*tuning[d])
        self.budget = budget_bytes
        self.steps_per_poll = steps_per_poll

    def close(self):
        for p in self.pools.values():
            p.close()

    def fits(self, req: Request) -> bool:
        free_blocks, largest, slots = self.pools[req.dim].capacity()
        return slots > 0 and largest >= req_blocks(req)

    def admit(self, req: Request) -> int:
        pool = self.pools[req.dim]
        return pool.add_case(case_desc(req), initial_state(req), req.t_end, req_cfl(req))

    def serve(self, requests, keep_results=False, max_wall=None, log=None):
        torch = self.torch
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        live = {2: {}, 3: {}}  # case id -> request
        done = []
        results = {}
        resident = 0.0
        occ_samples = []
        updates = 0
        t0 = time.perf_counter()
        last_log = t0
        while pending or waiting or live[2] or live[3]:
            now = time.perf_counter() - t0
            if max_wall is not None and now > max_wall:
                break
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # FIFO admission with backfill: a request that does not fit yet
            # does not block smaller ones behind it.
            still = deque()
            while waiting:
                req = waiting.popleft()
                if self.fits(req):
                    cid = self.admit(req)
                    live[req.dim][cid] = (req, time.perf_counter() - t0)
                    resident += req_blocks(req) * self.block_bytes[req.dim]
                else:
                    still.append(req)
            waiting = still
            if not (live[2] or live[3]):
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            occ_samples.append(resident)
            # Both pools advance concurrently on their own streams.
            for d in (2, 3):
                if live[d]:
                    self.pools[d].step(self.steps_per_poll)
            for d in (2, 3):
                if not live[d]:
                    continue
                st = self.pools[d].states()
                tnow = time.perf_counter() - t0
                for cid in list(live[d]):
                    rec = st[cid]
                    if rec.flags & CASE_STOP:
                        req, t_admit = live[d].pop(cid)
                        cells = req_blocks(req) * (16 ** req.dim)
                        updates += cells * rec.steps
                        if keep_results:
                            results[req.rid] = self.pools[d].download(cid)
                        self.pools[d].remove_case(cid)
                        resident -= req_blocks(req) * self.block_bytes[d]
                        done.append({"rid": req.rid, "dim": d, "nb": req.nb,
                                     "precision": req.precision, "arrival": req.arrival,
                                     "admitted": t_admit, "finished": tnow,
                                     "latency": tnow - req.arrival, "steps": rec.steps,
                                     "flags": rec.flags, "time": rec.time})
            if log and time.perf_counter() - last_log > 5.0:
                last_log = time.perf_counter()
                log(f"[serving] t={time.perf_counter() - t0:.1f}s done={len(done)} "
                    f"live={len(live[2]) + len(live[3])} waiting={len(waiting)} "
                    f"pending={len(pending)}")
        wall = time.perf_counter() - t0
        torch.cuda.synchronize()
        lat = np.array([r["latency"] for r in done]) if done else np.array([0.0])
        stats = {
            "requests": len(requests),
            "completed": len(done),
            "wall_s": wall,
            "cell_updates": updates,
            "cell_updates_per_s": updates / wall if wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()),
            "hbm_budget_bytes": self.budget,
            "hbm_resident_peak_bytes": max(occ_samples) if occ_samples else 0.0,
            "hbm_resident_mean_bytes": float(np.mean(occ_samples)) if occ_samples else 0.0,
            "nonfinite": sum(1 for r in done if r["flags"] & 1),
        }
        return stats, done, results


def main_bench(args):
    """bench.py --workload serving: one JSON line."""
    from . import autotune

    rate = getattr(args, "rate", 200.0)
    window = getattr(args, "window", 2.0)
    reqs = generate_requests(rate, window, seed=0)
    tuning = {d: autotune.lookup(d, gpu.FLUX_HLLC, gpu.FP64) for d in (2, 3)}
    server = Server(200e9, tuning=tuning)
    stats, done, _ = server.serve(reqs, log=lambda m: print(m, flush=True))
    server.close()
    line = {
        "metric": ("cell-updates/s WENO5-HLLC-RK3, FP64 vs mixed at stated L1 err; "
                   "p50/p99 case latency"),
        "value": stats["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64|f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"Poisson request stream {rate:g} req/s for {window:g} s: 70% 2D "
                                "four-quadrant 64^2-512^2, 30% 3D blob+shock 64^3-128^3, t_end "
                                "U[0.05,0.25], FP64/mixed 50/50, HBM budget 200 GB, continuous "
                                "batching")},
        "serving": stats,
    }
    print(json.dumps(line), flush=True)
