"""Continuous-batching serving loop over device-resident ensembles.

Requests (independent simulations) arrive as a Poisson stream.  Arrived
requests are admitted while their blocks fit the HBM budget (one block pool
shared by a 2D and a 3D ensemble; a request larger than the whole pool is
rejected on arrival) and the live cases hold fewer cells than a cap that
keeps the device saturated; among waiting requests the smallest estimated
work goes first (a request waiting past max_wait goes first regardless).  A
case stays resident until it reaches its end time, then its state is read
back and its blocks are released.  Between polls both
ensembles advance by a few macro steps, each on its own stream so their
kernels overlap on the device; every running case takes its own CFL step, so
a request's result does not depend on what else is running.

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


def req_cells(req: Request) -> int:
    return req_blocks(req) * 16 ** req.dim


def req_work(req: Request) -> float:
    """Estimated cell-updates of a request: cells x macro steps, with the
    step count taken as t_end * N / CFL (N cells per side; the CFL step is
    dx / (max signal speed) * CFL, and the signal speeds of the two case
    families are of order one)."""
    n = req.nb * 16
    return float(req_cells(req)) * req.t_end * n / req_cfl(req)


class Admission:
    """Host side of admission, with no device state: requests move from
    `pending` (not yet arrived) to `waiting` (arrived) and are admitted with
    backfill -- a request that does not fit yet does not block others behind
    it.  A request larger than the pool can ever hold (``max_blocks[dim]``) is
    rejected on arrival, so `waiting` only holds requests that fit once enough
    cases finish.

    order "fifo" admits in arrival order; "shortest" admits the smallest
    estimated work first (req_work), which minimises the mean time to
    solution when the device is the bottleneck.  With ``live_cells_cap``
    admission stops once the live cases hold that many cells: enough to keep
    the device saturated, while the rest wait instead of all sharing the
    device and all finishing late.  A request waiting longer than
    ``max_wait`` seconds goes first regardless of its size (no starvation);
    with nothing live the first candidate is always admitted."""

    def __init__(self, requests, max_blocks, order="fifo", live_cells_cap=None,
                 max_wait=float("inf")):
        if order not in ("fifo", "shortest"):
            raise ValueError(f"unknown admission order {order!r}")
        self.pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.waiting = deque()
        self.rejected = []
        self.max_blocks = dict(max_blocks)
        self.order = order
        self.live_cells_cap = live_cells_cap
        self.max_wait = max_wait

    def arrive(self, now: float):
        while self.pending and self.pending[0].arrival <= now:
            req = self.pending.popleft()
            if req_blocks(req) > self.max_blocks.get(req.dim, 0):
                self.rejected.append(req)
            else:
                self.waiting.append(req)

    def admit(self, fits, admit, live_cells=0, now=0.0):
        """Calls admit(req) for the waiting requests that fits(req), in the
        admission order, until the live cells reach live_cells_cap; returns
        how many were admitted."""
        if self.order == "shortest":
            overdue = [r for r in self.waiting if now - r.arrival > self.max_wait]
            rest = sorted((r for r in self.waiting if now - r.arrival <= self.max_wait),
                          key=lambda r: (req_work(r), r.arrival))
            queue = deque(overdue + rest)
        else:
            queue = self.waiting
        still = deque()
        n = 0
        cells = live_cells
        while queue:
            req = queue.popleft()
            capped = (self.live_cells_cap is not None and cells > 0 and
                      cells + req_cells(req) > self.live_cells_cap)
            if not capped and fits(req):
                admit(req)
                cells += req_cells(req)
                n += 1
            else:
                still.append(req)
        self.waiting = deque(sorted(still, key=lambda r: r.arrival))
        return n

    def next_arrival(self):
        return self.pending[0].arrival if self.pending else None

    def stuck(self, live: int) -> bool:
        """Nothing runs and nothing arrives, yet requests wait: with an empty
        pool every request within max_blocks fits, so this only happens when
        case slots or the pool are held elsewhere.  The loop must end then
        instead of spinning."""
        return live == 0 and not self.pending and bool(self.waiting)


class Server:
    """A 3D and a 2D ensemble admitting into one block pool sized from the HBM
    budget (a 3D block takes the room of 30 2D blocks), each stepping on its
    own stream so their kernels overlap on the device."""

    def __init__(self, budget_bytes: float, max_cases: int = 4096, steps_per_poll: int = 4,
                 tuning=None, order: str = "shortest", live_cells_cap: float | None = 8 << 20,
                 max_wait: float = 30.0):
        import torch

        self.torch = torch
        classes = [(gpu.FLUX_HLLC, gpu.FP64), (gpu.FLUX_HLLC, gpu.MIXED)]
        self.block_bytes = {d: 3 * gpu.block_doubles(d) * 8 for d in (2, 3)}
        max3 = int(budget_bytes // self.block_bytes[3])
        if max3 < 1:
            raise ValueError("HBM budget below one 3D block")
        self.pools = {3: Ensemble(3, max_blocks=max3, max_cases=max_cases, classes=classes)}
        self.pools[2] = Ensemble(2, max_blocks=0, max_cases=max_cases, classes=classes,
                                 share_with=self.pools[3])
        # One pool: a 3D block is 30 2D blocks.
        self.max_blocks = {3: max3, 2: 30 * max3}
        for d in (2, 3):
            if tuning and d in tuning:
                self.pools[d].set_tuning(*tuning[d])
        self.budget = budget_bytes
        self.steps_per_poll = steps_per_poll
        self.order = order
        self.live_cells_cap = live_cells_cap
        self.max_wait = max_wait

    def close(self):
        for d in (2, 3):  # the 2D ensemble shares the 3D one's pool
            self.pools[d].close()

    def fits(self, req: Request) -> bool:
        free_blocks, largest, slots = self.pools[req.dim].capacity()
        return slots > 0 and largest >= req_blocks(req)

    def admit(self, req: Request) -> int:
        pool = self.pools[req.dim]
        with self.torch.cuda.stream(pool.stream):
            st0 = initial_state(req)
        return pool.add_case(case_desc(req), st0, req.t_end, req_cfl(req), sync=False)

    def serve(self, requests, keep_results=False, max_wall=None, log=None):
        """Serve the stream; keep_results: True (every request's final state)
        or a set of request ids."""
        torch = self.torch
        queue = Admission(requests, self.max_blocks, order=self.order,
                          live_cells_cap=self.live_cells_cap, max_wait=self.max_wait)
        live = {2: {}, 3: {}}  # case id -> request
        live_cells = 0
        done = []
        results = {}
        resident = 0.0
        occ_samples = []
        updates = 0
        t0 = time.perf_counter()
        last_log = t0

        def admit(req):
            nonlocal resident, live_cells
            cid = self.admit(req)
            live[req.dim][cid] = (req, time.perf_counter() - t0)
            resident += req_blocks(req) * self.block_bytes[req.dim]
            live_cells += req_cells(req)

        while queue.pending or queue.waiting or live[2] or live[3]:
            now = time.perf_counter() - t0
            if max_wall is not None and now > max_wall:
                break
            queue.arrive(now)
            queue.admit(self.fits, admit, live_cells=live_cells, now=now)
            nlive = len(live[2]) + len(live[3])
            if nlive == 0:
                if queue.stuck(nlive):
                    break
                nxt = queue.next_arrival()
                if nxt is not None:
                    time.sleep(max(0.0, min(0.01, nxt - now)))
                continue
            occ_samples.append(resident)
            # Both ensembles advance concurrently on their own streams.
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
                        if keep_results is True or (keep_results and req.rid in keep_results):
                            results[req.rid] = self.pools[d].download(cid)
                        self.pools[d].remove_case(cid)
                        resident -= req_blocks(req) * self.block_bytes[d]
                        live_cells -= req_cells(req)
                        done.append({"rid": req.rid, "dim": d, "nb": req.nb,
                                     "precision": req.precision, "arrival": req.arrival,
                                     "admitted": t_admit, "finished": tnow,
                                     "latency": tnow - req.arrival, "steps": rec.steps,
                                     "flags": rec.flags, "time": rec.time})
            if log and time.perf_counter() - last_log > 5.0:
                last_log = time.perf_counter()
                log(f"[serving] t={time.perf_counter() - t0:.1f}s done={len(done)} "
                    f"live={len(live[2]) + len(live[3])} waiting={len(queue.waiting)} "
                    f"pending={len(queue.pending)}")
        wall = time.perf_counter() - t0
        torch.cuda.synchronize()
        lat = np.array([r["latency"] for r in done]) if done else np.array([0.0])
        stats = {
            "requests": len(requests),
            "completed": len(done),
            "rejected": len(queue.rejected),
            "unserved": len(queue.waiting) + len(queue.pending),
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


def _alone_fp64(req: Request):
    """A request run by itself in FP64 (its twin for the mixed error)."""
    twin = dataclasses.replace(req, precision=gpu.FP64)
    ens = Ensemble(twin.dim, max_blocks=req_blocks(twin), max_cases=1,
                   classes=[(gpu.FLUX_HLLC, gpu.FP64)])
    cid = ens.add_case(case_desc(twin), initial_state(twin), twin.t_end, req_cfl(twin))
    ens.run_until_done([cid], poll=8)
    out = ens.download(cid)
    ens.close()
    return out


def stream_summary(rate: float = 200.0, window: float = 2.0, seed: int = 0,
                   budget: float = 200e9, twins: int = 8, log=None, order: str = "shortest",
                   live_cells_cap=8 << 20, batched_reference: bool = True,
                   steps_per_poll: int = 16):
    """Serve a Poisson stream of the serving mix and summarise it: latency,
    sustained rate, HBM occupancy, and the L1 density error / mass and energy
    drift of the first `twins` mixed requests against their FP64 twins (run
    alone afterwards).  With batched_reference the same requests are also
    run as one batch (all admitted at t = 0, in arrival order, no cap): the
    engine's rate on exactly this mix, against which the stream's sustained
    rate is stated."""
    from . import autotune

    reqs = generate_requests(rate, window, seed=seed)
    tuning = {d: autotune.lookup(d, gpu.FLUX_HLLC, gpu.FP64) for d in (2, 3)}
    mixed = [r for r in reqs if r.precision == gpu.MIXED][:twins]
    server = Server(budget, tuning=tuning, order=order, live_cells_cap=live_cells_cap,
                    steps_per_poll=steps_per_poll)
    stats, done, results = server.serve(reqs, keep_results={r.rid for r in mixed}, log=log)
    server.close()
    stats["admission"] = {"order": order, "live_cells_cap": live_cells_cap,
                          "max_wait_s": server.max_wait, "steps_per_poll": steps_per_poll}
    if batched_reference:
        import torch

        torch.cuda.empty_cache()
        batch = [dataclasses.replace(r, arrival=0.0) for r in reqs]
        # A pool just large enough for every request at once.
        need = sum(req_blocks(r) * (30 if r.dim == 3 else 1) for r in batch)
        ref_budget = min(budget, 1.05 * (need + 60) * 3 * gpu.block_doubles(2) * 8)
        server = Server(ref_budget, tuning=tuning, order="fifo", live_cells_cap=None)
        bstats, _, _ = server.serve(batch, log=log)
        server.close()
        stats["batched_same_requests"] = {
            "cell_updates_per_s": bstats["cell_updates_per_s"], "wall_s": bstats["wall_s"],
            "latency_p50_s": bstats["latency_p50_s"], "latency_p99_s": bstats["latency_p99_s"],
            "note": "the same requests all admitted at t=0 (FIFO, no cap): the engine's rate on this mix"}
        stats["sustained_vs_batched"] = (stats["cell_updates_per_s"] /
                                         max(bstats["cell_updates_per_s"], 1e-30))
    errs = []
    for r in mixed:
        if r.rid not in results:
            continue
        m, f = results[r.rid], _alone_fp64(r)
        errs.append({"rid": r.rid, "dim": r.dim, "cells": req_blocks(r) * 16 ** r.dim,
                     "l1_rho_err_vs_fp64": float(np.abs(m[0] - f[0]).sum() / np.abs(f[0]).sum()),
                     "mass_rel_diff_vs_fp64": float(abs(m[0].sum() - f[0].sum()) / abs(f[0]).sum()),
                     "energy_rel_diff_vs_fp64": float(abs(m[1].sum() - f[1].sum()) /
                                                      abs(f[1]).sum())})
    stats["mixed_vs_fp64_twins"] = errs
    if errs:
        stats["mixed_l1_rho_err_worst"] = max(e["l1_rho_err_vs_fp64"] for e in errs)
    stats["workload"] = (f"Poisson request stream {rate:g} req/s for {window:g} s: 70% 2D "
                         "four-quadrant 64^2-512^2, 30% 3D blob+shock 64^3-128^3, t_end "
                         f"U[0.05,0.25], FP64/mixed 50/50, HBM budget {budget / 1e9:g} GB (one "
                         "block pool for 2D and 3D), continuous batching")
    return stats


def main_bench(args):
    """bench.py --workload serving: one JSON line."""
    rate = getattr(args, "rate", 200.0)
    window = getattr(args, "window", 2.0)
    stats = stream_summary(rate, window, log=lambda m: print(m, flush=True))
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
        "config": {"workload": stats["workload"]},
        "serving": stats,
    }
    print(json.dumps(line), flush=True)
