#!/usr/bin/env python3
"""Throughput benchmark of the FP64 WENO5-HLLC-RK3 path on one MI355X.

Workload (default): 3D ideal gas (gamma 1.4) on the periodic unit cube, 256^3
cells = 16^3 leaf blocks of 16^3 with 4 halo cells, WENO5-JS characteristic
reconstruction + HLLC (Einfeldt) + SSP-RK3 at CFL 0.5, FP64.  Initial state:
32 seeded Gaussian density/pressure blobs and a planar Mach-2 shock.  A step
is one full macro step (CFL reduction, three RK stages, halo fills) of the
whole domain; a cell-update is one internal cell advanced one macro step.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `roofline` covers the dominant kernel (the fused stage
kernel) timed with HIP events on its launch stream during the timed region
(its `from_profile` part holds the counter figures of the profiled run in
profiles/, not measured here); `modes` repeats the workload in the
FP64-with-FMA and mixed forms, each compared with FP64 after 100 steps
(density L1/Linf error, mass/energy drift); `roe_fp64` runs it with Roe flux
splitting in FP64; `mr_operators` times the multiresolution kernels on 512
parents and their 4096 children; `c0_2d_128` runs the 2D 128^2 configuration
to t_end in FP64 and mixed; `batched` runs 512 uneven 2D cases in one
ensemble against one at a time (FP64 and mixed, p50/p99 time to solution,
per-case mixed error); `serving` serves a 200 req/s Poisson stream of the
serving mix for 1 s (p50/p99 time to solution, sustained rate, HBM
occupancy, mixed error of sampled requests against FP64 twins);
`cpu_baseline` times the reference's own numerics (oracle/_ref, threads over
blocks) on a bounded sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("cell-updates/s WENO5-HLLC-RK3, FP64 vs mixed at stated L1 err; "
          "p50/p99 case latency")

# Algorithmic FLOP per cell face and per cell-update, counted from the
# reference code (+ - * / sqrt = 1; min/max/abs/compare = 0).
FLOP_PER_FACE = {(3, 0): 1476.0, (3, 1): 1996.0, (2, 0): 1102.0, (2, 1): 1405.0}
# Per cell and RK stage: D faces + divergence + RK combination + cell primes.
FLOP_PER_CELL_STAGE = {(3, 0): 4549.0, (3, 1): 6109.0, (2, 0): 2257.0, (2, 1): 2863.0}
FLOP_PER_CELL_UPDATE = {(3, 0): 13.7e3, (3, 1): 18.3e3, (2, 0): 6.8e3, (2, 1): 8.6e3}
PRECISIONS = {"fp64": 0, "mixed": 1, "fp64-fma": 2}
PEAK_FP64 = 78.6e12   # MI355X FP64 vector/matrix, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X FP32 vector (packed), FLOP/s
PEAK_HBM = 8.0e12     # bytes/s


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def host_core_counts():
    """Cores this process may run on (sched_getaffinity) and the host's
    logical CPUs (lscpu's CPU(s) line; os.cpu_count() if lscpu is absent)."""
    import subprocess

    host = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("CPU(s):"):
                host = int(ln.split(":", 1)[1])
                break
    except (OSError, ValueError, subprocess.SubprocessError):
        host = None
    return len(os.sched_getaffinity(0)), host or os.cpu_count()


def cpu_baseline(flux: int, threads: int, steps: int = 5):
    """The reference's own numerics (oracle/_ref: its translation units,
    compiled here; the C restatement if that is not built) on a 128^3
    periodic sample of the same workload (the 8^3-block form of the blob +
    shock state), `steps` macro steps, threads over blocks."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from alpaca_amd import cases

    nb = [8, 8, 8]
    state = cases.blobs_shock_3d(8, seed=0)
    which = "ref" if oracle_py.ref_available(3) else "oracle"
    t0 = time.perf_counter()
    oracle_py.run(3, nb, [1, 1, 1], flux, 0.5, 1e300, steps, state, cell_size=1.0 / 128,
                  threads=threads, which=which)
    dt = time.perf_counter() - t0
    cells = 128 ** 3
    affinity, host = host_core_counts()
    return {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": threads,
            "kind": "reference" if which == "ref" else "port",
            "cores_affinity": affinity, "cores_host_lscpu": host,
            "binary": ("oracle/_ref/libalpaca_ref3d.so: the reference's translation units "
                       "built from source by oracle/Makefile" if which == "ref"
                       else "oracle/liboracle.so (C restatement)"),
            "sample": f"128^3 periodic blob+shock (8^3 blocks), {steps} steps, "
                      f"{'HLLC' if flux == 0 else 'Roe'} FP64, {threads} threads over blocks, "
                      f"{dt:.2f} s wall ({dt * threads:.0f} core-seconds)"}


def load_profile(kernel_tag: str):
    """Counter-derived figures of a kernel (profiles/traffic.json, written from
    the rocprofv3 --pmc runs the profile summaries in profiles/ describe), or {}."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(kernel_tag) or {}
    except (OSError, ValueError):
        return {}


def run_mode(args, flux, precision, nb, keep_state=False, state_after=0):
    """Time args.steps macro steps of one arithmetic mode; returns a dict.
    With keep_state, the density and the equation sums are kept after
    max(state_after, warmup + steps) steps (the untimed rest run after the
    timed region)."""
    import torch

    from alpaca_amd import cases, gpu

    prec = PRECISIONS[precision]
    desc = gpu.make_desc(3, [nb, nb, nb], [gpu.BC_PERIODIC] * 3, flux=flux, precision=prec)
    ic = cases.blobs_shock_3d_torch(nb, seed=0)
    case = gpu.GpuCase(desc, ic, cfl=0.5)
    del ic
    torch.cuda.synchronize()
    case.advance(args.warmup)
    torch.cuda.synchronize()
    gpu.profile_enable(True)
    for k in range(gpu.KERNEL_STAGE + 1):
        gpu.profile_read(k, reset=True)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    case.advance(args.steps)
    stop.record()
    torch.cuda.synchronize()
    elapsed = start.elapsed_time(stop) / 1e3
    gpu.profile_enable(False)
    rec = case.record()
    if rec.flags & 1:
        raise SystemExit(f"bench: non-finite state ({precision})")
    cells = (nb * 16) ** 3
    stage_ms, stage_launches, stage_cells = gpu.profile_read(gpu.KERNEL_STAGE)
    ts_ms, ts_launches, _ = gpu.profile_read(gpu.KERNEL_TIMESTEP)
    out = {
        "value": cells * args.steps / elapsed,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "stage_launches": stage_launches,
        "stage_avg_ms": stage_ms / max(stage_launches, 1),
        "cells_per_launch": stage_cells / max(stage_launches, 1),
        "timestep_ms_total": ts_ms,
        "time": rec.time,
        "steps_total": rec.steps,
    }
    if keep_state:
        rest = state_after - (args.warmup + args.steps)
        if rest > 0:
            case.advance(rest)
            torch.cuda.synchronize()
        out["compared_after_steps"] = case.record().steps
        out["rho"] = gpu.from_blocks(case.u, 3, case.nb)[0].clone()
        out["sums"] = gpu.from_blocks(case.u, 3, case.nb).sum(dim=(1, 2, 3)).cpu().tolist()
    del case
    torch.cuda.empty_cache()
    return out


def batched_specs(ncases, seed=0):
    """512 uneven 2D four-quadrant cases (cases.batched_cases)."""
    from alpaca_amd import cases

    return cases.batched_cases(ncases, seed)


def batched_summary(args):
    """All cases in one ensemble (one persistent kernel per RK stage, graph
    step) versus the same cases run one at a time on the same engine; FP64 and
    mixed, with the mixed density error against FP64 per case."""
    import numpy as np
    import torch

    from alpaca_amd import autotune, gpu
    from alpaca_amd.ensemble import CASE_STOP, Ensemble

    specs = batched_specs(args.cases)
    total_blocks = sum(nb * nb for nb, _, _ in specs)
    out = {}
    finals = {}
    for prec_name in ("fp64", "mixed"):
        prec = PRECISIONS[prec_name]
        descs = [gpu.make_desc(2, [nb, nb], [0, 0], flux=gpu.FLUX_HLLC, precision=prec,
                               gamma=g) for nb, g, _ in specs]
        if args.tune:
            sample = [(d, s0) for d, (_, _, s0) in list(zip(descs, specs))[:64]]
            tile, gpc, table = autotune.tune_and_store(2, gpu.FLUX_HLLC, prec, sample)
        else:
            tile, gpc = autotune.lookup(2, gpu.FLUX_HLLC, prec)
            table = None
        ens = Ensemble(2, max_blocks=total_blocks, max_cases=len(specs),
                       classes=[(gpu.FLUX_HLLC, prec)])
        ens.set_tuning(tile, gpc)
        ids = [ens.add_case(d, s0, 0.2, 0.6) for d, (_, _, s0) in zip(descs, specs)]
        ens.step(0)
        ens.stream.synchronize()
        done_at = {}
        t0 = time.perf_counter()
        while len(done_at) < len(ids):
            ens.step(args.poll)
            st = ens.states()
            now = time.perf_counter() - t0
            for c in ids:
                if c not in done_at and st[c].flags & CASE_STOP:
                    done_at[c] = now
        wall = time.perf_counter() - t0
        st = ens.states()
        updates = sum(nb * nb * 256 * st[c].steps for c, (nb, _, _) in zip(ids, specs))
        lat = np.array(sorted(done_at.values()))
        finals[prec_name] = [ens.download(c)[0] for c in ids]
        ens.close()
        torch.cuda.empty_cache()
        # One at a time on the same engine (same kernels and graph step).
        solo = Ensemble(2, max_blocks=max(nb * nb for nb, _, _ in specs), max_cases=1,
                        classes=[(gpu.FLUX_HLLC, prec)])
        solo.set_tuning(tile, gpc)
        solo_wall = 0.0
        for d, (_, _, s0) in zip(descs, specs):
            cid = solo.add_case(d, s0, 0.2, 0.6)
            solo.stream.synchronize()
            t1 = time.perf_counter()
            while not solo.states(1)[cid].flags & CASE_STOP:
                solo.step(args.poll)
            solo_wall += time.perf_counter() - t1
            solo.remove_case(cid)
        solo.close()
        out[prec_name] = {
            "cell_updates_per_s": updates / wall,
            "wall_s": wall,
            "cell_updates": updates,
            "one_at_a_time_cell_updates_per_s": updates / solo_wall,
            "one_at_a_time_wall_s": solo_wall,
            "batched_speedup": solo_wall / wall,
            "case_latency_p50_s": float(np.percentile(lat, 50)),
            "case_latency_p99_s": float(np.percentile(lat, 99)),
            "tile_config": tile,
            "groups_per_cu": gpc,
        }
        if table is not None:
            out[prec_name]["tuning_table"] = table
    errs = [float(np.abs(m - f).sum() / np.abs(f).sum())
            for m, f in zip(finals["mixed"], finals["fp64"])]
    out["mixed"]["l1_rho_err_vs_fp64_max_over_cases"] = max(errs)
    out["mixed"]["l1_rho_err_vs_fp64_mean_over_cases"] = float(np.mean(errs))
    out["mixed"]["speedup_vs_fp64"] = out["mixed"]["cell_updates_per_s"] / out["fp64"]["cell_updates_per_s"]
    out["cases"] = len(specs)
    out["cells"] = sum(nb * nb * 256 for nb, _, _ in specs)
    out["leaf_blocks"] = total_blocks
    out["workload"] = (f"{len(specs)} uneven 2D four-quadrant cases (32^2..256^2, weights "
                       ".4/.3/.2/.1, gamma U[1.2,1.67]), HLLC, CFL 0.6, t_end 0.2, one persistent "
                       "kernel per RK stage, graph-captured step; latency = time to solution "
                       "with all cases admitted at t=0")
    return out


def run_batched(args):
    out = batched_summary(args)
    specs_n, cells, total_blocks = out["cases"], out["cells"], out["leaf_blocks"]
    line = {
        "metric": METRIC,
        "value": out["fp64"]["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": out["workload"], "cases": specs_n, "cells": cells,
                   "leaf_blocks": total_blocks},
        "batched": out,
    }
    print(json.dumps(line), flush=True)


def c0_summary():
    """The CPU-runnable configuration on the GPU: 2D four-quadrant (seed 0),
    128^2 cells = 8x8 blocks, HLLC, CFL 0.6, t_end 0.2, zero-gradient sides,
    FP64 and mixed: throughput and the mixed density error and mass/energy
    difference against FP64 at t_end."""
    import numpy as np
    import torch

    from alpaca_amd import cases, gpu

    st0 = cases.four_quadrant_2d(8, 8, seed=0)
    out = {"workload": "2D four-quadrant seed 0, 128^2 cells (8x8 blocks of 16^2), HLLC, CFL 0.6, "
                       "t_end 0.2, zero-gradient sides"}
    finals = {}
    for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
        desc = gpu.make_desc(2, [8, 8], [0, 0], flux=gpu.FLUX_HLLC, precision=prec)
        for sync_every in (1, 32):
            case = gpu.GpuCase(desc, st0, cfl=0.6, t_end=0.2)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            steps, _ = case.run(max_steps=100000, sync_every=sync_every)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            if sync_every == 1:
                finals[name] = case.download()
                out[name] = {"steps": steps, "wall_s": wall,
                             "cell_updates_per_s": 128 * 128 * steps / wall,
                             "time": case.record().time}
            else:
                assert np.array_equal(case.download(), finals[name])
                out[name]["chunked"] = {"sync_every": sync_every, "steps": steps, "wall_s": wall,
                                        "cell_updates_per_s": 128 * 128 * steps / wall}
    f, m = finals["fp64"], finals["mixed"]
    out["mixed"]["l1_rho_err_vs_fp64"] = float(np.abs(m[0] - f[0]).sum() / np.abs(f[0]).sum())
    out["mixed"]["linf_rho_err_vs_fp64"] = float(np.abs(m[0] - f[0]).max() / np.abs(f[0]).max())
    out["mixed"]["mass_rel_diff_vs_fp64"] = float(abs(m[0].sum() - f[0].sum()) / abs(f[0]).sum())
    out["mixed"]["energy_rel_diff_vs_fp64"] = float(abs(m[1].sum() - f[1].sum()) / abs(f[1]).sum())
    out["note"] = ("64 blocks cannot fill 256 CUs: the rate is launch- and latency-bound; "
                   "the main figures read the case record back every step (the reference's "
                   "ComputeLoop checks the end time every step); 'chunked' reads it every 32 "
                   "steps, the device skipping the steps after t_end (same final state)")
    return out


def run_accuracy(args):
    """Mixed vs FP64 on the 256^3 HLLC case: relative L1 / Linf density error
    and mass / energy drift of the mixed form against the FP64 path after each
    checkpoint step count."""
    import numpy as np
    import torch

    from alpaca_amd import cases, gpu

    nb = args.blocks
    checkpoints = sorted({int(c) for c in args.checkpoints.split(",")})
    runs = {}
    for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
        desc = gpu.make_desc(3, [nb] * 3, [gpu.BC_PERIODIC] * 3, flux=gpu.FLUX_HLLC,
                             precision=prec)
        case = gpu.GpuCase(desc, cases.blobs_shock_3d_torch(nb, seed=0), cfl=0.5)
        g0 = gpu.from_blocks(case.u, 3, case.nb)
        sums0 = g0.sum(dim=(1, 2, 3)).cpu().numpy()
        del g0
        done = 0
        snaps = {}
        for c in checkpoints:
            case.advance(c - done)
            done = c
            g = gpu.from_blocks(case.u, 3, case.nb)
            snaps[c] = {"rho": g[0].clone(), "sums": g.sum(dim=(1, 2, 3)).cpu().numpy(),
                        "time": case.record().time}
            del g
            print(f"[accuracy] {name} step {c} t={snaps[c]['time']:.6f}", flush=True)
        runs[name] = (sums0, snaps)
        del case
        torch.cuda.empty_cache()
    rows = []
    for c in checkpoints:
        ref = runs["fp64"][1][c]
        mix = runs["mixed"][1][c]
        d = (mix["rho"] - ref["rho"]).abs()
        s0 = runs["fp64"][0]
        rows.append({
            "steps": c,
            "time_fp64": ref["time"], "time_mixed": mix["time"],
            "l1_rho_err": float(d.sum() / ref["rho"].abs().sum()),
            "linf_rho_err": float(d.max() / ref["rho"].abs().max()),
            "mass_drift_mixed": float(abs(mix["sums"][0] - s0[0]) / abs(s0[0])),
            "energy_drift_mixed": float(abs(mix["sums"][1] - s0[1]) / abs(s0[1])),
            "mass_drift_fp64": float(abs(ref["sums"][0] - s0[0]) / abs(s0[0])),
            "energy_drift_fp64": float(abs(ref["sums"][1] - s0[1]) / abs(s0[1])),
        })
    print(json.dumps({"workload": f"3D {nb * 16}^3 periodic blob+shock HLLC, mixed vs FP64",
                      "rows": rows}), flush=True)


# Algorithmic HBM bytes per link of the multiresolution kernels (3D, five
# equations, doubles): averaging reads the child's 16^3 internal cells and
# writes the parent's 8^3 octant; prediction reads the parent's stencil window
# (predicted cells + 2 per side: 16^3 for the whole block, 12^3 for the
# internal cells) and writes the child cells (24^3 or 16^3); the details read
# the parent window and all 24^3 child cells of Mass and Energy.
MR_BYTES_PER_LINK = {
    "average": (16 ** 3 + 8 ** 3) * 5 * 8,
    "prediction_block": (16 ** 3 + 24 ** 3) * 5 * 8,
    "prediction_internal": (12 ** 3 + 16 ** 3) * 5 * 8,
    "details": (16 ** 3 + 24 ** 3) * 2 * 8,
    # level-wise integration, per listed block: buffer preparation reads two
    # and writes one 24^3 x 5 block; IntegrateNode reads and writes the jump
    # surfaces (cons, flux: read 2, write 2 of 6 x 5 x 16^2) and the internal
    # cells (read 2, write 1 of 16^3 x 5).
    "lts_prepare_buffer": 3 * 24 ** 3 * 5 * 8,
    "lts_integrate_node": (4 * 6 * 5 * 16 ** 2 + 3 * 16 ** 3 * 5) * 8,
    # halo update across a level jump, per (child, face location): writes the
    # 4 x 16 x 16 slab, reads the parent window of its 2 x 8 x 8 parent cells
    # + 2 per side (6 x 12 x 12).
    "jump_halos_faces": (4 * 16 * 16 + 6 * 12 * 12) * 5 * 8,
}


def run_mr(args, nparents=512):
    """Multiresolution kernels on 512 parents and their 4096 children (3D;
    the 256^3 cells of level 1): each operator over all links per launch,
    timed with HIP events on the launch stream."""
    import torch

    from alpaca_amd import gpu
    from alpaca_amd import multiresolution as mr

    dim = 3
    n = nparents * 9
    bd = gpu.block_doubles(dim)
    g = torch.Generator(device="cuda").manual_seed(5)
    blocks = (torch.rand(n * bd, dtype=torch.float64, device="cuda", generator=g) + 0.5)
    rows = [(p * 9 + 1 + pos, p * 9, pos, 1) for p in range(nparents) for pos in range(8)]
    links = torch.as_tensor(mr.make_links(rows), device="cuda")
    nl = len(rows)
    m = mr.Multiresolution(mr.ThresholdComputer(dim, 1, 1, 1e-3))
    out = (torch.empty(nl, dtype=torch.float64, device="cuda"),
           torch.empty(nl, dtype=torch.int32, device="cuda"))
    ops = {
        "average": lambda: m.average(blocks, dim, links),
        "prediction_block": lambda: m.prediction(blocks, dim, links),
        "prediction_internal": lambda: m.prediction(blocks, dim, links, (4, 16, 4, 16, 4, 16)),
        "details": lambda: m.child_needs_remeshing(blocks, dim, links, out=out),
    }
    # Level-wise integration on the 4096 children as one level (LTS stage
    # operations): jump surfaces for the exchange-free operations.
    from alpaca_amd import lts as lts_mod

    per_surf = 6 * (dim + 2) * 16 * 16
    jc = torch.rand(n * per_surf, dtype=torch.float64, device="cuda", generator=g)
    jf = torch.rand(n * per_surf, dtype=torch.float64, device="cuda", generator=g)
    init = torch.rand(n * bd, dtype=torch.float64, device="cuda", generator=g)
    rhs = torch.rand(n * bd, dtype=torch.float64, device="cuda", generator=g)
    level = torch.as_tensor([r[0] for r in rows], dtype=torch.int32, device="cuda")
    ti = lts_mod.LevelTimeIntegrator(dim)
    for dt in (1e-4, 1.1e-4):
        ti.append_micro_timestep(dt)
    ops["lts_prepare_buffer"] = lambda: ti.prepare_buffer_for_integration(blocks, init, level, 1)
    ops["lts_integrate_node"] = lambda: ti.integrate_node(blocks, rhs, jc, jf, level, 1, 2)
    # Jump halos: every child's six face slabs (East ... Bottom) in one launch.
    jrows = [r for r in rows for _ in range(6)]
    jlinks = torch.as_tensor(mr.make_links(jrows), device="cuda")
    jlocs = torch.as_tensor([loc for _ in rows for loc in range(6)], dtype=torch.int32,
                            device="cuda")
    ops["jump_halos_faces"] = lambda: m.jump_halo_update(blocks, dim, jlinks, jlocs)
    counts = {"jump_halos_faces": len(jrows)}
    stream = torch.cuda.current_stream()
    reps = max(args.steps, 1)
    res = {}
    for name, op in ops.items():
        for _ in range(max(args.warmup, 1)):
            op()
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(reps):
            op()
        t1.record(stream)
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / reps
        cnt = counts.get(name, nl)
        bytes_launch = MR_BYTES_PER_LINK[name] * cnt
        gbs = bytes_launch / (ms / 1e3) / 1e9
        res[name] = {"avg_launch_ms": ms, "links": cnt, "bytes_per_link": MR_BYTES_PER_LINK[name],
                     "hbm_GBs": gbs, "frac": gbs / (PEAK_HBM / 1e9)}
    return {"workload": f"3D MR operators, {nparents} parents x 8 children (256^3 on level 1), "
                        "FP64; lts_* over the 4096 children as one level; jump_halos_faces: the six "
                        "face slabs of every child", "kernels": res}


def run_multilevel(warmup: int = 1, steps: int = 3):
    """The multi-level mode on the GPU backend: 3D periodic blob, 4^3 level-0
    nodes of size 0.25, maximum level 2 (multiresolution with local time
    stepping, remeshing every micro step), HLLC FP64.  A leaf of level l
    advances 2^l times per macro step; the rate counts those leaf
    cell-updates."""
    import time

    import torch

    from alpaca_amd import cases, gpu, multilevel

    dim, lmax, nb0 = 3, 2, [4, 4, 4]
    cap = 64 * (1 + 8 + 64)
    be = multilevel.GpuBackend(dim, cap, flux=gpu.FLUX_HLLC)
    sol = multilevel.MultiLevelSolver(be, dim, nb0, 0.25, lmax, [gpu.BC_PERIODIC] * 6,
                                      cases.blob_node_initial(dim, 0.25), 0.5, 10.0, 1e-3, lmax)
    t0 = time.perf_counter()
    sol.create_new_simulation()
    torch.cuda.synchronize()
    create_s = time.perf_counter() - t0
    for _ in range(warmup):
        sol.advance()
    torch.cuda.synchronize()
    updates = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        lv = [len(sol.topo.leaves_on_level(level)) for level in range(lmax + 1)]
        updates += sum(n * 16 ** 3 * (1 << level) for level, n in enumerate(lv))
        sol.advance()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    leaves = [len(sol.topo.leaves_on_level(level)) for level in range(lmax + 1)]
    out = {"workload": ("3D periodic blob, 4^3 level-0 nodes, maximum level 2 (MR + local time "
                        "stepping, remeshing each micro step), HLLC FP64, GPU backend"),
           "create_s": create_s, "macro_steps": steps, "macro_step_ms": wall * 1e3 / steps,
           "leaves_per_level_after": leaves,
           "leaf_cell_updates_per_s": updates / wall,
           "finest_uniform_cells": 64 * 4 ** 3 * 16 ** 3,
           "note": ("leaf cells are a fraction of the finest level's uniform cells; the rate "
                    "counts every leaf's micro steps")}
    be.close()
    del be, sol
    torch.cuda.empty_cache()
    return out


def run_copy_probe():
    """A plain device copy of known size, for scaling the HBM counters."""
    import torch

    n = 1 << 28  # 2 GiB of doubles
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read_per_launch": n * 8,
                      "copy_probe_bytes_written_per_launch": n * 8}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--flux", choices=["hllc", "roe"], default="hllc")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="fp64")
    ap.add_argument("--blocks", type=int, default=16, help="leaf blocks per side (16 -> 256^3)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-modes", action="store_true",
                    help="time only --precision (skip the other arithmetic modes)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the C0, batched and serving legs of the default line")
    ap.add_argument("--compare-steps", type=int, default=100,
                    help="uniform: step count after which the modes are compared with FP64")
    ap.add_argument("--serving-window", type=float, default=1.0,
                    help="arrival window (s) of the serving leg of the default line")
    ap.add_argument("--workload", choices=["uniform", "batched", "serving", "accuracy", "mr",
                                           "multilevel", "c0"],
                    default="uniform",
                    help="uniform: the 256^3 case (default); batched: 512 uneven 2D cases; "
                         "serving: Poisson request stream; c0: the 2D 128^2 case to t_end")
    ap.add_argument("--cases", type=int, default=512, help="batched: number of cases")
    ap.add_argument("--poll", type=int, default=8, help="batched: steps between polls")
    ap.add_argument("--tune", action="store_true", help="batched: autotune the tile shape")
    ap.add_argument("--rate", type=float, default=200.0, help="serving: requests per second")
    ap.add_argument("--window", type=float, default=2.0, help="serving: arrival window (s)")
    ap.add_argument("--checkpoints", default="100,1000", help="accuracy: step counts")
    ap.add_argument("--copy-probe", action="store_true",
                    help="time nothing: run a plain 2 GiB device copy (counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding of the block grid is a design note, not a run mode.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    torch.cuda.set_device(0)
    if args.copy_probe:
        run_copy_probe()
        return
    if args.workload == "accuracy":
        run_accuracy(args)
        return
    if args.workload == "batched":
        run_batched(args)
        return
    if args.workload == "mr":
        print(json.dumps(run_mr(args)), flush=True)
        return
    if args.workload == "multilevel":
        print(json.dumps(run_multilevel()), flush=True)
        return
    if args.workload == "c0":
        print(json.dumps(c0_summary()), flush=True)
        return
    if args.workload == "serving":
        from alpaca_amd import serving

        serving.main_bench(args)
        return
    flux = 0 if args.flux == "hllc" else 1
    nb = args.blocks
    n = nb * 16
    cells = n ** 3
    modes = [args.precision] if args.no_modes else ["fp64", "fp64-fma", "mixed"]
    if args.precision not in modes:
        modes.insert(0, args.precision)
    results = {}
    for m in modes:
        results[m] = run_mode(args, flux, m, nb, keep_state=not args.no_modes,
                              state_after=args.compare_steps)
    head = results[args.precision]
    value = head["value"]

    flop_cell_stage = FLOP_PER_CELL_STAGE[(3, flux)]
    avg_launch_s = head["stage_avg_ms"] / 1e3
    achieved = head["cells_per_launch"] * flop_cell_stage / avg_launch_s
    peak = PEAK_FP32 if args.precision == "mixed" else PEAK_FP64
    tag = f"stage_3d_{args.flux}_{args.precision}"
    variant = {"fp64": "exact::", "fp64-fma": "fast::", "mixed": "fast::"}[args.precision]
    rtype = "float" if args.precision == "mixed" else "double"
    prof = load_profile(tag)
    bound = "valu_fp32" if args.precision == "mixed" else "valu_fp64"
    roofline = {
        "bound": bound,
        "achieved": achieved / 1e12,
        "peak": peak / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": prof.get("hbm_bytes_per_launch"),
        "kernel": f"{variant}stage_kernel<3,{'HLLC' if flux == 0 else 'Roe'},{rtype},8,8,8,256>",
        "note": (f"compute roofline (the contract's 'mfma' slot): a VALU stencil, no contraction "
                 f"maps to MFMA, so the peak is the {'FP32' if args.precision == 'mixed' else 'FP64'} "
                 f"vector figure; achieved = {flop_cell_stage:.0f} algorithmic FLOP per cell and "
                 f"RK stage x {head['cells_per_launch']:.0f} cells per launch / avg launch "
                 f"{head['stage_avg_ms']:.3f} ms over {head['stage_launches']} launches (HIP "
                 f"events on the launch stream, this run); traffic = HBM bytes per launch from "
                 f"the rocprofv3 FETCH_SIZE/WRITE_SIZE runs in profiles/"),
    }
    if prof:
        # Counter figures of the profiled run (profiles/traffic.json, from the
        # rocprofv3 --pmc runs its summary in profiles/ gives): not measured in
        # this run.  Rates are these per-launch figures over this run's launch time.
        fp = {"source": "profiles/traffic.json (rocprofv3 --pmc)", "kernel": prof.get("kernel")}
        if prof.get("hbm_bytes_per_launch"):
            fp["hbm_bytes_per_launch"] = prof["hbm_bytes_per_launch"]
            fp["algorithmic_bytes_per_launch"] = prof.get("algorithmic_bytes_per_launch")
            fp["hbm_GBs_at_this_launch_time"] = prof["hbm_bytes_per_launch"] / avg_launch_s / 1e9
            fp["hbm_frac"] = fp["hbm_GBs_at_this_launch_time"] / 8000.0
        if prof.get("hw_flop_per_launch"):
            fp["hw_flop_per_launch"] = prof["hw_flop_per_launch"]
            fp["hw_valu_flop_frac_at_this_launch_time"] = (prof["hw_flop_per_launch"] /
                                                           avg_launch_s / peak)
        for k in ("valu_busy", "valu_issue_frac", "valu_instr_per_cell"):
            if prof.get(k) is not None:
                fp[k] = prof[k]
        roofline["from_profile"] = fp
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "mixed" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D {n}^3 periodic blob+Mach-2 shock, WENO5-JS char + "
                         f"{'HLLC(Einfeldt)' if flux == 0 else 'Roe flux splitting'} + SSP-RK3, "
                         f"CFL 0.5, {args.precision}"),
            "cells": cells,
            "leaf_blocks": nb ** 3,
            "block": "16^3 internal + 4 halo",
            "flux": args.flux,
            "precision": args.precision,
        },
        "roofline": roofline,
    }
    if len(results) > 1:
        summary = {}
        for m, r in results.items():
            pk = PEAK_FP32 if m == "mixed" else PEAK_FP64
            summary[m] = {"cell_updates_per_s": r["value"], "ms_per_step": r["ms_per_step"],
                          "stage_avg_ms": r["stage_avg_ms"],
                          "roofline_frac": (r["cells_per_launch"] * flop_cell_stage /
                                            (r["stage_avg_ms"] / 1e3) / pk),
                          "roofline_peak": "fp32 vector" if m == "mixed" else "fp64 vector"}
        ref = results.get("fp64")
        if ref is not None and "rho" in ref:
            for m, r in results.items():
                if m == "fp64" or "rho" not in r:
                    continue
                d = (r["rho"] - ref["rho"]).abs()
                summary[m]["l1_rho_err_vs_fp64"] = float(d.sum() / ref["rho"].abs().sum())
                summary[m]["linf_rho_err_vs_fp64"] = float(d.max() / ref["rho"].abs().max())
                summary[m]["mass_drift"] = abs(r["sums"][0] - ref["sums"][0]) / abs(ref["sums"][0])
                summary[m]["energy_drift"] = abs(r["sums"][1] - ref["sums"][1]) / abs(ref["sums"][1])
                summary[m]["speedup_vs_fp64"] = r["value"] / ref["value"]
            summary["compared_after_steps"] = ref.get("compared_after_steps")
        line["modes"] = summary
        if flux == 0:
            # The same workload with Roe flux splitting in FP64: the 100-step
            # oracle and conservation configuration, reported beside the
            # HLLC headline.
            roe = run_mode(args, 1, "fp64", nb, keep_state=False)
            line["roe_fp64"] = {"cell_updates_per_s": roe["value"], "ms_per_step": roe["ms_per_step"],
                                "stage_avg_ms": roe["stage_avg_ms"],
                                "roofline_frac": (roe["cells_per_launch"] * FLOP_PER_CELL_STAGE[(3, 1)]
                                                  / (roe["stage_avg_ms"] / 1e3) / PEAK_FP64)}
        # The multiresolution operators (averaging, prediction, details) of
        # the multi-level mode at the same cell count on level 1, and a
        # multi-level run through the driver.
        line["mr_operators"] = run_mr(args)
        print("[bench] multilevel", file=sys.stderr, flush=True)
        line["multilevel"] = run_multilevel()
    if not (args.no_extra or args.no_modes):
        # The other configurations at their stated sizes: C0 (2D 128^2), the
        # 512-case batched ensemble and the Poisson serving stream.
        print("[bench] c0", file=sys.stderr, flush=True)
        line["c0_2d_128"] = c0_summary()
        print("[bench] batched", file=sys.stderr, flush=True)
        line["batched"] = batched_summary(args)
        print("[bench] serving", file=sys.stderr, flush=True)
        from alpaca_amd import serving

        line["serving"] = serving.stream_summary(
            args.rate, args.serving_window,
            log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(flux, cpu_threads())
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
