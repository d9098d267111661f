"""Tile-shape autotuner for the ensemble stage kernels.

For a (dimension, flux, precision) class it times every tile configuration
(3D: 8x8x8 cells / 256 threads, 16x8x8 / 512; 2D: 16x16 / 128, 16x16 / 256,
16x8 / 128) and a few persistent-grid sizes (workgroups per CU) on a
throwaway ensemble holding a sample of the real case mix, and keeps the
fastest.  The choice changes only scheduling: every configuration computes
the same faces with the same arithmetic.  Results are cached as JSON keyed by
GPU architecture in ~/.cache/alpaca_amd/ (ALPACA_TUNE_CACHE overrides), and
the shipped defaults in tune_defaults.json (measured on gfx950) are used when
tuning is not requested.
"""
from __future__ import annotations

import json
import os
import time

from . import gpu
from .ensemble import Ensemble

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS = os.path.join(_HERE, "tune_defaults.json")
TILES = {2: [0, 1, 2], 3: [0, 1]}
GROUPS = {2: [0, 4, 8, 12], 3: [0, 1, 2, 4]}


def _arch():
    import torch

    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:
        return "unknown"


def _cache_path():
    path = os.environ.get("ALPACA_TUNE_CACHE")
    if path:
        return path
    return os.path.join(os.path.expanduser("~"), ".cache", "alpaca_amd", f"tune_{_arch()}.json")


def _key(dim, flux, precision):
    return f"dim{dim}_flux{flux}_prec{precision}"


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def lookup(dim, flux, precision):
    """(tile_config, groups_per_cu) from the cache, else the shipped
    defaults, else (0, 0)."""
    key = _key(dim, flux, precision)
    for path in (_cache_path(), DEFAULTS):
        entry = _load(path).get(_arch(), {}).get(key)
        if entry:
            return int(entry["tile"]), int(entry["groups_per_cu"])
    return 0, 0


def tune(dim, flux, precision, sample_cases, steps=6, repeats=2, verbose=False):
    """Time every candidate on a fresh ensemble holding sample_cases
    [(desc, state0), ...]; returns (tile, groups_per_cu, table)."""
    import torch

    nblocks = sum(d.nb[0] * d.nb[1] * (d.nb[2] if d.dim == 3 else 1) for d, _ in sample_cases)
    table = []
    best = None
    for tile in TILES[dim]:
        for gpc in GROUPS[dim]:
            ens = Ensemble(dim, max_blocks=nblocks, max_cases=len(sample_cases),
                           classes=[(flux, precision)])
            ens.set_tuning(tile, gpc)
            for d, s0 in sample_cases:
                ens.add_case(d, s0, 1e300, 0.5)
            ens.step(1)  # capture + warm up
            ens.stream.synchronize()
            ms = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                ens.step(steps)
                ens.stream.synchronize()
                ms.append((time.perf_counter() - t0) * 1e3 / steps)
            ens.close()
            torch.cuda.empty_cache()
            t = min(ms)
            table.append({"tile": tile, "groups_per_cu": gpc, "ms_per_step": t})
            if verbose:
                print(f"  tile {tile} groups/CU {gpc}: {t:.3f} ms/step", flush=True)
            if best is None or t < best[2]:
                best = (tile, gpc, t)
    return best[0], best[1], table


def tune_and_store(dim, flux, precision, sample_cases, path=None, **kw):
    tile, gpc, table = tune(dim, flux, precision, sample_cases, **kw)
    path = path or _cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    data = _load(path)
    data.setdefault(_arch(), {})[_key(dim, flux, precision)] = {
        "tile": tile, "groups_per_cu": gpc, "table": table}
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1)
    return tile, gpc, table


def precision_name(p):
    return {gpu.FP64: "fp64", gpu.FP64_FMA: "fp64-fma", gpu.MIXED: "mixed"}[p]
