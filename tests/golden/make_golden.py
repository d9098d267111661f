"""Write the golden vectors that pin the CPU oracle.

Run in a container that has the reference tree: it builds nothing itself;
``make -C oracle ref`` must have produced oracle/_ref/libalpaca_ref{2,3}d.so,
which link the reference's own translation units (eigendecomposition,
finite-volume and flux-splitting schemes, WENO5, stiffened-gas EOS, prime
state conversion, RungeKutta3TVD, zero-gradient BC).

Outputs (small, committed):
  golden_vectors.npz  per-function inputs/outputs (WENO5, EOS, HLLC) and
                      per-block face fluxes / right-hand sides (2D in full;
                      3D as a strided sample plus a SHA-256 of the full array);
  golden_runs.json    full trajectories: step count, every dt, final time,
                      SHA-256 of the final state's bytes, per-field L1 sums.

    python tests/golden/make_golden.py

With ``--full`` it instead runs the full-size 3D case (256^3 cells = 16^3
blocks, periodic, blob + Mach-2 shock seed 0, CFL 0.5, 100 steps, FP64) with
the reference's Roe flux splitting and HLLC schemes on all host threads (about
half an hour per scheme on 8 cores) and writes

  golden_full_256.json  per scheme: SHA-256 of the initial and final states,
                        every dt, final time, conservation sums;
  golden_full_256.npz   the final state at 4096 fixed cells (all equations).

    python tests/golden/make_golden.py --full [--threads N] [--flux roe|hllc]

With ``--ids`` it writes level0_ids.json only (level-0 node ids, used to pin
the restart files' node-id encoding); with ``--edge`` edge_blocks.json only
(hashes of the reference's per-block fluxes and right-hand sides on the
special-branch states of tests/edge_states.py); with ``--mr`` mr_golden.json
only (the reference's multiresolution averaging, prediction, details and
thresholds on the block pairs of tests/mr_states.py).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from alpaca_amd import cases  # noqa: E402
import oracle_py  # noqa: E402

GAMMA = 1.4


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def random_states(rng, n, dim, gamma=GAMMA, b=0.0):
    rho = rng.uniform(0.1, 2.0, n)
    p = rng.uniform(0.05, 2.0, n)
    vel = [rng.uniform(-1.5, 1.5, n) for _ in range(dim)]
    c = cases.primes_to_conservatives(dim, gamma, b, rho, vel, p)
    return np.ascontiguousarray(c.T)


def weno_inputs(rng):
    rows = [np.ones(6), np.array([0, 0, 0, 1, 1, 1.0]), np.array([1, 1, 1, 0, 0, 0.0]),
            np.zeros(6), np.arange(6, dtype=float), (np.arange(6, dtype=float)) ** 2]
    rows += list(rng.uniform(-2, 2, (40, 6)))
    rows += list(np.sin(np.linspace(0, 1, 6)[None, :] * rng.uniform(0.1, 6, (10, 1))))
    return np.array(rows)


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    # WENO5 (weno5.h:88-158).
    wi = weno_inputs(rng)
    out["weno_in"] = wi
    out["weno_out"] = np.array([oracle_py.weno5(v, which="ref") for v in wi])
    # EOS known answers in the form of ref_eos, for two materials.
    for dim in (2, 3):
        for tag, gamma, b in (("ideal", 1.4, 0.0), ("stiff", 1.6, 2.0)):
            cons = rng.uniform(0.5, 3.0, (20, 5))
            cons[:, 4] = rng.uniform(20.0, 40.0, 20)  # energy large enough for p > -B
            if dim == 2:
                cons[:, 3] = 0.0  # z momentum absent in 2D
            cons_in = cons[:, [0, 1, 2, 3, 4]]
            prim = np.column_stack([rng.uniform(0.5, 3.0, 20), rng.uniform(-1, 1, 20),
                                    rng.uniform(-1, 1, 20), rng.uniform(-1, 1, 20),
                                    rng.uniform(0.5, 4.0, 20)])
            if dim == 2:
                prim[:, 3] = 0.0
            res = np.array([oracle_py.eos(dim, gamma, b, c, p, which="ref")
                            for c, p in zip(cons_in, prim)])
            out[f"eos{dim}_{tag}_cons"] = cons_in
            out[f"eos{dim}_{tag}_prim"] = prim
            out[f"eos{dim}_{tag}_out"] = res
    # HLLC (hllc_riemann_solver.h:52-198) on random face states, every direction.
    for dim in (2, 3):
        cl = random_states(rng, 30, dim)
        cr = random_states(rng, 30, dim)
        # Include identical states (consistency) and supersonic pairs.
        cr[:5] = cl[:5]
        for d in range(dim):
            res = np.array([oracle_py.hllc(dim, d, GAMMA, 0.0, a, b_, which="ref")
                            for a, b_ in zip(cl, cr)])
            out[f"hllc{dim}_dir{d}"] = res
        out[f"hllc{dim}_left"] = cl
        out[f"hllc{dim}_right"] = cr
    # Per-block face fluxes and right-hand sides.
    meta = {"blocks": {}}
    for dim in (2, 3):
        for flux in (0, 1):
            for seed, smooth in ((11, False), (12, True)):
                blk = cases.random_block(dim, seed, smooth=smooth)
                f = oracle_py.block_fluxes(dim, flux, GAMMA, 0.0, blk, 1.0 / 128, which="ref")
                r = oracle_py.block_rhs(dim, flux, GAMMA, 0.0, blk, 1.0 / 128, which="ref")
                key = f"block{dim}_flux{flux}_seed{seed}"
                meta["blocks"][key] = {"dim": dim, "flux": flux, "seed": seed, "smooth": smooth,
                                       "fluxes_sha256": sha(f), "rhs_sha256": sha(r)}
                if dim == 2:
                    out[key + "_fluxes"] = f
                    out[key + "_rhs"] = r
                else:
                    out[key + "_fluxes_sample"] = f.ravel()[::97].copy()
                    out[key + "_rhs_sample"] = r.ravel()[::13].copy()
    np.savez_compressed(os.path.join(HERE, "golden_vectors.npz"), **out)

    runs = []
    # 2D four-quadrant, 128^2, zero-gradient, CFL 0.6, t_end 0.2.
    for flux in (0, 1):
        for seed in (0, 1, 2):
            st0 = cases.four_quadrant_2d(8, 8, seed=seed)
            st, dts, tf = oracle_py.run(2, [8, 8], [0, 0], flux, 0.6, 0.2, 100000, st0, which="ref")
            runs.append({"name": f"quadrant2d_seed{seed}_flux{flux}", "dim": 2, "nb": [8, 8, 1],
                         "bc": [0, 0, 0], "flux": flux, "cfl": 0.6, "t_end": 0.2,
                         "max_steps": 100000, "ic": {"kind": "four_quadrant_2d", "seed": seed},
                         "steps": int(len(dts)), "dts": dts.tolist(), "t_final": tf,
                         "sha256": sha(st), "l1": np.abs(st).sum(axis=(1, 2, 3)).tolist()})
            print(runs[-1]["name"], runs[-1]["steps"], flush=True)
    # 3D periodic blob + Mach-2 shock, 32^3, CFL 0.5, 100 steps.
    for flux in (1, 0):
        st0 = cases.blobs_shock_3d(2, seed=1)
        st, dts, tf = oracle_py.run(3, [2, 2, 2], [1, 1, 1], flux, 0.5, 1e300, 100, st0,
                                    which="ref")
        runs.append({"name": f"blobs3d_seed1_flux{flux}", "dim": 3, "nb": [2, 2, 2],
                     "bc": [1, 1, 1], "flux": flux, "cfl": 0.5, "t_end": 1e300,
                     "max_steps": 100, "ic": {"kind": "blobs_shock_3d", "seed": 1},
                     "steps": int(len(dts)), "dts": dts.tolist(), "t_final": tf,
                     "sha256": sha(st), "l1": np.abs(st).sum(axis=(1, 2, 3)).tolist()})
        print(runs[-1]["name"], runs[-1]["steps"], flush=True)
    # 2D Sod along x on [0,1]x[0,0.25] (64x16 cells, 4x1 blocks), t_end 0.2: the
    # reference testsuite's DetailedSod level-0 case.
    for flux in (0, 1):
        st0 = sod_state()
        st, dts, tf = oracle_py.run(2, [4, 1], [0, 0], flux, 0.6, 0.2, 100000, st0,
                                    cell_size=1.0 / 64, which="ref")
        runs.append({"name": f"sodx2d_flux{flux}", "dim": 2, "nb": [4, 1, 1],
                     "bc": [0, 0, 0], "flux": flux, "cfl": 0.6, "t_end": 0.2,
                     "max_steps": 100000, "cell_size": 1.0 / 64, "ic": {"kind": "sod_x"},
                     "steps": int(len(dts)), "dts": dts.tolist(), "t_final": tf,
                     "sha256": sha(st), "l1": np.abs(st).sum(axis=(1, 2, 3)).tolist()})
    # Symmetry sides (SymmetryBoundaryCondition): 2D four-quadrant with
    # symmetry in x and y, and a 3D case with symmetry / zero-gradient /
    # periodic in x / y / z.
    for flux in (0, 1):
        st0 = cases.four_quadrant_2d(6, 4, seed=7)
        st, dts, tf = oracle_py.run(2, [6, 4], [2, 2], flux, 0.6, 0.2, 100000, st0, which="ref")
        runs.append({"name": f"quadrant2d_symmetry_seed7_flux{flux}", "dim": 2, "nb": [6, 4, 1],
                     "bc": [2, 2, 0], "flux": flux, "cfl": 0.6, "t_end": 0.2,
                     "max_steps": 100000, "ic": {"kind": "four_quadrant_2d", "seed": 7},
                     "steps": int(len(dts)), "dts": dts.tolist(), "t_final": tf,
                     "sha256": sha(st), "l1": np.abs(st).sum(axis=(1, 2, 3)).tolist()})
        st0 = cases.blobs_shock_3d(2, seed=5)
        st, dts, tf = oracle_py.run(3, [2, 2, 2], [2, 0, 1], flux, 0.5, 1e300, 60, st0,
                                    which="ref")
        runs.append({"name": f"blobs3d_mixedbc_seed5_flux{flux}", "dim": 3, "nb": [2, 2, 2],
                     "bc": [2, 0, 1], "flux": flux, "cfl": 0.5, "t_end": 1e300,
                     "max_steps": 60, "ic": {"kind": "blobs_shock_3d", "seed": 5},
                     "steps": int(len(dts)), "dts": dts.tolist(), "t_final": tf,
                     "sha256": sha(st), "l1": np.abs(st).sum(axis=(1, 2, 3)).tolist()})
        print(runs[-2]["name"], runs[-2]["steps"], runs[-1]["name"], flush=True)
    meta["runs"] = runs
    with open(os.path.join(HERE, "golden_runs.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


def sod_state():
    """Sod shock tube along x: (1, 0, 1) for x < 0.5, (0.125, 0, 0.1) else."""
    nx, ny = 64, 16
    x = (np.arange(nx) + 0.5) / 64
    X = np.repeat(x[:, None], ny, axis=1)
    left = X < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    return cases.primes_to_conservatives(2, GAMMA, 0.0, rho, (z, z), p)[..., None]


FULL_NB = 16
FULL_STEPS = 100
FULL_SAMPLES = 4096


def full_sample_index(n=FULL_NB * 16, count=FULL_SAMPLES):
    """Fixed cells (i, j, k) of the 256^3 domain where the final state is kept."""
    rng = np.random.default_rng(256)
    return rng.integers(0, n, size=(count, 3))


def main_full(threads, fluxes):
    path_json = os.path.join(HERE, "golden_full_256.json")
    path_npz = os.path.join(HERE, "golden_full_256.npz")
    meta = json.load(open(path_json)) if os.path.exists(path_json) else {"runs": {}}
    arrays = dict(np.load(path_npz)) if os.path.exists(path_npz) else {}
    idx = full_sample_index()
    arrays["sample_index"] = idx
    nb = FULL_NB
    for flux in fluxes:
        st0 = cases.blobs_shock_3d(nb, seed=0)
        t0 = time.time()
        st, dts, tf = oracle_py.run(3, [nb] * 3, [1, 1, 1], flux, 0.5, 1e300, FULL_STEPS, st0,
                                    threads=threads, which="ref")
        wall = time.time() - t0
        name = "roe" if flux == 1 else "hllc"
        meta["runs"][name] = {
            "dim": 3, "nb": [nb] * 3, "bc": [1, 1, 1], "flux": flux, "cfl": 0.5,
            "steps": int(len(dts)), "ic": {"kind": "blobs_shock_3d", "seed": 0},
            "ic_sha256": sha(st0), "sha256": sha(st), "dts": dts.tolist(), "t_final": tf,
            "sums0": st0.sum(axis=(1, 2, 3)).tolist(), "sums": st.sum(axis=(1, 2, 3)).tolist(),
            "reference_wall_s": wall, "threads": threads}
        arrays[f"{name}_final_sample"] = st[:, idx[:, 0], idx[:, 1], idx[:, 2]].T.copy()
        print(f"{name}: {len(dts)} steps, t={tf:.6e}, {wall:.0f} s on {threads} threads",
              flush=True)
        with open(path_json, "w") as fh:
            json.dump(meta, fh, indent=1)
        np.savez_compressed(path_npz, **arrays)


def main_edge():
    """edge_blocks.json: per dimension, flux family and edge kind (see
    tests/edge_states.py) the SHA-256 of the reference's per-block face
    fluxes and right-hand side, NaNs made canonical."""
    sys.path.insert(0, os.path.dirname(HERE))
    import edge_states

    meta = {}
    with np.errstate(all="ignore"):
        for dim in (2, 3):
            for flux in (0, 1):
                for kind in edge_states.KINDS:
                    blk = edge_states.edge_block(dim, kind, seed=3)
                    fl = oracle_py.block_fluxes(dim, flux, GAMMA, 0.0, blk, 1.0 / 128, which="ref")
                    rhs = oracle_py.block_rhs(dim, flux, GAMMA, 0.0, blk, 1.0 / 128, which="ref")
                    meta[f"{dim}d_{'roe' if flux else 'hllc'}_{kind}"] = {
                        "fluxes_sha256": hashlib.sha256(edge_states.canonical(fl)).hexdigest(),
                        "rhs_sha256": hashlib.sha256(edge_states.canonical(rhs)).hexdigest(),
                        "nan_fluxes": int(np.isnan(fl).sum())}
    with open(os.path.join(HERE, "edge_blocks.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


ID_SHAPES_FULL = ([1, 1, 1], [4, 1, 1], [3, 5, 1], [2, 3, 4], [5, 4, 3])
ID_SHAPES_HASHED = ([16, 16, 16], [128, 2, 1], [1, 128, 2], [2, 1, 128], [64, 64, 1])


def main_ids():
    """level0_ids.json: the reference's level-0 node ids (TopologyManager's
    creation loop over the id arithmetic of id_information.cpp) in creation
    order, in full for small grids and as SHA-256 of the uint64 array for
    large ones."""
    meta = {"full": {}, "sha256": {}}
    for nb in ID_SHAPES_FULL:
        meta["full"][",".join(map(str, nb))] = [int(v) for v in oracle_py.ref_level0_ids(nb)]
    for nb in ID_SHAPES_HASHED:
        ids = oracle_py.ref_level0_ids(nb)
        meta["sha256"][",".join(map(str, nb))] = hashlib.sha256(
            np.ascontiguousarray(ids, dtype=np.uint64).tobytes()).hexdigest()
    with open(os.path.join(HERE, "level0_ids.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


# Halo range (start, size) integrated by the IntegrateJumpHalos vectors: the
# +y halo slab over the internal x (and z) cells.
LTS_HALO = {2: ((4, 20, 0), (16, 4, 1)), 3: ((4, 20, 4), (16, 4, 16))}


def main_mr():
    """mr_golden.json: per dimension, pair kind (tests/mr_states.py) and
    sibling position, the reference's Multiresolution::Average output (the
    parent block after averaging), Prediction output per range of
    mr_states.ranges (every field, child cells outside the range untouched)
    as SHA-256 with NaNs canonical, and the L-infinity, L-one and L-two
    details of ChildNeedsRemeshing (read off its decisions,
    oracle_py.ref_mr_detail) as hex floats; the reference's AverageJumpBuffer
    on seeded jump surfaces (seed 40 + dim, child then parent) for every
    sibling position; the reference's RungeKutta3TVD on one Node
    (PrepareBufferForIntegration + IntegrateNode, IntegrateJumpHalos) with
    four of eight seeded micro time steps per stage; plus ThresholdOnLevel
    for a few settings."""
    sys.path.insert(0, os.path.dirname(HERE))
    import edge_states
    import mr_states

    meta = {"pairs": {}, "thresholds": []}
    with np.errstate(all="ignore"):
        for dim in (2, 3):
            for kind in mr_states.KINDS:
                for pos in range(2 ** dim):
                    parent, child = mr_states.pair(dim, kind, pos)
                    avg = oracle_py.mr_average(dim, child, parent, pos, which="ref")
                    preds = []
                    for rng in mr_states.ranges(dim):
                        pred = np.stack([oracle_py.mr_prediction(dim, parent[e], child[e], pos, rng,
                                                                 which="ref")
                                         for e in range(dim + 2)])
                        preds.append(hashlib.sha256(edge_states.canonical(pred)).hexdigest())
                    details = [float(oracle_py.ref_mr_detail(dim, parent, child, pos, norm)).hex()
                               for norm in (oracle_py.MR_LINF, oracle_py.MR_L1, oracle_py.MR_L2)]
                    meta["pairs"][f"{dim}d_{kind}_{pos}"] = {
                        "average_sha256": hashlib.sha256(edge_states.canonical(avg)).hexdigest(),
                        "prediction_sha256": preds,
                        "detail": details[0],
                        "detail_l1": details[1],
                        "detail_l2": details[2]}
        meta["jump"] = {}
        for dim in (2, 3):
            icz = 16 if dim == 3 else 1
            rng = np.random.default_rng(40 + dim)
            for pos in range(2 ** dim):
                child = rng.standard_normal((6, dim + 2, 16, icz))
                parent = rng.standard_normal((6, dim + 2, 16, icz))
                out = oracle_py.mr_average_jump(dim, child, parent, pos, which="ref")
                meta["jump"][f"{dim}d_{pos}"] = hashlib.sha256(out.tobytes()).hexdigest()
        meta["lts"] = {}
        for dim in (2, 3):
            tz, icz = (24, 16) if dim == 3 else (1, 1)
            for op in (0, 1):
                for stage in range(3):
                    rng = np.random.default_rng(1000 * dim + 10 * op + stage)
                    micro = rng.uniform(1e-4, 1e-3, 8)
                    bufs = [rng.standard_normal((dim + 2, 24, 24, tz)) for _ in range(3)]
                    bufs += [rng.standard_normal((6, dim + 2, 16, icz)) for _ in range(2)]
                    start, size = LTS_HALO[dim]
                    out = oracle_py.ref_lts_integrate(dim, op, stage, 4, micro, start, size, *bufs)
                    meta["lts"][f"{dim}d_op{op}_stage{stage}"] = hashlib.sha256(
                        b"".join(o.tobytes() for o in out)).hexdigest()
        for dim in (2, 3):
            for ml, rl, eps in ((1, 1, 1e-3), (4, 2, 1e-3), (6, 6, 0.01), (3, 0, 0.5)):
                for level in range(ml + 1):
                    meta["thresholds"].append(
                        [dim, ml, rl, eps, level,
                         float(oracle_py.mr_threshold(dim, ml, rl, eps, level, which="ref")).hex()])
    with open(os.path.join(HERE, "mr_golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    ap.add_argument("--flux", choices=["roe", "hllc", "both"], default="both")
    ap.add_argument("--ids", action="store_true", help="write level0_ids.json only")
    ap.add_argument("--edge", action="store_true", help="write edge_blocks.json only")
    ap.add_argument("--mr", action="store_true", help="write mr_golden.json only")
    a = ap.parse_args()
    if a.mr:
        main_mr()
    elif a.ids:
        main_ids()
    elif a.edge:
        main_edge()
    elif a.full:
        main_full(a.threads, {"roe": [1], "hllc": [0], "both": [1, 0]}[a.flux])
    else:
        main()
