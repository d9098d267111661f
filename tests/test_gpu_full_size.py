"""Full-size parity: the 256^3 periodic blob + Mach-2 shock case (16^3 leaf
blocks of 16^3 cells), 100 FP64 steps at CFL 0.5, against the reference's own
Roe flux-splitting and HLLC schemes run on the CPU (fixtures written by
``tests/golden/make_golden.py --full`` from the reference's translation units).

Both FP64 forms must stay within the FP64 bound, 1e-10 maximum relative
error, at 4096 fixed cells and in every time step, and conserve mass,
momentum and energy to 1e-13 relative under the periodic boundaries.  The
exact form (no contraction) must moreover be bit for bit the reference's: the
SHA-256 of all 83.9 M final conservative values and every time step equal the
fixture's.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from alpaca_amd import cases, gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "golden_full_256.json")
SAMPLES = os.path.join(HERE, "golden", "golden_full_256.npz")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def _fixture(name):
    if not os.path.exists(FIXTURE):
        pytest.skip("full-size fixture not generated")
    runs = json.load(open(FIXTURE))["runs"]
    if name not in runs:
        pytest.skip(f"full-size fixture has no {name} run")
    return runs[name], np.load(SAMPLES)


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP64_FMA])
@pytest.mark.parametrize("name", ["roe", "hllc"])
def test_full_256_matches_reference(torch, name, precision):
    """Both FP64 forms: the exact one (the reference's operation order, no
    contraction) and the one with FMA contraction are held to the same bound."""
    run, arrays = _fixture(name)
    print(f"[full 256^3] scheme={name} precision={precision} steps={run['steps']}", flush=True)
    nb = run["nb"][0]
    st0 = cases.blobs_shock_3d(nb, seed=run["ic"]["seed"])
    assert _sha(st0) == run["ic_sha256"], "host initial state differs from the fixture's"
    desc = gpu.make_desc(3, run["nb"], [gpu.BC_PERIODIC] * 3, flux=run["flux"],
                         precision=precision)
    case = gpu.GpuCase(desc, st0, cfl=run["cfl"])
    steps, dts = case.run(run["steps"])
    assert steps == run["steps"]
    want_dts = np.array(run["dts"])
    dt_err = float(np.max(np.abs(np.array(dts) - want_dts) / want_dts))
    print(f"  dt max rel err {dt_err:.3e} identical {np.array_equal(np.array(dts), want_dts)}",
          flush=True)
    assert dt_err <= 1e-10
    got = case.download()
    del case
    idx = arrays["sample_index"]
    sample = got[:, idx[:, 0], idx[:, 1], idx[:, 2]].T
    want = arrays[f"{name}_final_sample"]
    scale = np.maximum(np.abs(want), 1e-300)
    rel = np.abs(sample - want) / scale
    # Momentum components can be ~0 at sampled cells: measure them relative to
    # the field maximum instead.
    field_max = np.abs(got).reshape(5, -1).max(axis=1)
    rel_field = np.abs(sample - want) / field_max[None, :]
    err = float(np.minimum(rel, rel_field).max())
    print(f"  sampled max rel err {err:.3e}", flush=True)
    assert err <= 1e-10
    identical = _sha(got) == run["sha256"]
    print(f"  bit-identical to the reference: {identical}", flush=True)
    if precision == gpu.FP64:
        assert identical, "exact FP64 form is no longer bit-identical to the reference"
        assert np.array_equal(np.array(dts), want_dts)
    # Conservation: relative to sum |q| per field; the momentum components
    # against the total momentum magnitude (the initial y/z momenta are 0).
    s0 = st0.sum(axis=(1, 2, 3))
    s1 = got.sum(axis=(1, 2, 3))
    a0 = np.abs(st0).sum(axis=(1, 2, 3))
    a1 = np.abs(got).sum(axis=(1, 2, 3))
    scale = np.maximum(a0, a1)
    scale[2:] = max(a0[2:].sum(), a1[2:].sum())
    drift = np.abs(s1 - s0) / scale
    print(f"  drift per equation {drift}", flush=True)
    assert np.all(drift <= 1e-13)


def test_device_initial_state_is_host_state(torch):
    """The device-side generator (used by bench.py and serving) yields the
    same bits as the host generator the fixtures were made from."""
    for nb, seed in ((2, 3), (16, 0)):
        host = cases.blobs_shock_3d(nb, seed=seed)
        dev = cases.blobs_shock_3d_torch(nb, seed=seed).cpu().numpy()
        assert np.array_equal(host, dev), f"nb={nb}: max diff {np.abs(host - dev).max()}"


def test_mixed_256_hllc_against_fp64_100_and_1000_steps(torch):
    """The mixed form at its stated size: 256^3 HLLC, 100 and 1000 steps,
    against the exact FP64 path (which the test above pins to the reference):
    L1 density error <= 1e-6, and mass and energy drift <= 1e-13 under the
    periodic sides for both forms."""
    nb = 16
    st0 = cases.blobs_shock_3d_torch(nb, seed=0)
    runs = {}
    for prec in (gpu.FP64, gpu.MIXED):
        desc = gpu.make_desc(3, [nb] * 3, [gpu.BC_PERIODIC] * 3, flux=gpu.FLUX_HLLC,
                             precision=prec)
        runs[prec] = gpu.GpuCase(desc, st0, cfl=0.5)
    g0 = gpu.from_blocks(runs[gpu.FP64].u, 3, [nb] * 3)
    s0 = g0.sum(dim=(1, 2, 3))
    a0 = g0.abs().sum(dim=(1, 2, 3))
    del g0
    done = 0
    for target in (100, 1000):
        for case in runs.values():
            case.advance(target - done)
        done = target
        torch.cuda.synchronize()
        f = gpu.from_blocks(runs[gpu.FP64].u, 3, [nb] * 3)
        m = gpu.from_blocks(runs[gpu.MIXED].u, 3, [nb] * 3)
        l1 = float((m[0] - f[0]).abs().sum() / f[0].abs().sum())
        linf = float((m[0] - f[0]).abs().max() / f[0].abs().max())
        drift = {}
        for name, g in (("fp64", f), ("mixed", m)):
            s = g.sum(dim=(1, 2, 3))
            drift[name] = [float((s[e] - s0[e]).abs() / a0[e]) for e in (0, 1)]
        tf, tm = runs[gpu.FP64].record().time, runs[gpu.MIXED].record().time
        print(f"[mixed 256^3] {target} steps: L1 rho {l1:.3e} Linf rho {linf:.3e} "
              f"t {tf:.6f}/{tm:.6f} drift mass/energy fp64 {drift['fp64']} mixed {drift['mixed']}",
              flush=True)
        del f, m
        assert l1 <= 1e-6
        for name in drift:
            assert max(drift[name]) <= 1e-13, (name, drift[name])
