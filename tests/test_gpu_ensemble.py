"""Ensembles of independent cases: every case's result is the same alone and
in the mix (bit for bit), including cases admitted while others run, and
matches the CPU oracle."""
import numpy as np
import pytest

from alpaca_amd import cases, gpu
from alpaca_amd.ensemble import CASE_CANCELLED, CASE_DONE, CASE_STOP, Ensemble
import oracle_py

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _mix_2d(precision):
    """Uneven 2D four-quadrant cases (sizes, gamma, t_end, sides differ)."""
    specs = []
    rng = np.random.default_rng(3)
    for i, (nbx, nby) in enumerate([(2, 2), (4, 4), (3, 2), (8, 8), (2, 5), (4, 2)]):
        gamma = float(rng.uniform(1.2, 1.67))
        bc = [int(rng.integers(0, 2)), int(rng.integers(0, 2))]
        t_end = float(rng.uniform(0.05, 0.15))
        st0 = cases.four_quadrant_2d(nbx, nby, seed=100 + i, gamma=gamma)
        desc = gpu.make_desc(2, [nbx, nby], bc, flux=gpu.FLUX_HLLC, precision=precision,
                             gamma=gamma)
        specs.append((desc, st0, t_end))
    return specs


def _alone(desc, st0, t_end, tile=0):
    """The case run by itself: a one-case ensemble (same kernels)."""
    ens = Ensemble(desc.dim, max_blocks=64, max_cases=1,
                   classes=[(desc.flux, desc.precision)])
    ens.set_tuning(tile)
    cid = ens.add_case(desc, st0, t_end, 0.6)
    ens.run_until_done([cid], poll=1)
    rec = ens.states()[cid]
    out = ens.download(cid)
    ens.close()
    return out, rec.steps, rec.time


def _single_case_path(desc, st0, t_end):
    case = gpu.GpuCase(desc, st0, cfl=0.6, t_end=t_end)
    steps, _ = case.run(max_steps=100000)
    return case.download(), steps, case.record().time


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.MIXED])
@pytest.mark.parametrize("tile", [0, 1, 2])
def test_ensemble_alone_equals_mix_2d(torch, precision, tile):
    print(f"[ensemble 2d] precision={precision} tile={tile}", flush=True)
    specs = _mix_2d(precision)
    ens = Ensemble(2, max_blocks=200, max_cases=16, classes=[(gpu.FLUX_HLLC, precision)])
    ens.set_tuning(tile)
    ids = [ens.add_case(d, s, t, 0.6) for d, s, t in specs]
    ens.run_until_done(ids, poll=4)
    st = ens.states()
    for cid, (d, s0, t_end) in zip(ids, specs):
        assert st[cid].flags & CASE_DONE and not (st[cid].flags & (CASE_STOP & ~CASE_DONE))
        got = ens.download(cid)
        want, steps, t = _alone(d, s0, t_end, tile)
        assert st[cid].steps == steps and st[cid].time == t == t_end
        assert np.array_equal(got, want), f"case {cid}"
        if precision == gpu.FP64:
            single, steps1, _ = _single_case_path(d, s0, t_end)
            assert steps1 == steps and np.array_equal(got, single), f"case {cid}"
    # FP64 also equals the CPU oracle.
    if precision == gpu.FP64:
        d, s0, t_end = specs[1]
        want, _, _ = oracle_py.run(2, [d.nb[0], d.nb[1]], [d.bc[0], d.bc[1]], 0, 0.6, t_end,
                                   100000, s0, gamma=d.gamma)
        assert np.array_equal(ens.download(ids[1]), want)
    ens.close()


def test_ensemble_continuous_admission_and_cancel(torch):
    """Cases admitted while others run, removed when done, one cancelled: the
    finished ones equal their alone results bit for bit."""
    print("[ensemble admission]", flush=True)
    specs = _mix_2d(gpu.FP64)
    ens = Ensemble(2, max_blocks=90, max_cases=4, classes=[(gpu.FLUX_HLLC, gpu.FP64)])
    pending = list(range(len(specs)))
    live = {}
    results = {}
    cancelled = None
    guard = 0
    while (pending or live) and guard < 20000:
        guard += 1
        # Admit while capacity allows (continuous batching).
        while pending:
            i = pending[0]
            d, s0, t_end = specs[i]
            try:
                cid = ens.add_case(d, s0, t_end, 0.6)
            except MemoryError:
                break
            live[cid] = i
            pending.pop(0)
            if cancelled is None and i == 2:
                ens.cancel_case(cid)
                cancelled = i
        ens.step(3)
        st = ens.states()
        for cid in list(live):
            if st[cid].flags & CASE_STOP:
                i = live.pop(cid)
                results[i] = (ens.download(cid), st[cid])
                ens.remove_case(cid)
    assert not pending and not live
    for i, (d, s0, t_end) in enumerate(specs):
        got, rec = results[i]
        if i == cancelled:
            assert rec.flags & CASE_CANCELLED and rec.steps == 0
            continue
        want, steps, _ = _alone(d, s0, t_end)
        assert rec.steps == steps
        assert np.array_equal(got, want), f"case {i}"
    ens.close()


@pytest.mark.parametrize("tile", [0, 1])
def test_ensemble_3d_mixed_classes(torch, tile):
    """3D cases of both flux families and all three arithmetic modes in one
    pool: each equals its case run alone; the FP64 ones equal the single-case
    kernel bit for bit, the FP64-with-FMA one is within 1e-10 of it."""
    print(f"[ensemble 3d] tile={tile}", flush=True)
    st0 = cases.blobs_shock_3d(2, seed=8)
    combos = [(gpu.FLUX_HLLC, gpu.FP64), (gpu.FLUX_ROE, gpu.FP64), (gpu.FLUX_HLLC, gpu.MIXED),
              (gpu.FLUX_HLLC, gpu.FP64_FMA)]
    ens = Ensemble(3, max_blocks=40, max_cases=4, classes=combos)
    ens.set_tuning(tile)
    ids = []
    for flux, prec in combos:
        d = gpu.make_desc(3, [2, 2, 2], [1, 1, 1], flux=flux, precision=prec)
        ids.append((ens.add_case(d, st0, 1e300, 0.5), d))
    ens.step(12)
    for cid, d in ids:
        solo = Ensemble(3, max_blocks=8, max_cases=1, classes=[(d.flux, d.precision)])
        solo.set_tuning(tile)
        sid = solo.add_case(d, st0, 1e300, 0.5)
        solo.step(12)
        assert np.array_equal(ens.download(cid), solo.download(sid))
        if d.precision == gpu.FP64:
            single = gpu.GpuCase(d, st0, cfl=0.5)
            single.advance(12)
            assert np.array_equal(ens.download(cid), single.download())
        if d.precision == gpu.FP64_FMA:
            exact = gpu.make_desc(3, [2, 2, 2], [1, 1, 1], flux=d.flux, precision=gpu.FP64)
            single = gpu.GpuCase(exact, st0, cfl=0.5)
            single.advance(12)
            want = single.download()
            err = float(np.max(np.abs(ens.download(cid) - want)) / np.max(np.abs(want)))
            print(f"  fp64-fma in the pool vs exact single case: {err:.3e}")
            assert err <= 1e-10
        solo.close()
    ens.close()
