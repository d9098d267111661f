"""The batched workload at its stated size: 512 independent 2D four-quadrant
cases (32^2..256^2 cells, weights .4/.3/.2/.1, gamma ~ U[1.2, 1.67], HLLC,
CFL 0.6, t_end 0.2) in one ensemble, FP64 and mixed.

* FP64 cases equal the CPU oracle bit for bit (one case of every size and a
  further sample of small ones);
* a case of every size gives the same bits alone as in the 512-case mix;
* every mixed case stays within L1 density error 1e-6 of its FP64 twin.
"""
import numpy as np
import pytest

from alpaca_amd import cases, gpu
from alpaca_amd.ensemble import CASE_DONE, CASE_STOP, Ensemble
import oracle_py

pytestmark = pytest.mark.gpu

T_END, CFL = 0.2, 0.6
MIXED_L1_BOUND = 1e-6


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


@pytest.fixture(scope="module")
def specs():
    return cases.batched_cases(512, seed=0)


def _desc(nb, gamma, precision):
    return gpu.make_desc(2, [nb, nb], [0, 0], flux=gpu.FLUX_HLLC, precision=precision,
                         gamma=gamma)


def _run_all(specs, precision):
    total = sum(nb * nb for nb, _, _ in specs)
    ens = Ensemble(2, max_blocks=total, max_cases=len(specs),
                   classes=[(gpu.FLUX_HLLC, precision)])
    ids = [ens.add_case(_desc(nb, g, precision), s0, T_END, CFL) for nb, g, s0 in specs]
    ens.run_until_done(ids, poll=16)
    st = ens.states()
    out = []
    for cid in ids:
        assert st[cid].flags & CASE_DONE and not (st[cid].flags & (CASE_STOP & ~CASE_DONE))
        assert st[cid].time == T_END
        out.append((ens.download(cid), st[cid].steps))
    ens.close()
    return out


@pytest.fixture(scope="module")
def results(torch, specs):
    print("[batched 512] running FP64 and mixed ensembles", flush=True)
    return {p: _run_all(specs, p) for p in (gpu.FP64, gpu.MIXED)}


def _first_of_each_size(specs):
    seen = {}
    for i, (nb, _, _) in enumerate(specs):
        seen.setdefault(nb, i)
    return sorted(seen.values())


def test_fp64_cases_match_oracle(specs, results):
    small = [i for i, (nb, _, _) in enumerate(specs) if nb == 2][:8]
    picks = sorted(set(_first_of_each_size(specs) + small))
    for i in picks:
        nb, g, s0 = specs[i]
        print(f"  oracle case {i}: {nb * 16}^2 gamma {g:.3f}", flush=True)
        want, dts, tf = oracle_py.run(2, [nb, nb], [0, 0], 0, CFL, T_END, 100000, s0, gamma=g,
                                      threads=16)
        got, steps = results[gpu.FP64][i]
        assert steps == len(dts) and tf == T_END
        assert np.array_equal(got, want), f"case {i}: max diff {np.abs(got - want).max()}"


def test_alone_equals_mix(specs, results):
    for prec in (gpu.FP64, gpu.MIXED):
        for i in _first_of_each_size(specs):
            nb, g, s0 = specs[i]
            ens = Ensemble(2, max_blocks=nb * nb, max_cases=1, classes=[(gpu.FLUX_HLLC, prec)])
            cid = ens.add_case(_desc(nb, g, prec), s0, T_END, CFL)
            ens.run_until_done([cid], poll=16)
            alone = ens.download(cid)
            ens.close()
            assert np.array_equal(alone, results[prec][i][0]), f"precision {prec} case {i}"


def test_mixed_within_bound_of_fp64(specs, results):
    errs = []
    for (m, _), (f, _) in zip(results[gpu.MIXED], results[gpu.FP64]):
        errs.append(float(np.abs(m[0] - f[0]).sum() / np.abs(f[0]).sum()))
    errs = np.array(errs)
    worst = int(errs.argmax())
    print(f"  mixed L1 rho vs FP64 over 512 cases: mean {errs.mean():.3e}, worst {errs.max():.3e} "
          f"(case {worst}, {specs[worst][0] * 16}^2)", flush=True)
    assert errs.max() <= MIXED_L1_BOUND
