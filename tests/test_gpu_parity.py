"""GPU parity: the HIP path against the CPU oracle through the C ABI.

FP64 is compared bit for bit where the reference arithmetic is reproduced
exactly (per-block fluxes and right-hand sides, trajectories); the stated
bound of the FP64 path is max relative error <= 1e-10 after 100 steps.
"""
import ctypes

import numpy as np
import pytest

from alpaca_amd import cases, gpu
import oracle_py

pytestmark = pytest.mark.gpu

REL_TOL = 1e-10  # FP64 GPU vs CPU oracle, max relative error


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _maxrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_block_fluxes_and_rhs_bitexact(torch, dim, flux):
    print(f"[block parity] dim={dim} flux={flux}", flush=True)
    lib = gpu.load_library()
    blocks = [cases.random_block(dim, s, smooth=(s % 2 == 1)) for s in range(4)]
    host = np.stack(blocks)
    dev = torch.as_tensor(host, device="cuda").contiguous()
    desc = gpu.make_desc(dim, [1, 1, 1], [0, 0, 0], flux=flux, cell_size=1.0 / 128)
    nfl = oracle_py.nflux(dim)
    out = torch.zeros(len(blocks) * nfl, dtype=torch.float64, device="cuda")
    s = gpu._stream_ptr()
    gpu.check(lib.alpaca_convective_fluxes(ctypes.byref(desc), len(blocks),
                                           ctypes.c_void_p(dev.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), s), "fluxes")
    rhs = torch.zeros_like(dev)
    gpu.check(lib.alpaca_right_hand_side(ctypes.byref(desc), len(blocks),
                                         ctypes.c_void_p(dev.data_ptr()),
                                         ctypes.c_void_p(rhs.data_ptr()), s), "rhs")
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(len(blocks), -1)
    got_rhs = gpu.from_blocks(rhs, dim, [len(blocks), 1, 1]).cpu().numpy()
    icz = 16 if dim == 3 else 1
    for bi, blk in enumerate(blocks):
        want = oracle_py.block_fluxes(dim, flux, 1.4, 0.0, blk, 1.0 / 128).ravel()
        assert np.array_equal(got[bi], want), f"block {bi} flux max diff {np.max(np.abs(got[bi]-want))}"
        want_rhs = oracle_py.block_rhs(dim, flux, 1.4, 0.0, blk, 1.0 / 128)
        g = got_rhs[:, bi * 16:(bi + 1) * 16, :, :icz]
        assert np.array_equal(g, want_rhs), f"block {bi} rhs max diff {np.max(np.abs(g-want_rhs))}"


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_four_quadrant_2d_to_t_end(torch, flux):
    """128^2 four-quadrant problem, zero-gradient sides, CFL 0.6, t_end 0.2."""
    print(f"[2d trajectory] flux={flux}", flush=True)
    nb = [8, 8]
    state0 = cases.four_quadrant_2d(8, 8, seed=0)
    desc = gpu.make_desc(2, nb, [0, 0], flux=flux)
    case = gpu.GpuCase(desc, state0, cfl=0.6, t_end=0.2)
    steps, dts = case.run(max_steps=10000)
    got = case.download()
    want, want_dts, t_final = oracle_py.run(2, nb, [0, 0], flux, 0.6, 0.2, 10000, state0)
    assert steps == len(want_dts)
    assert np.array_equal(np.array(dts), want_dts)
    assert case.record().time == t_final == 0.2
    err = _maxrel(got, want)
    print(f"  steps={steps} max rel err={err:.3e} bitexact={np.array_equal(got, want)}")
    assert err <= REL_TOL


@pytest.mark.parametrize("flux", [gpu.FLUX_ROE, gpu.FLUX_HLLC])
def test_blobs_shock_3d_periodic_100_steps(torch, flux):
    """32^3 periodic blob + Mach-2 shock, CFL 0.5, 100 steps vs the oracle."""
    print(f"[3d trajectory] flux={flux}", flush=True)
    nb = [2, 2, 2]
    state0 = cases.blobs_shock_3d(2, seed=1)
    desc = gpu.make_desc(3, nb, [1, 1, 1], flux=flux)
    case = gpu.GpuCase(desc, state0, cfl=0.5)
    case.advance(100)
    got = case.download()
    want, want_dts, _ = oracle_py.run(3, nb, [1, 1, 1], flux, 0.5, 1e300, 100, state0)
    rec = case.record()
    assert rec.steps == 100 and rec.flags == 0
    err = _maxrel(got, want)
    print(f"  max rel err={err:.3e} bitexact={np.array_equal(got, want)}")
    assert err <= REL_TOL
