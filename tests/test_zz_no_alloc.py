"""alpaca_right_hand_side_noalloc: the right-hand side without allocation.

Its face fluxes live in the library's fixed per-GPU scratch array.  Two calls
in a row (the second on fewer, different blocks) must each give
alpaca_right_hand_side's bits, which the oracle's per-block right-hand side
pins; calls up to the block limit leave the device's free memory unchanged,
and a call above the limit is refused before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from alpaca_amd import cases, gpu
import oracle_py

pytestmark = pytest.mark.gpu

DX = 1.0 / 128
MAX_BLOCKS = {2: 3855, 3: 128}


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _blocks(torch, dim, n, seed0):
    host = np.stack([cases.random_block(dim, seed0 + s, smooth=(s % 2 == 1)) for s in range(n)])
    return host, torch.as_tensor(host, device="cuda").contiguous()


def _rhs_of(rhs, dim, n):
    icz = 16 if dim == 3 else 1
    g = gpu.from_blocks(rhs, dim, [n, 1, 1]).cpu().numpy()
    return [g[:, i * 16:(i + 1) * 16, :, :icz] for i in range(n)]


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.MIXED])
@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_equals_rhs_on_two_calls(torch, dim, flux, precision):
    print(f"[noalloc] dim={dim} flux={flux} precision={precision}", flush=True)
    lib = gpu.load_library()
    desc = gpu.make_desc(dim, [1, 1, 1], [0, 0, 0], flux=flux, precision=precision, cell_size=DX)
    d = ctypes.byref(desc)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n, seed in ((6, 40 + dim), (3, 90 + dim)):  # the second call: fewer, other blocks
        host, dev = _blocks(torch, dim, n, seed)
        got = torch.full_like(dev, np.nan)
        want = torch.full_like(dev, np.nan)
        gpu.check(lib.alpaca_right_hand_side_noalloc(d, n, P(dev), P(got), sp), "noalloc")
        gpu.check(lib.alpaca_right_hand_side(d, n, P(dev), P(want), sp), "rhs")
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"n={n}: noalloc differs from alpaca_right_hand_side"
        if precision == gpu.FP64:
            ref = [oracle_py.block_rhs(dim, flux, 1.4, 0.0, b, DX) for b in host]
            for b, (g, w) in enumerate(zip(_rhs_of(got, dim, n), ref)):
                assert np.array_equal(g, w), f"n={n} block {b} vs oracle"


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_limit_and_free_memory(torch, dim):
    """At the block limit the call runs and the device's free memory stays
    the same; one block more is refused with ALPACA_ERR_CAPACITY."""
    print(f"[noalloc limit] dim={dim}", flush=True)
    lib = gpu.load_library()
    desc = gpu.make_desc(dim, [1, 1, 1], [0, 0, 0], cell_size=DX)
    d = ctypes.byref(desc)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = MAX_BLOCKS[dim]
    one, _ = _blocks(torch, dim, 1, 5)
    src = torch.as_tensor(np.repeat(one, n + 1, axis=0), device="cuda").contiguous()
    out = torch.zeros_like(src)
    gpu.check(lib.alpaca_right_hand_side_noalloc(d, 1, P(src), P(out), sp), "warm-up")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    gpu.check(lib.alpaca_right_hand_side_noalloc(d, n, P(src), P(out), sp), "at the limit")
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    want = torch.zeros_like(src)
    gpu.check(lib.alpaca_right_hand_side(d, n, P(src), P(want), sp), "rhs")
    torch.cuda.synchronize()
    assert torch.equal(out[:n], want[:n])
    assert lib.alpaca_right_hand_side_noalloc(d, n + 1, P(src), P(out), sp) == 5
