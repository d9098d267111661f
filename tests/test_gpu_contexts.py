"""Scratch ownership of the per-block entries: calls on different streams or
contexts never share face-flux scratch, growth needs no synchronisation, and
every form gives the same bits as the oracle's per-block right-hand side."""
import ctypes

import numpy as np
import pytest

from alpaca_amd import cases, gpu
import oracle_py

pytestmark = pytest.mark.gpu

DX = 1.0 / 128


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _blocks(torch, dim, n, seed0):
    host = np.stack([cases.random_block(dim, seed0 + s, smooth=(s % 2 == 1)) for s in range(n)])
    return host, torch.as_tensor(host, device="cuda").contiguous()


def _rhs_oracle(dim, flux, host):
    return [oracle_py.block_rhs(dim, flux, 1.4, 0.0, b, DX) for b in host]


def _rhs_of(rhs, dim, n):
    icz = 16 if dim == 3 else 1
    g = gpu.from_blocks(rhs, dim, [n, 1, 1]).cpu().numpy()
    return [g[:, i * 16:(i + 1) * 16, :, :icz] for i in range(n)]


@pytest.mark.parametrize("dim", [2, 3])
def test_two_streams_concurrent_rhs_bitexact(torch, dim):
    """alpaca_right_hand_side on two streams at once, with different block
    counts (so each stream's scratch grows while the other's kernels run),
    equals the oracle block by block."""
    print(f"[two streams] dim={dim}", flush=True)
    lib = gpu.load_library()
    flux = gpu.FLUX_HLLC
    desc = gpu.make_desc(dim, [1, 1, 1], [0, 0, 0], flux=flux, cell_size=DX)
    sizes = {0: [2, 6, 3], 1: [5, 1, 7]}
    inputs = {k: [_blocks(torch, dim, n, 100 * k + 10 * i) for i, n in enumerate(v)]
              for k, v in sizes.items()}
    outs = {k: [torch.zeros_like(d) for _, d in v] for k, v in inputs.items()}
    streams = {k: torch.cuda.Stream() for k in sizes}
    torch.cuda.synchronize()
    for i in range(3):
        for k in (0, 1):
            host, dev = inputs[k][i]
            sp = ctypes.c_void_p(streams[k].cuda_stream)
            gpu.check(lib.alpaca_right_hand_side(ctypes.byref(desc), len(host),
                                                 ctypes.c_void_p(dev.data_ptr()),
                                                 ctypes.c_void_p(outs[k][i].data_ptr()), sp),
                      "rhs")
    torch.cuda.synchronize()
    for k in (0, 1):
        for i in range(3):
            host, _ = inputs[k][i]
            got = _rhs_of(outs[k][i], dim, len(host))
            for b, (g, w) in enumerate(zip(got, _rhs_oracle(dim, flux, host))):
                assert np.array_equal(g, w), f"stream {k} call {i} block {b}"


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_explicit_context_forms_equal_stream_forms(torch, flux):
    """The alpaca_ctx_* forms give the stream forms' bits (fluxes, RHS, RHS
    with the jump-flux save), also after a reserve and a growth in between."""
    print(f"[contexts] flux={flux}", flush=True)
    lib = gpu.load_library()
    dim = 3
    desc = gpu.make_desc(dim, [1, 1, 1], [0, 0, 0], flux=flux, cell_size=DX)
    s = torch.cuda.Stream()
    ctx = ctypes.c_void_p()
    with torch.cuda.stream(s):
        gpu.check(lib.alpaca_context_create(ctypes.c_void_p(s.cuda_stream), ctypes.byref(ctx)),
                  "context_create")
        gpu.check(lib.alpaca_context_reserve(ctx, dim, 2), "reserve")
        try:
            for n, seed in ((2, 7), (5, 11)):  # the second call outgrows the reserve
                host, dev = _blocks(torch, dim, n, seed)
                nfl = oracle_py.nflux(dim)
                jump_n = n * 6 * gpu.NE(dim) * 16 * 16
                f_ctx = torch.zeros(n * nfl, dtype=torch.float64, device="cuda")
                f_str = torch.zeros_like(f_ctx)
                r_ctx, r_str, r_jmp = (torch.zeros_like(dev) for _ in range(3))
                j_ctx = torch.zeros(jump_n, dtype=torch.float64, device="cuda")
                j_str = torch.zeros_like(j_ctx)
                d = ctypes.byref(desc)
                pa = ctypes.c_void_p(dev.data_ptr())
                sp = ctypes.c_void_p(s.cuda_stream)
                P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
                gpu.check(lib.alpaca_ctx_convective_fluxes(ctx, d, n, pa, P(f_ctx)), "ctx fluxes")
                gpu.check(lib.alpaca_convective_fluxes(d, n, pa, P(f_str), sp), "fluxes")
                gpu.check(lib.alpaca_ctx_right_hand_side(ctx, d, n, pa, P(r_ctx)), "ctx rhs")
                gpu.check(lib.alpaca_right_hand_side(d, n, pa, P(r_str), sp), "rhs")
                gpu.check(lib.alpaca_ctx_right_hand_side_jump_fluxes(ctx, d, n, pa, P(r_jmp),
                                                                     P(j_ctx)), "ctx jump")
                gpu.check(lib.alpaca_right_hand_side_jump_fluxes(d, n, pa, P(r_str), P(j_str),
                                                                 sp), "jump")
                s.synchronize()
                assert torch.equal(f_ctx, f_str)
                assert torch.equal(r_ctx, r_str) and torch.equal(r_jmp, r_str)
                assert torch.equal(j_ctx, j_str)
                want = _rhs_oracle(dim, flux, host)
                for b, (g, w) in enumerate(zip(_rhs_of(r_ctx, dim, n), want)):
                    assert np.array_equal(g, w), f"n={n} block {b}"
        finally:
            gpu.check(lib.alpaca_context_destroy(ctx), "context_destroy")
