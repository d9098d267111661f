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


def _face_slab_reference(blocks, dim, nb, bc):
    """Expected face halo slabs of every block (numpy): neighbour's internal
    cells, wrapped on periodic sides, first/last internal layer on
    zero-gradient sides, mirrored internal cells with the normal momentum
    negated on symmetry sides (internal_halo_manager.cpp:439-466,
    boundary_constants.h:169-253, symmetry_boundary_condition.h:89-109)."""
    nbk = blocks.shape[0]
    exp = blocks.copy()
    nb3 = list(nb) + [1] * (3 - len(nb))
    for b in range(nbk):
        c = [b // (nb3[1] * nb3[2]), (b // nb3[2]) % nb3[1], b % nb3[2]]
        for d in range(dim):
            for hi in (0, 1):
                edge = c[d] == (nb3[d] - 1 if hi else 0)
                sl_dst = [slice(4, 20)] * dim
                sl_src = [slice(4, 20)] * dim
                if edge and bc[d] == gpu.BC_ZERO_GRADIENT:
                    layer = 19 if hi else 4
                    for h in range(4):
                        sl_dst[d] = (20 + h) if hi else h
                        sl_src[d] = layer
                        exp[(b, slice(None)) + tuple(sl_dst)] = blocks[(b, slice(None)) + tuple(sl_src)]
                    continue
                if edge and bc[d] == gpu.BC_SYMMETRY:
                    for h in range(4):
                        hidx = (20 + h) if hi else h
                        sl_dst[d] = hidx
                        sl_src[d] = (39 if hi else 7) - hidx
                        for e in range(dim + 2):
                            sign = -1.0 if e == 2 + d else 1.0
                            exp[(b, e) + tuple(sl_dst)] = sign * blocks[(b, e) + tuple(sl_src)]
                    continue
                n = list(c)
                n[d] = (c[d] + (1 if hi else -1)) % nb3[d]
                nbi = (n[0] * nb3[1] + n[1]) * nb3[2] + n[2]
                sl_dst[d] = slice(20, 24) if hi else slice(0, 4)
                sl_src[d] = slice(4, 8) if hi else slice(16, 20)
                exp[(b, slice(None)) + tuple(sl_dst)] = blocks[(nbi, slice(None)) + tuple(sl_src)]
    return exp


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_symmetry_sides_match_oracle(torch, flux):
    """Symmetry sides: 2D four-quadrant with symmetry in x and y to t = 0.2,
    and 3D blob + shock with symmetry / zero-gradient / periodic sides for 60
    steps, both bit-exact with the oracle (itself bit-exact with the
    reference's SymmetryBoundaryCondition by the golden trajectories)."""
    print(f"[symmetry] flux={flux}", flush=True)
    st0 = cases.four_quadrant_2d(6, 4, seed=7)
    case = gpu.GpuCase(gpu.make_desc(2, [6, 4], [2, 2], flux=flux), st0, cfl=0.6, t_end=0.2)
    steps, dts = case.run(100000)
    want, want_dts, _ = oracle_py.run(2, [6, 4], [2, 2], flux, 0.6, 0.2, 100000, st0)
    assert steps == len(want_dts) and np.array_equal(np.array(dts), want_dts)
    assert np.array_equal(case.download(), want)
    st0 = cases.blobs_shock_3d(2, seed=5)
    case = gpu.GpuCase(gpu.make_desc(3, [2, 2, 2], [2, 0, 1], flux=flux), st0, cfl=0.5)
    case.advance(60)
    want, _, _ = oracle_py.run(3, [2, 2, 2], [2, 0, 1], flux, 0.5, 1e300, 60, st0)
    got = case.download()
    print(f"  3D max rel err={_maxrel(got, want):.3e}", flush=True)
    assert np.array_equal(got, want)


def _face_slab_mask(dim):
    m = np.zeros((24,) * dim, dtype=bool)
    for d in range(dim):
        for lo, hi in ((0, 4), (20, 24)):
            sl = [slice(4, 20)] * dim
            sl[d] = slice(lo, hi)
            m[tuple(sl)] = True
    return m


@pytest.mark.parametrize("dim,nb,bc", [(2, [3, 2], [0, 1]), (3, [2, 3, 2], [1, 0, 1]),
                                       (3, [1, 1, 2], [1, 1, 0]), (2, [3, 2], [2, 2]),
                                       (3, [2, 1, 3], [2, 1, 0]), (3, [1, 2, 2], [0, 2, 2])])
def test_fill_halos_face_slabs(torch, dim, nb, bc):
    print(f"[halo] dim={dim} nb={nb} bc={bc}", flush=True)
    lib = gpu.load_library()
    rng = np.random.default_rng(5)
    nbk = int(np.prod(nb))
    blocks = rng.uniform(0.5, 2.0, (nbk, dim + 2) + (24,) * dim)
    dev = torch.as_tensor(blocks, device="cuda").contiguous()
    desc = gpu.make_desc(dim, nb, bc)
    gpu.check(lib.alpaca_fill_halos(ctypes.byref(desc), ctypes.c_void_p(dev.data_ptr()),
                                    gpu._stream_ptr()), "fill_halos")
    got = dev.cpu().numpy()
    exp = _face_slab_reference(blocks, dim, nb, bc)
    mask = _face_slab_mask(dim)
    assert np.array_equal(got[..., mask], exp[..., mask])
    # Internal cells untouched.
    inner = (slice(None), slice(None)) + (slice(4, 20),) * dim
    assert np.array_equal(got[inner], blocks[inner])


@pytest.mark.parametrize("dim", [2, 3])
def test_stage_api_matches_advance_and_fills_halos(torch, dim):
    """Three alpaca_rk3_stage calls (reference buffer semantics: stage 0 fills
    the initial buffer) equal one alpaca_advance step, and each stage's output
    halos equal a separate alpaca_fill_halos of that output."""
    print(f"[stage api] dim={dim}", flush=True)
    lib = gpu.load_library()
    if dim == 2:
        nb, bc, st0 = [3, 2], [0, 1], cases.four_quadrant_2d(3, 2, seed=4)
    else:
        nb, bc, st0 = [2, 2, 1], [1, 0, 1], cases.blobs_shock_3d(2, seed=2)[:, :, :, :16].copy()
    desc = gpu.make_desc(dim, nb, bc, flux=gpu.FLUX_HLLC)
    a = gpu.GpuCase(desc, st0, cfl=0.5)
    b = gpu.GpuCase(desc, st0, cfl=0.5)
    s = gpu._stream_ptr()
    d = ctypes.byref(desc)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    gpu.check(lib.alpaca_timestep(d, p(b.u), p(b.state), s), "timestep")
    init = torch.zeros_like(b.u)
    mask = _face_slab_mask(dim)
    cur = b.u
    outs = [b.s1, b.s2, torch.zeros_like(b.u)]
    for stage in range(3):
        gpu.check(lib.alpaca_rk3_stage(d, stage, p(cur), p(init), p(outs[stage]), p(b.state), s),
                  f"stage {stage}")
        o = outs[stage].clone()
        gpu.check(lib.alpaca_fill_halos(d, p(o), s), "fill_halos")
        torch.cuda.synchronize()
        assert np.array_equal(outs[stage].cpu().numpy()[..., mask], o.cpu().numpy()[..., mask])
        cur = outs[stage]
    a.advance(1)
    torch.cuda.synchronize()
    want = gpu.from_blocks(a.u, dim, a.nb).cpu().numpy()
    got = gpu.from_blocks(cur, dim, a.nb).cpu().numpy()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_conservation_256_periodic_100_steps(torch, flux):
    """Full 256^3 periodic case: mass, momentum and energy sums after 100
    steps equal the initial sums to <= 1e-13 (relative to sum |q|, the total
    momentum magnitude for the momentum components)."""
    print(f"[conservation 256^3] flux={flux}", flush=True)
    desc = gpu.make_desc(3, [16, 16, 16], [1, 1, 1], flux=flux)
    ic = cases.blobs_shock_3d_torch(16, seed=0)
    case = gpu.GpuCase(desc, ic, cfl=0.5)

    def sums():
        g = gpu.from_blocks(case.u, 3, case.nb)
        return g.sum(dim=(1, 2, 3)).cpu().numpy(), g.abs().sum(dim=(1, 2, 3)).cpu().numpy()

    s0, a0 = sums()
    case.advance(100)
    s1, a1 = sums()
    rec = case.record()
    assert rec.steps == 100 and rec.flags == 0
    # Momentum components are measured against the total momentum magnitude:
    # the initial y/z momenta are exactly zero.
    scale = np.maximum(a0, a1)
    scale[2:] = np.maximum(a0[2:].sum(), a1[2:].sum())
    drift = np.abs(s1 - s0) / scale
    print(f"  drift per field {drift}")
    assert np.all(drift <= 1e-13)


def test_fp64_fma_within_bound(torch):
    """FP64 with FMA contraction: not bit-exact, but within the FP64 bound
    (max error <= 1e-10 of the field maximum) after 100 steps, 3D periodic."""
    print("[fp64-fma 3d]", flush=True)
    nb = [2, 2, 2]
    state0 = cases.blobs_shock_3d(2, seed=1)
    for flux in (gpu.FLUX_HLLC, gpu.FLUX_ROE):
        desc = gpu.make_desc(3, nb, [1, 1, 1], flux=flux, precision=gpu.FP64_FMA)
        case = gpu.GpuCase(desc, state0, cfl=0.5)
        case.advance(100)
        got = case.download()
        want, _, _ = oracle_py.run(3, nb, [1, 1, 1], flux, 0.5, 1e300, 100, state0)
        err = _maxrel(got, want)
        print(f"  flux={flux} max rel err={err:.3e}")
        assert err <= REL_TOL


@pytest.mark.parametrize("nb", [[8, 8], [16, 16]])
def test_fp64_fma_within_bound_2d(torch, nb):
    """FP64 with FMA contraction in 2D, both single-case tilings (8x8 tiles
    below 256 blocks, 16x16 tiles from 256 blocks): four-quadrant problem,
    zero-gradient sides, max error <= 1e-10 of the field maximum."""
    steps = 100 if nb[0] == 8 else 30
    print(f"[fp64-fma 2d] nb={nb} steps={steps}", flush=True)
    state0 = cases.four_quadrant_2d(nb[0], nb[1], seed=0)
    for flux in (gpu.FLUX_HLLC, gpu.FLUX_ROE):
        desc = gpu.make_desc(2, nb, [0, 0], flux=flux, precision=gpu.FP64_FMA)
        case = gpu.GpuCase(desc, state0, cfl=0.6)
        case.advance(steps)
        got = case.download()
        want, _, _ = oracle_py.run(2, nb, [0, 0], flux, 0.6, 1e300, steps, state0)
        err = _maxrel(got, want)
        print(f"  flux={flux} max rel err={err:.3e}")
        assert err <= REL_TOL


MIXED_L1_RHO_TOL = 1e-6  # L1 density error of the mixed form vs the FP64 oracle


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_mixed_density_error(torch, flux):
    """Mixed form (FP32 faces, FP64 pressure recovery and RK): relative L1
    density error vs the FP64 oracle after 100 steps, and mass drift."""
    print(f"[mixed 3d] flux={flux}", flush=True)
    nb = [2, 2, 2]
    state0 = cases.blobs_shock_3d(2, seed=1)
    desc = gpu.make_desc(3, nb, [1, 1, 1], flux=flux, precision=gpu.MIXED)
    case = gpu.GpuCase(desc, state0, cfl=0.5)
    case.advance(100)
    got = case.download()
    want, _, _ = oracle_py.run(3, nb, [1, 1, 1], flux, 0.5, 1e300, 100, state0)
    l1 = float(np.abs(got[0] - want[0]).sum() / np.abs(want[0]).sum())
    mass = abs(got[0].sum() - state0[0].sum()) / state0[0].sum()
    print(f"  L1 rho err={l1:.3e} mass drift={mass:.3e}")
    assert l1 <= MIXED_L1_RHO_TOL
    assert mass <= 1e-13


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_edge_blocks_bitexact(torch, dim, flux):
    """The special branches (face skips, S* = 0, the eigensystem skip, the
    scaled square-root and division paths; tests/edge_states.py): the FP64
    kernels' per-block fluxes and right-hand sides equal the oracle's bit for
    bit, NaNs in the same places."""
    import edge_states

    print(f"[edge blocks] dim={dim} flux={flux}", flush=True)
    lib = gpu.load_library()
    kinds = edge_states.KINDS
    host = np.stack([edge_states.edge_block(dim, k, seed=3) for k in kinds])
    dev = torch.as_tensor(host, device="cuda").contiguous()
    desc = gpu.make_desc(dim, [1, 1, 1], [0, 0, 0], flux=flux, cell_size=1.0 / 128)
    nfl = oracle_py.nflux(dim)
    out = torch.zeros(len(kinds) * nfl, dtype=torch.float64, device="cuda")
    rhs = torch.zeros_like(dev)
    s = gpu._stream_ptr()
    gpu.check(lib.alpaca_convective_fluxes(ctypes.byref(desc), len(kinds),
                                           ctypes.c_void_p(dev.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), s), "fluxes")
    gpu.check(lib.alpaca_right_hand_side(ctypes.byref(desc), len(kinds),
                                         ctypes.c_void_p(dev.data_ptr()),
                                         ctypes.c_void_p(rhs.data_ptr()), s), "rhs")
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(len(kinds), -1)
    got_rhs = gpu.from_blocks(rhs, dim, [len(kinds), 1, 1]).cpu().numpy()
    icz = 16 if dim == 3 else 1
    with np.errstate(all="ignore"):
        for bi, kind in enumerate(kinds):
            want = oracle_py.block_fluxes(dim, flux, 1.4, 0.0, host[bi], 1.0 / 128).ravel()
            assert np.array_equal(got[bi], want, equal_nan=True), kind
            want_rhs = oracle_py.block_rhs(dim, flux, 1.4, 0.0, host[bi], 1.0 / 128)
            g = got_rhs[:, bi * 16:(bi + 1) * 16, :, :icz]
            assert np.array_equal(g, want_rhs, equal_nan=True), kind


@pytest.mark.parametrize("prec", [gpu.FP64, gpu.MIXED])
def test_chunked_run_stops_at_end_time(torch, prec):
    """Advancing in chunks of 7 or 1000 steps with one read-back per chunk
    (the device skips the steps after t_end; 1000 exceeds the run's step
    count) gives the per-step run's state bits, step count and final time."""
    print(f"[chunked run] precision={prec}", flush=True)
    st0 = cases.four_quadrant_2d(6, 4, seed=3)
    desc = gpu.make_desc(2, [6, 4], [0, 0], flux=gpu.FLUX_HLLC, precision=prec)
    out = {}
    for every in (1, 7, 1000):
        case = gpu.GpuCase(desc, st0, cfl=0.6, t_end=0.2)
        steps, _ = case.run(max_steps=100000, sync_every=every)
        rec = case.record()
        out[every] = (case.download(), steps, rec.time, rec.steps, rec.flags)
    a, sa, ta, ra, fa = out[1]
    assert sa == ra < 1000 and not ta < 0.2 and fa == 4
    for every in (7, 1000):
        b, sb, tb, rb, fb = out[every]
        assert sb == rb == sa and tb == ta and fb == fa
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
