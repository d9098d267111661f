"""Level-wise integration on the GPU (alpaca_lts_prepare_buffer,
alpaca_lts_integrate for internal cells and halo ranges,
alpaca_lts_integrate_jump_conservatives) through alpaca_amd.lts against the
oracle, bit for bit: a pool of blocks of which a shuffled list (with an
out-of-range entry, skipped) is treated, every stage and several micro-step
counts, blocks off the list untouched."""
import numpy as np
import pytest

from alpaca_amd import gpu, lts
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _pools(torch, dim, nb, seed):
    tz, icz = (24, 16) if dim == 3 else (1, 1)
    rng = np.random.default_rng(seed)
    host = {k: rng.standard_normal((nb, dim + 2, 24, 24, tz)) for k in ("avg", "rhs", "init")}
    host.update({k: rng.standard_normal((nb, 6, dim + 2, 16, icz)) for k in ("jc", "jf")})
    dev = {k: torch.as_tensor(v, device="cuda").contiguous() for k, v in host.items()}
    return host, dev


@pytest.mark.parametrize("dim", [2, 3])
def test_level_stage_operations_bitexact(torch, dim):
    nb = 7
    rng = np.random.default_rng(dim)
    chosen = [5, 1, 3, 6]
    lst = torch.as_tensor(np.array(chosen + [nb + 3], dtype=np.int32), device="cuda")
    micro = rng.uniform(1e-4, 1e-3, 8)
    for stage in range(3):
        for n in (1, 4):
            print(f"[lts] dim={dim} stage={stage} n={n}", flush=True)
            host, dev = _pools(torch, dim, nb, 10 * stage + n + dim)
            ti = lts.LevelTimeIntegrator(dim)
            for dt in micro:
                ti.append_micro_timestep(dt)
            ti.prepare_buffer_for_integration(dev["avg"], dev["init"], lst, stage)
            ti.integrate_node(dev["avg"], dev["rhs"], dev["jc"], dev["jf"], lst, stage, n)
            start, size = (20, 4, 4 if dim == 3 else 0), (4, 16, 16 if dim == 3 else 1)
            ti.integrate_jump_halos(dev["avg"], dev["init"], lst, stage, n, start, size)
            torch.cuda.synchronize()
            got = {k: v.cpu().numpy() for k, v in dev.items()}
            k = 2 * dim
            for b in range(nb):
                want = {key: host[key][b].copy() for key in host}
                if b in chosen:
                    want["avg"] = O.lts_prepare_buffer(dim, want["avg"], want["init"], stage)
                    c, f = O.lts_integrate_jump(dim, want["jc"][:k], want["jf"][:k],
                                                O.lts_timestep(micro, n, stage, True))
                    want["jc"][:k], want["jf"][:k] = c, f
                    want["rhs"] = O.lts_integrate(dim, want["avg"], want["rhs"], None,
                                                  O.lts_timestep(micro, n, stage))
                    rng6 = (start[0], size[0], start[1], size[1], start[2], size[2])
                    want["init"] = O.lts_integrate(dim, want["avg"], want["init"], rng6,
                                                   O.lts_timestep(micro, n, stage))
                for key in host:
                    assert np.array_equal(got[key][b], want[key]), (b, key)


def test_fill_initial_buffer_and_bad_arguments(torch):
    dim, nb = 3, 3
    host, dev = _pools(torch, dim, nb, 4)
    lst = torch.as_tensor(np.array([2, 0], dtype=np.int32), device="cuda")
    ti = lts.LevelTimeIntegrator(dim)
    ti.fill_initial_buffer(dev["avg"], dev["init"], lst, 0)
    torch.cuda.synchronize()
    init = dev["init"].cpu().numpy()
    assert np.array_equal(init[0], host["avg"][0]) and np.array_equal(init[2], host["avg"][2])
    assert np.array_equal(init[1], host["init"][1])
    lib = gpu.load_library()
    p = gpu.ctypes.c_void_p(dev["avg"].data_ptr())
    q = gpu.ctypes.c_void_p(dev["rhs"].data_ptr())
    l = gpu.ctypes.c_void_p(lst.data_ptr())
    st = gpu._stream_ptr()
    bad = (gpu.ctypes.c_int * 6)(20, 8, 0, 24, 0, 24)
    assert lib.alpaca_lts_integrate(3, p, q, nb, l, 2, bad, gpu.ctypes.c_double(1.0), st) == 1
    assert lib.alpaca_lts_prepare_buffer(3, p, q, nb, l, 2, 3, st) == 1
    assert lib.alpaca_lts_integrate_jump_conservatives(3, p, p, nb, l, 2,
                                                       gpu.ctypes.c_double(1.0), st) == 1


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", [0, 1])
def test_right_hand_side_with_jump_fluxes(torch, dim, flux):
    """alpaca_right_hand_side_jump_fluxes: the per-block right-hand side as
    alpaca_right_hand_side gives it, and the boundary face fluxes added into
    the jump-flux surfaces, both bit for bit against the oracle."""
    from alpaca_amd import cases

    print(f"[rhs+jump] dim={dim} flux={flux}", flush=True)
    icz = 16 if dim == 3 else 1
    nb = 3
    blocks = np.stack([cases.random_block(dim, 40 + b + 3 * flux) for b in range(nb)])
    rng = np.random.default_rng(flux + dim)
    jump0 = rng.standard_normal((nb, 6, dim + 2, 16, icz))
    dev = torch.as_tensor(blocks, device="cuda").contiguous()
    rhs = torch.empty_like(dev)
    jump = torch.as_tensor(jump0, device="cuda").contiguous()
    desc = gpu.make_desc(dim, [1, 1, 1][:dim], [0] * dim, flux=flux, cell_size=1.0 / 64)
    lib = gpu.load_library()
    gpu.check(lib.alpaca_right_hand_side_jump_fluxes(gpu.ctypes.byref(desc), nb,
                                                     gpu.ctypes.c_void_p(dev.data_ptr()),
                                                     gpu.ctypes.c_void_p(rhs.data_ptr()),
                                                     gpu.ctypes.c_void_p(jump.data_ptr()),
                                                     gpu._stream_ptr()), "rhs+jump")
    torch.cuda.synchronize()
    got_rhs = rhs.cpu().numpy()
    got_jump = jump.cpu().numpy()
    hs, ic = 4, 16
    zs = slice(hs, hs + ic) if dim == 3 else slice(0, 1)
    for b in range(nb):
        want = O.block_rhs(dim, flux, 1.4, 0.0, blocks[b], 1.0 / 64)
        assert np.array_equal(got_rhs[b][:, hs:hs + ic, hs:hs + ic, zs], want), b
        assert np.array_equal(got_jump[b], O.block_jump_fluxes(dim, flux, 1.4, 0.0, blocks[b],
                                                               jump0[b])), b


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_flux_adjustment_and_reset_bitexact(torch, dim):
    """JumpFluxAdjustment's exchange on three leaves with different sets of
    jump faces (a parent neighbour per face, others -1) against the oracle,
    bit for bit, cells and jump surfaces; then the reset zeroes the listed
    blocks' faces."""
    print(f"[lts exchange] dim={dim}", flush=True)
    tz, icz = (24, 16) if dim == 3 else (1, 1)
    rng = np.random.default_rng(60 + dim)
    nb = 6  # blocks 0..2 leaves, 3..5 non-leaf neighbours
    cells = rng.standard_normal((nb, dim + 2, 24, 24, tz))
    jump = rng.standard_normal((nb, 6, dim + 2, 16, icz))
    faces = [[3, -1, 4, -1, 5, -1], [-1, 5, -1, -1, -1, 3], [4, 3, 5, 4, 3, 5]]
    entries = np.array([[leaf] + faces[leaf] for leaf in range(3)], dtype=np.int32)
    dc = torch.as_tensor(cells, device="cuda").contiguous()
    dj = torch.as_tensor(jump, device="cuda").contiguous()
    s = 48.0
    lts.jump_flux_adjustment(dim, dc, dj, torch.as_tensor(entries, device="cuda"), s)
    torch.cuda.synchronize()
    got_c, got_j = dc.cpu().numpy(), dj.cpu().numpy()
    for leaf in range(3):
        nbs = [jump[x] if (x >= 0 and loc < 2 * dim) else None for loc, x in enumerate(faces[leaf])]
        wc, wj = O.lts_jump_flux_adjustment(dim, cells[leaf], jump[leaf], nbs, s)
        assert np.array_equal(got_c[leaf], wc), leaf
        assert np.array_equal(got_j[leaf], wj), leaf
    for b in range(3, nb):
        assert np.array_equal(got_c[b], cells[b]) and np.array_equal(got_j[b], jump[b])
    lts.reset_jump_conservatives(dim, dj, torch.as_tensor(np.array([0, 4], dtype=np.int32),
                                                          device="cuda"))
    torch.cuda.synchronize()
    after = dj.cpu().numpy()
    for b in (0, 4):
        assert np.all(after[b, :2 * dim] == 0.0)
        assert np.array_equal(after[b, 2 * dim:], got_j[b, 2 * dim:])
    assert np.array_equal(after[1], got_j[1])


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_halos_of_a_level_in_one_call(torch, dim):
    """alpaca_lts_integrate_jump_halos over (block, location) entries equals
    the oracle's per-location halo integration (IntegrateJumpHalos with
    GetStartIndicesHaloRecv / GetHaloSize), bit for bit; entries with a block
    outside the pool or (2D) a Top/Bottom location are skipped."""
    import ctypes
    from alpaca_amd import multiresolution as mr
    print(f"[lts jump halos] dim={dim}", flush=True)
    nb = 6
    host, dev = _pools(torch, dim, nb, 77 + dim)
    locs = [loc for loc in range(26) if dim == 3 or mr.location_direction(loc)[2] == 0]
    rng = np.random.default_rng(dim)
    entries = [(int(b), loc) for b in (4, 0, 5) for loc in rng.permutation(locs)[:7].tolist()]
    entries += [(nb + 1, locs[0]), (2, 26)] + ([(3, 4)] if dim == 2 else [])
    ent = torch.as_tensor(np.array(entries, dtype=np.int32), device="cuda")
    ts = 3.25e-4
    lib = gpu.load_library()
    gpu.check(lib.alpaca_lts_integrate_jump_halos(
        dim, ctypes.c_void_p(dev["avg"].data_ptr()), ctypes.c_void_p(dev["rhs"].data_ptr()), nb,
        ctypes.c_void_p(ent.data_ptr()), len(entries), ts, gpu._stream_ptr()), "jump halos")
    torch.cuda.synchronize()
    got = dev["rhs"].cpu().numpy()
    want = host["rhs"].copy()
    for b, loc in entries[:21]:
        r = mr.halo_recv_range(dim, loc)
        want[b] = O.lts_integrate(dim, host["avg"][b], want[b], tuple(r) + (0, 1) * (3 - dim), ts)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    assert torch.equal(dev["avg"].cpu(), torch.as_tensor(host["avg"]))
