"""Multiresolution operators on the GPU (alpaca_mr_average,
alpaca_mr_average_jump_buffers, alpaca_mr_prediction,
alpaca_mr_child_needs_remeshing) against the oracle.

Averaging and prediction use only additions, subtractions and
multiplications in the reference's order, so they are compared bit for bit
(NaN positions equal).  The detail divides; it is compared to the oracle's
within one unit in the last place (the hardware FP64 division sequence may
round a near-midpoint quotient the other way), and the remeshing decisions,
taken on thresholds away from the details, exactly.  At the full size (512
parents, 4096 children: the 256^3 cells of level 1) the properties that do
not depend on size are checked: a child predicted from its parent over the
whole block has detail 0 (decision Coarse), averaging the eight children of
a parent conserves every equation's sum, and sampled links agree with the
oracle bit for bit.
"""
import ctypes

import numpy as np
import pytest

from alpaca_amd import gpu
from alpaca_amd import multiresolution as mr
from oracle import oracle_py as O

import mr_states

pytestmark = pytest.mark.gpu

IC, HS, TC = 16, 4, 24


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _pool(torch, dim, kinds):
    """Pool [p0, c0, p1, c1, ...] of every pair kind and position, with links
    (child 2i+1, parent 2i, position, level i % 3)."""
    blocks, rows, pairs = [], [], []
    for kind in kinds:
        for pos in range(2 ** dim):
            p, c = mr_states.pair(dim, kind, pos)
            rows.append((len(blocks) + 1, len(blocks), pos, len(rows) % 3))
            blocks += [p, c]
            pairs.append((kind, pos, p, c))
    host = np.ascontiguousarray(np.stack(blocks))
    dev = torch.as_tensor(host, device="cuda").contiguous()
    links = torch.as_tensor(mr.make_links(rows), device="cuda")
    return host, dev, links, pairs


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("dim", [2, 3])
def test_average_bitexact(torch, dim):
    print(f"[mr average] dim={dim}", flush=True)
    host, dev, links, pairs = _pool(torch, dim, mr_states.KINDS)
    mr.Multiresolution.average(dev, dim, links)
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    with np.errstate(all="ignore"):
        for i, (kind, pos, p, c) in enumerate(pairs):
            want = O.mr_average(dim, c, p, pos)
            assert _same(got[2 * i], want), (kind, pos)
            assert _same(got[2 * i + 1], c), (kind, pos)  # the child is read only


@pytest.mark.parametrize("dim", [2, 3])
def test_prediction_bitexact_every_range(torch, dim):
    for rng in mr_states.ranges(dim):
        print(f"[mr prediction] dim={dim} range={rng}", flush=True)
        host, dev, links, pairs = _pool(torch, dim, mr_states.KINDS)
        mr.Multiresolution.prediction(dev, dim, links, rng)
        torch.cuda.synchronize()
        got = dev.cpu().numpy()
        with np.errstate(all="ignore"):
            for i, (kind, pos, p, c) in enumerate(pairs):
                want = np.stack([O.mr_prediction(dim, p[e], c[e], pos, rng)
                                 for e in range(dim + 2)])
                assert _same(got[2 * i + 1], want), (kind, pos, rng)
                assert _same(got[2 * i], p), (kind, pos)


@pytest.mark.parametrize("dim", [2, 3])
def test_details_and_decisions(torch, dim):
    print(f"[mr details] dim={dim}", flush=True)
    host, dev, links, pairs = _pool(torch, dim, mr_states.KINDS)
    with np.errstate(all="ignore"):
        want = np.array([O.mr_detail_linf(dim, p, c, pos) for _, pos, p, c in pairs])
    # thresholds on levels 0..2 (link i on level i % 3) that split the smooth
    # details (2e-4 .. 4e-3) so that all three decisions occur, each away from
    # its link's detail by more than 1 ulp
    th = mr.ThresholdComputer(dim, 2, 2, 4.0e-3)
    for i, w in enumerate(want):
        eps = th.threshold_on_level(i % 3)
        if np.isfinite(w):
            assert abs(w - eps) > 1e-12 * eps and abs(w - 32 * eps) > 1e-12 * eps
    details, decisions = mr.Multiresolution(th).child_needs_remeshing(dev, dim, links)
    torch.cuda.synchronize()
    got = details.cpu().numpy()
    dec = decisions.cpu().numpy()
    for i, (kind, pos, _, _) in enumerate(pairs):
        w, g = want[i], got[i]
        assert g == w or abs(g - w) <= np.spacing(w), (kind, pos, g, w)
        assert dec[i] == O.mr_decision(w, th.threshold_on_level(i % 3)), (kind, pos, dec[i], w)
    assert set(dec.tolist()) == {mr.NEUTRAL, mr.REFINE, mr.COARSE}
    # the blocks are read only
    assert _same(dev.cpu().numpy(), host)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("norm", [mr.L1, mr.L2])
def test_details_l1_l2(torch, dim, norm):
    """The L-one and L-two details: tree-reduced sums, equal to the oracle's
    sequential sums to 1e-13 relative; non-finite sums classified alike."""
    print(f"[mr details norm {norm}] dim={dim}", flush=True)
    host, dev, links, pairs = _pool(torch, dim, mr_states.KINDS)
    with np.errstate(all="ignore"):
        want = np.array([O.mr_detail(dim, p, c, pos, norm) for _, pos, p, c in pairs])
    th = mr.ThresholdComputer(dim, 2, 2, 4.0e-3)
    details, decisions = mr.Multiresolution(th, norm).child_needs_remeshing(dev, dim, links)
    torch.cuda.synchronize()
    got = details.cpu().numpy()
    dec = decisions.cpu().numpy()
    for i, (kind, pos, _, _) in enumerate(pairs):
        w, g = want[i], got[i]
        if np.isfinite(w):
            assert abs(g - w) <= 1e-13 * abs(w), (kind, pos, g, w)
            eps = th.threshold_on_level(i % 3)
            if abs(w - eps) > 1e-10 * eps and abs(w - 32 * eps) > 1e-10 * eps:
                assert dec[i] == O.mr_decision(w, eps), (kind, pos, dec[i], w)
        else:
            assert g == w, (kind, pos, g, w)  # +inf (NaN sums are dropped)


@pytest.mark.parametrize("dim", [2, 3])
def test_all_children_of_one_parent(torch, dim):
    """The octree use: one parent, all 2^dim children in one call."""
    print(f"[mr siblings] dim={dim}", flush=True)
    n = 2 ** dim
    parent, _ = mr_states.pair(dim, "smooth", 0)
    children = [mr_states.pair(dim, "smooth", pos)[1] for pos in range(n)]
    host = np.ascontiguousarray(np.stack([parent] + children))
    dev = torch.as_tensor(host, device="cuda").contiguous()
    links = torch.as_tensor(mr.make_links([(1 + pos, 0, pos, 1) for pos in range(n)]),
                            device="cuda")
    mr.Multiresolution.average(dev, dim, links)
    torch.cuda.synchronize()
    want = parent
    for pos in range(n):
        want = O.mr_average(dim, children[pos], want, pos)
    assert _same(dev.cpu().numpy()[0], want)
    # every internal parent cell was written: the internal mean is the mean
    # of the children's internal means
    inner = (slice(None), slice(HS, HS + IC), slice(HS, HS + IC),
             slice(HS, HS + IC) if dim == 3 else slice(0, 1))
    got_mean = dev.cpu().numpy()[0][inner].mean(axis=(1, 2, 3))
    child_mean = np.mean([c[inner].mean(axis=(1, 2, 3)) for c in children], axis=0)
    assert np.allclose(got_mean, child_mean, rtol=1e-13, atol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_buffer_average_bitexact(torch, dim):
    """AverageJumpBuffer: three parents with all their children (shuffled
    link order, one call): every parent's surfaces equal the oracle's
    children-by-children accumulation bit for bit; children unchanged."""
    print(f"[mr jump average] dim={dim}", flush=True)
    n = 2 ** dim
    icz = 16 if dim == 3 else 1
    rng = np.random.default_rng(21 + dim)
    nparents = 3
    nb = nparents * (n + 1)
    host = rng.standard_normal((nb, 6, dim + 2, 16, icz))
    rows = [(p * (n + 1) + 1 + pos, p * (n + 1), pos, 1) for p in range(nparents) for pos in range(n)]
    order = rng.permutation(len(rows))
    links = torch.as_tensor(mr.make_links([rows[k] for k in order]), device="cuda")
    dev = torch.as_tensor(host, device="cuda").contiguous()
    mr.Multiresolution.average_jump_buffer(dev, dim, links)
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    want = host.copy()
    for child, parent, pos, _ in rows:
        want[parent] = O.mr_average_jump(dim, host[child], want[parent], pos)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_skipped_links_and_bad_arguments(torch, dim):
    print(f"[mr skipped] dim={dim}", flush=True)
    host, dev, _, pairs = _pool(torch, dim, ["smooth"])
    nb = host.shape[0]
    bad = torch.as_tensor(mr.make_links([(nb, 0, 0, 1), (1, 1, 0, 1), (1, 0, 2 ** dim, 1),
                                         (-1, 0, 0, 1), (1, nb + 5, 0, 1)]), device="cuda")
    mr.Multiresolution.average(dev, dim, bad)
    mr.Multiresolution.prediction(dev, dim, bad)
    th = mr.ThresholdComputer(dim, 1, 1, 1e-3)
    details, decisions = mr.Multiresolution(th).child_needs_remeshing(dev, dim, bad)
    torch.cuda.synchronize()
    assert _same(dev.cpu().numpy(), host)
    assert np.all(np.isnan(details.cpu().numpy()))
    assert np.all(decisions.cpu().numpy() == -1)
    # a level outside the threshold table: detail computed, decision -1
    lvl = torch.as_tensor(mr.make_links([(1, 0, 0, 5)]), device="cuda")
    details, decisions = mr.Multiresolution(th).child_needs_remeshing(dev, dim, lvl)
    torch.cuda.synchronize()
    assert np.isfinite(details.cpu().numpy()[0]) and decisions.cpu().numpy()[0] == -1

    lib = gpu.load_library()
    good = torch.as_tensor(mr.make_links([(1, 0, 0, 1)]), device="cuda")
    args = (dim, ctypes.c_void_p(dev.data_ptr()), nb, ctypes.c_void_p(good.data_ptr()), 1)
    st = gpu._stream_ptr()
    z = (0, 1) if dim == 2 else (0, TC)
    for rng in [(1, 4, 0, TC) + z, (0, TC + 2, 0, TC) + z, (-2, 4, 0, TC) + z,
                (0, TC, 20, 6) + z]:
        r = (ctypes.c_int * 6)(*rng)
        assert lib.alpaca_mr_prediction(*args, r, st) == 1, rng
    eps = (ctypes.c_double * 33)(*([1e-3] * 33))
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    dec = torch.empty(1, dtype=torch.int32, device="cuda")
    assert lib.alpaca_mr_child_needs_remeshing(*args, 0, eps, 33, ctypes.c_void_p(out.data_ptr()),
                                               ctypes.c_void_p(dec.data_ptr()), st) \
        == 1
    assert lib.alpaca_mr_child_needs_remeshing(*args, 0, None, 0, ctypes.c_void_p(out.data_ptr()),
                                               ctypes.c_void_p(dec.data_ptr()), st) \
        == 1
    assert lib.alpaca_mr_child_needs_remeshing(*args, 3, eps, 2, ctypes.c_void_p(out.data_ptr()),
                                               ctypes.c_void_p(dec.data_ptr()), st) \
        == 1
    assert lib.alpaca_mr_average(4, *args[1:], st) == 1
    # details alone (no decisions) need no threshold table
    assert lib.alpaca_mr_child_needs_remeshing(*args, 0, None, 0, ctypes.c_void_p(out.data_ptr()),
                                               None, st) == 0
    torch.cuda.synchronize()
    want = O.mr_detail_linf(dim, host[0], host[1], 0)
    g = out.cpu().numpy()[0]
    assert g == want or abs(g - want) <= np.spacing(want)


def test_full_size_level_one_properties(torch):
    """512 parents and their 4096 children (3D, 256^3 cells on level 1)."""
    dim, nparents = 3, 512
    print(f"[mr full size] parents={nparents}", flush=True)
    n = nparents * 9
    bd = gpu.block_doubles(dim)
    g = torch.Generator(device="cuda").manual_seed(11)
    dev = (torch.rand(n * bd, dtype=torch.float64, device="cuda", generator=g) + 0.5).contiguous()
    rows = [(p * 9 + 1 + pos, p * 9, pos, 1) for p in range(nparents) for pos in range(8)]
    links = torch.as_tensor(mr.make_links(rows), device="cuda")
    blocks = dev.view(n, dim + 2, TC, TC, TC)
    sample = [0, 8 * 77 + 3, 8 * 300 + 6, 8 * 511 + 7]
    before = {k: (blocks[rows[k][1]].cpu().numpy(), blocks[rows[k][0]].cpu().numpy())
              for k in sample}

    # averaging: bit-exact on sampled links, every parent's internal sum
    # equals the mean of its children's internal sums
    mr.Multiresolution.average(dev, dim, links)
    torch.cuda.synchronize()
    inner = (slice(None), slice(None), slice(HS, HS + IC), slice(HS, HS + IC), slice(HS, HS + IC))
    parents = blocks[0::9][inner].sum(dim=(2, 3, 4))
    kids = blocks.view(nparents, 9, dim + 2, TC, TC, TC)[:, 1:][..., HS:HS + IC, HS:HS + IC,
                                                            HS:HS + IC].sum(dim=(3, 4, 5))
    assert torch.allclose(parents * 8.0, kids.sum(dim=1), rtol=1e-13, atol=0)
    for k in sample:
        p, c = before[k]
        want = O.mr_average(dim, c, p, rows[k][2])
        got = blocks[rows[k][1]].cpu().numpy()
        x0 = [TC // 2 if (rows[k][2] >> a) & 1 else HS for a in range(3)]
        sl = (slice(None),) + tuple(slice(x, x + IC // 2) for x in x0)
        assert np.array_equal(got[sl], want[sl]), k

    # prediction over the whole block, then details: every child is its own
    # parent's prediction, so every detail is 0 and every decision Coarse
    mr.Multiresolution.prediction(dev, dim, links)
    th = mr.ThresholdComputer(dim, 1, 1, 1e-3)
    details, decisions = mr.Multiresolution(th).child_needs_remeshing(dev, dim, links)
    torch.cuda.synchronize()
    assert float(details.abs().max()) == 0.0
    assert bool((decisions == mr.COARSE).all())
    for k in sample:
        p = blocks[rows[k][1]].cpu().numpy()
        want = np.stack([O.mr_prediction(dim, p[e], np.zeros((TC, TC, TC)), rows[k][2])
                         for e in range(dim + 2)])
        assert np.array_equal(blocks[rows[k][0]].cpu().numpy(), want), k


def _valid_locations(dim):
    return [loc for loc in range(26) if dim == 3 or mr.location_direction(loc)[2] == 0]


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_halo_update_bitexact(torch, dim):
    """alpaca_mr_jump_halo_update: each link's halo slab at its own location
    equals the oracle's prediction over that location's receive range; the
    rest of the child and the parent are untouched; an invalid location (and
    in 2D one with a Top/Bottom part) leaves the child as it was."""
    print(f"[mr jump halo] dim={dim}", flush=True)
    host, dev, links, pairs = _pool(torch, dim, mr_states.KINDS)
    valid = _valid_locations(dim)
    bad = [26, -1] + ([4, 8, 25] if dim == 2 else [])
    locs = [valid[i % len(valid)] if i % 7 else bad[(i // 7) % len(bad)]
            for i in range(len(pairs))]
    mr.Multiresolution.jump_halo_update(dev, dim, links,
                                        torch.tensor(locs, dtype=torch.int32, device="cuda"))
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    with np.errstate(all="ignore"):
        for i, (kind, pos, p, c) in enumerate(pairs):
            if locs[i] in valid:
                rng = mr.halo_recv_range(dim, locs[i])
                want = np.stack([O.mr_prediction(dim, p[e], c[e], pos, rng)
                                 for e in range(dim + 2)])
            else:
                want = c
            assert _same(got[2 * i + 1], want), (kind, pos, locs[i])
            assert _same(got[2 * i], p), (kind, pos)


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_halo_update_every_location_of_one_child(torch, dim):
    """One launch over every location of the same child (disjoint slabs of
    one block written by concurrent workgroups) equals the oracle applying
    the ranges one after the other: the whole halo is predicted, the internal
    cells keep their values."""
    print(f"[mr jump halo, all locations] dim={dim}", flush=True)
    valid = _valid_locations(dim)
    for pos in range(2 ** dim):
        p, c = mr_states.pair(dim, "smooth", pos, seed=5)
        dev = torch.as_tensor(np.ascontiguousarray(np.stack([p, c])), device="cuda").contiguous()
        links = torch.as_tensor(mr.make_links([(1, 0, pos, 1)] * len(valid)), device="cuda")
        mr.Multiresolution.jump_halo_update(dev, dim, links,
                                            torch.tensor(valid, dtype=torch.int32, device="cuda"))
        torch.cuda.synchronize()
        want = c.copy()
        for loc in valid:
            rng = mr.halo_recv_range(dim, loc)
            for e in range(dim + 2):
                want[e] = O.mr_prediction(dim, p[e], want[e], pos, rng)
        assert _same(dev.cpu().numpy()[1], want), pos


def test_jump_halo_update_bad_arguments(torch):
    lib = gpu.load_library()
    dev = torch.zeros(2 * gpu.block_doubles(3), dtype=torch.float64, device="cuda")
    links = torch.as_tensor(mr.make_links([(1, 0, 0, 1)]), device="cuda")
    st = ctypes.c_void_p(0)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.alpaca_mr_jump_halo_update(3, ptr(dev), 2, ptr(links), None, 1, st) == 1
    assert lib.alpaca_mr_jump_halo_update(4, ptr(dev), 2, ptr(links), ptr(links), 1, st) == 1
    assert lib.alpaca_mr_jump_halo_update(3, ptr(dev), 2, None, None, 0, st) == 0
    with pytest.raises(gpu.AlpacaError):
        mr.Multiresolution.jump_halo_update(dev.view(2, -1), 3, links,
                                            torch.zeros(2, dtype=torch.int32, device="cuda"))
