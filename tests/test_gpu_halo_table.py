"""Halo fill over a neighbour table (alpaca_fill_halos_table) on the GPU.

Expected haloed blocks come from a numpy restatement of the reference's halo
semantics on the whole domain: pad each axis by four cells -- wrap for a
periodic direction, the edge value for zero gradient (boundary_constants.h:
212-253), the mirror for symmetry with that axis's momentum negated
(boundary_constants.h:39-48,169-208) -- axis after axis (so edges and
corners compose), then cut every block's 24-cell window.  Every halo cell,
faces, edges and corners, is compared bit for bit; the face slabs also
against the grid entry alpaca_fill_halos.  A shuffled block numbering with
the table renumbered to match must give the same blocks.
"""
import numpy as np
import pytest

from alpaca_amd import gpu

pytestmark = pytest.mark.gpu

IC, HS, TC = 16, 4, 24


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _padded(g, dim, bc):
    """The domain array [NE][NX][NY][NZ] with every axis padded by HS."""
    out = g
    for a in range(dim):
        width = [(0, 0)] * out.ndim
        width[1 + a] = (HS, HS)
        mode = {gpu.BC_PERIODIC: "wrap", gpu.BC_ZERO_GRADIENT: "edge",
                gpu.BC_SYMMETRY: "symmetric"}[bc[a]]
        out = np.pad(out, width, mode=mode)
        if bc[a] == gpu.BC_SYMMETRY:
            idx = [slice(None)] * out.ndim
            idx[0] = 2 + a
            lo, hi = list(idx), list(idx)
            lo[1 + a] = slice(0, HS)
            hi[1 + a] = slice(out.shape[1 + a] - HS, None)
            out[tuple(lo)] *= -1.0
            out[tuple(hi)] *= -1.0
    return out


def _expected_blocks(g, dim, nb, bc):
    p = _padded(g, dim, bc)
    nb3 = list(nb) + [1] * (3 - len(nb))
    tz = TC if dim == 3 else 1
    out = np.empty((nb3[0] * nb3[1] * nb3[2], dim + 2, TC, TC, tz))
    for bx in range(nb3[0]):
        for by in range(nb3[1]):
            for bz in range(nb3[2]):
                k = (bx * nb3[1] + by) * nb3[2] + bz
                w = p[:, bx * IC:bx * IC + TC, by * IC:by * IC + TC]
                w = w[..., bz * IC:bz * IC + TC] if dim == 3 else w
                out[k] = w
    return out


CASES = [
    (2, [3, 2], [gpu.BC_PERIODIC, gpu.BC_ZERO_GRADIENT]),
    (2, [2, 3], [gpu.BC_SYMMETRY, gpu.BC_PERIODIC]),
    (2, [1, 1], [gpu.BC_SYMMETRY, gpu.BC_ZERO_GRADIENT]),
    (3, [2, 2, 2], [gpu.BC_PERIODIC, gpu.BC_PERIODIC, gpu.BC_PERIODIC]),
    (3, [3, 2, 2], [gpu.BC_ZERO_GRADIENT, gpu.BC_SYMMETRY, gpu.BC_PERIODIC]),
    (3, [1, 2, 3], [gpu.BC_SYMMETRY, gpu.BC_SYMMETRY, gpu.BC_ZERO_GRADIENT]),
]


@pytest.mark.parametrize("dim,nb,bc", CASES)
def test_table_fill_matches_padded_domain(torch, dim, nb, bc):
    print(f"[halo table] dim={dim} nb={nb} bc={bc}", flush=True)
    rng = np.random.default_rng(sum(nb) + 10 * dim)
    shape = (dim + 2,) + tuple(n * IC for n in nb) + ((1,) if dim == 2 else ())
    g = rng.standard_normal(shape)
    want = _expected_blocks(g, dim, nb, bc)
    dev = gpu.to_blocks(torch.as_tensor(g, device="cuda"), dim,
                        list(nb) + [1] * (3 - len(nb))).contiguous()
    table = torch.as_tensor(gpu.grid_neighbors(dim, nb, bc), device="cuda")
    side = [bc[a] for a in range(dim) for _ in (0, 1)]
    gpu.fill_halos_table(dev, dim, table, side)
    torch.cuda.synchronize()
    got = dev.cpu().numpy().reshape(want.shape)
    assert np.array_equal(got, want)
    # the grid entry writes the face slabs only: they must agree
    desc = gpu.make_desc(dim, nb, bc)
    ref = gpu.to_blocks(torch.as_tensor(g, device="cuda"), dim,
                        list(nb) + [1] * (3 - len(nb))).contiguous()
    gpu.check(gpu.load_library().alpaca_fill_halos(gpu.ctypes.byref(desc),
                                                   gpu.ctypes.c_void_p(ref.data_ptr()),
                                                   gpu._stream_ptr()), "fill_halos")
    torch.cuda.synchronize()
    ref = ref.cpu().numpy().reshape(want.shape)
    inner = slice(HS, HS + IC)
    zi = inner if dim == 3 else slice(0, 1)
    for a in range(dim):
        for side_slice in (slice(0, HS), slice(HS + IC, TC)):
            idx = [slice(None), slice(None), inner, inner, zi]
            idx[2 + a] = side_slice
            assert np.array_equal(got[tuple(idx)], ref[tuple(idx)]), (a, side_slice)


def test_table_indirection_with_shuffled_blocks(torch):
    dim, nb, bc = 3, [2, 3, 2], [gpu.BC_PERIODIC, gpu.BC_ZERO_GRADIENT, gpu.BC_SYMMETRY]
    rng = np.random.default_rng(5)
    shape = (dim + 2,) + tuple(n * IC for n in nb)
    g = rng.standard_normal(shape)
    want = _expected_blocks(g, dim, nb, bc)
    table = gpu.grid_neighbors(dim, nb, bc)
    n = table.shape[0]
    perm = rng.permutation(n)          # new index of block k is perm[k]
    inv = np.argsort(perm)
    shuffled_table = np.where(table[inv] >= 0, perm[np.maximum(table[inv], 0)], -1).astype(np.int32)
    blocks = gpu.to_blocks(torch.as_tensor(g, device="cuda"), dim, nb).contiguous()
    shuffled = blocks[torch.as_tensor(inv, device="cuda")].contiguous()
    gpu.fill_halos_table(shuffled, dim, torch.as_tensor(shuffled_table, device="cuda"),
                         [bc[a] for a in range(dim) for _ in (0, 1)])
    torch.cuda.synchronize()
    got = shuffled.cpu().numpy()[perm]
    assert np.array_equal(got, want)


def test_table_fill_rejects_bad_arguments(torch):
    lib = gpu.load_library()
    side = (gpu.ctypes.c_int * 6)(0, 0, 0, 0, 0, 0)
    assert lib.alpaca_fill_halos_table(4, None, 1, None, side, None) == 1
    bad = (gpu.ctypes.c_int * 6)(0, 0, 7, 0, 0, 0)
    t = torch.zeros(16, dtype=torch.float64, device="cuda")
    n = torch.zeros(9, dtype=torch.int32, device="cuda")
    assert lib.alpaca_fill_halos_table(2, gpu.ctypes.c_void_p(t.data_ptr()), 1,
                                       gpu.ctypes.c_void_p(n.data_ptr()), bad, None) == 1
