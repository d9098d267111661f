"""The C ABI library loads on a host without a GPU and exports every entry
point declared in include/alpaca_gpu.h (no compute calls here)."""
import ctypes
import os
import re

import pytest

from alpaca_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alpaca_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(alpaca_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("alpaca_convective_fluxes", "alpaca_right_hand_side", "alpaca_rk3_stage",
                 "alpaca_fill_halos", "alpaca_timestep", "alpaca_advance"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.fail("libalpaca_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) <= set(gpu.exported_symbols())


def test_host_only_entry_points():
    lib = gpu.load_library()
    assert lib.alpaca_abi_version() == 1
    assert lib.alpaca_block_doubles(2) == 4 * 24 * 24
    assert lib.alpaca_block_doubles(3) == 5 * 24 ** 3
    assert lib.alpaca_block_doubles(4) == 0
    assert lib.alpaca_status_string(3) == b"non-finite state"


def test_invalid_descriptors_are_rejected_before_any_launch():
    lib = gpu.load_library()
    bad = gpu.make_desc(2, [2, 2], [0, 0])
    bad.dim = 4
    assert lib.alpaca_fill_halos(ctypes.byref(bad), ctypes.c_void_p(1), None) == 1
    bad = gpu.make_desc(3, [2, 2, 2], [1, 1, 1])
    bad.bc[1] = 7
    assert lib.alpaca_timestep(ctypes.byref(bad), ctypes.c_void_p(1), ctypes.c_void_p(1),
                               None) == 1
    good = gpu.make_desc(3, [2, 2, 2], [1, 1, 1])
    assert lib.alpaca_rk3_stage(ctypes.byref(good), 3, ctypes.c_void_p(1), None,
                                ctypes.c_void_p(2), ctypes.c_void_p(3), None) == 1
    assert lib.alpaca_advance(ctypes.byref(good), ctypes.c_void_p(1), ctypes.c_void_p(1),
                              ctypes.c_void_p(2), ctypes.c_void_p(3), 1, None) == 1


def test_oversized_counts_are_rejected_before_any_launch():
    """Block and link counts beyond 2^24 (every launch grid derived from them
    fits 32 bits below that) are refused with ALPACA_ERR_INVALID."""
    lib = gpu.load_library()
    huge = gpu.make_desc(3, [256, 256, 257], [1, 1, 1])
    assert lib.alpaca_fill_halos(ctypes.byref(huge), ctypes.c_void_p(16), None) == 1
    big = 1 << 25
    side = (ctypes.c_int * 6)(0, 0, 0, 0, 0, 0)
    assert lib.alpaca_fill_halos_table(3, ctypes.c_void_p(16), big, ctypes.c_void_p(16), side,
                                       None) == 1
    p = ctypes.c_void_p(256)
    assert lib.alpaca_mr_average(3, p, big, p, 1, None) == 1
    assert lib.alpaca_mr_average(3, p, 10, p, big, None) == 1
    assert lib.alpaca_mr_average_jump_buffers(2, p, 10, p, -1, None) == 1
    rng = (ctypes.c_int * 6)(0, 24, 0, 24, 0, 24)
    assert lib.alpaca_mr_prediction(3, p, 10, p, big, rng, None) == 1
    assert lib.alpaca_mr_jump_halo_update(3, p, 10, p, p, big, None) == 1
    assert lib.alpaca_mr_jump_halo_update(3, p, big, p, p, 1, None) == 1
    assert lib.alpaca_mr_jump_halo_update(3, p, 10, p, None, 1, None) == 1  # no locations
    assert lib.alpaca_mr_jump_halo_update(3, ctypes.c_void_p(264), 10, p, p, 1, None) == 1
    assert lib.alpaca_mr_jump_halo_update(1, p, 10, p, p, 1, None) == 1
    assert lib.alpaca_mr_jump_halo_update(2, p, 10, None, None, 0, None) == 0
    eps = (ctypes.c_double * 1)(1e-3)
    assert lib.alpaca_mr_child_needs_remeshing(3, p, big, p, 1, 0, eps, 1, p, None, None) == 1
    # misaligned pools and unknown norms are refused as well
    assert lib.alpaca_mr_average(3, ctypes.c_void_p(264), 10, p, 1, None) == 1
    assert lib.alpaca_mr_child_needs_remeshing(3, p, 10, p, 1, 5, eps, 1, p, None, None) == 1
    assert lib.alpaca_ensemble_bytes(2, big, 4) == 0
    out = ctypes.c_void_p()
    assert lib.alpaca_ensemble_create(2, big, 4, 1, None, ctypes.byref(out)) == 1
    # no links: nothing to do, and no device call is made
    assert lib.alpaca_mr_average(3, p, 10, None, 0, None) == 0


def test_grid_neighbour_table():
    """alpaca_fill_halos_table's table for a grid: offsets reciprocate
    (B at +o of A means A at -o of B), periodic directions wrap, other
    directions end in -1, and the centre slot is the block itself."""
    import numpy as np

    for dim, nb, bc in ((2, [3, 2], [gpu.BC_PERIODIC, gpu.BC_ZERO_GRADIENT]),
                        (3, [2, 3, 4], [gpu.BC_SYMMETRY, gpu.BC_PERIODIC, gpu.BC_ZERO_GRADIENT])):
        t = gpu.grid_neighbors(dim, nb, bc)
        slots = t.shape[1]
        assert slots == (27 if dim == 3 else 9)
        centre = slots // 2
        assert np.array_equal(t[:, centre], np.arange(t.shape[0]))
        for a in range(t.shape[0]):
            for s in range(slots):
                b = t[a, s]
                if b >= 0:
                    assert t[b, slots - 1 - s] == a
        # a block on a non-periodic side has no neighbour beyond it
        nb3 = list(nb) + [1] * (3 - len(nb))
        for ax in range(dim):
            if bc[ax] == gpu.BC_PERIODIC:
                assert (t >= 0).sum() > 0
                continue
            off = [0, 0, 0]
            off[ax] = -1
            slot = ((off[0] + 1) * 3 + (off[1] + 1)) * 3 + (off[2] + 1) if dim == 3 else \
                (off[0] + 1) * 3 + (off[1] + 1)
            for k in range(t.shape[0]):
                pos = [k // (nb3[1] * nb3[2]), (k // nb3[2]) % nb3[1], k % nb3[2]]
                assert (t[k, slot] < 0) == (pos[ax] == 0)


def test_noalloc_and_exact_ops_refuse_before_any_device_call():
    """The no-allocation right-hand side refuses block counts above its fixed
    scratch (128 3D / 3855 2D blocks) with ALPACA_ERR_CAPACITY, and bad
    arguments with ALPACA_ERR_INVALID, before any HIP call; so does the exact
    FP64 operations hook."""
    lib = gpu.load_library()
    p, q = ctypes.c_void_p(256), ctypes.c_void_p(512)
    d3 = gpu.make_desc(3, [1, 1, 1], [0, 0, 0], cell_size=1.0 / 128)
    d2 = gpu.make_desc(2, [1, 1], [0, 0], cell_size=1.0 / 128)
    assert lib.alpaca_right_hand_side_noalloc(ctypes.byref(d3), 129, p, q, None) == 5
    assert lib.alpaca_right_hand_side_noalloc(ctypes.byref(d2), 3856, p, q, None) == 5
    assert lib.alpaca_right_hand_side_noalloc(ctypes.byref(d3), 0, p, q, None) == 1
    assert lib.alpaca_right_hand_side_noalloc(ctypes.byref(d3), 1, p, p, None) == 1
    bad = gpu.make_desc(3, [1, 1, 1], [0, 0, 0])
    bad.dim = 5
    assert lib.alpaca_right_hand_side_noalloc(ctypes.byref(bad), 1, p, q, None) == 1
    assert lib.alpaca_exact_fp64_ops(5, p, p, q, 10, None) == 1
    assert lib.alpaca_exact_fp64_ops(-1, p, p, q, 10, None) == 1
    assert lib.alpaca_exact_fp64_ops(0, p, None, q, 10, None) == 1
    assert lib.alpaca_exact_fp64_ops(3, p, None, q, 10, None) == 1
    assert lib.alpaca_exact_fp64_ops(0, p, p, q, -1, None) == 1
    assert lib.alpaca_exact_fp64_ops(2, p, None, q, 0, None) == 0
    assert lib.alpaca_exact_fp64_ops(4, p, None, q, 0, None) == 0


def test_pool_list_entries_refuse_before_any_device_call():
    """The block-list forms over a node pool (right-hand side with the jump
    fluxes, buffer swap and copy, CFL reduction, jump-halo integration)
    refuse bad arguments with ALPACA_ERR_INVALID, and an empty list is done
    without a device call."""
    lib = gpu.load_library()
    p, q, r, lst = (ctypes.c_void_p(v) for v in (256, 512, 768, 1024))
    ctx = ctypes.c_void_p(2048)  # never dereferenced on these paths
    d3 = gpu.make_desc(3, [4, 1, 1], [0, 0, 0], cell_size=1.0 / 64)
    rhs = lib.alpaca_ctx_right_hand_side_jump_fluxes_list
    assert rhs(None, ctypes.byref(d3), 8, lst, 4, p, q, r) == 1      # no context
    assert rhs(ctx, ctypes.byref(d3), 0, lst, 4, p, q, r) == 1       # no pool
    assert rhs(ctx, ctypes.byref(d3), 1 << 25, lst, 4, p, q, r) == 1
    assert rhs(ctx, ctypes.byref(d3), 8, None, 4, p, q, r) == 1      # no list
    assert rhs(ctx, ctypes.byref(d3), 8, lst, -1, p, q, r) == 1
    assert rhs(ctx, ctypes.byref(d3), 8, lst, 4, p, p, r) == 1       # avg is rhs
    assert rhs(ctx, ctypes.byref(d3), 8, lst, 4, p, q, None) == 1
    assert rhs(ctx, ctypes.byref(d3), 8, None, 0, p, q, r) == 0      # nothing to do
    swap = lib.alpaca_swap_blocks_list
    assert swap(4, p, q, 8, lst, 4, None) == 1
    assert swap(3, p, p, 8, lst, 4, None) == 1
    assert swap(3, p, q, 0, lst, 4, None) == 1
    assert swap(3, p, q, 8, None, 4, None) == 1
    assert swap(3, p, q, 8, lst, 1 << 25, None) == 1
    assert swap(2, p, q, 8, None, 0, None) == 0
    copy = lib.alpaca_copy_blocks_list
    assert copy(3, p, p, 8, lst, 4, None) == 1
    assert copy(3, None, q, 8, lst, 4, None) == 1
    assert copy(3, p, q, 8, lst, -1, None) == 1
    assert copy(3, p, q, 8, None, 0, None) == 0
    jh = lib.alpaca_lts_integrate_jump_halos
    assert jh(4, p, q, 8, lst, 4, 1e-3, None) == 1
    assert jh(3, None, q, 8, lst, 4, 1e-3, None) == 1
    assert jh(3, p, q, 0, lst, 4, 1e-3, None) == 1
    assert jh(3, p, q, 8, None, 4, 1e-3, None) == 1
    assert jh(3, p, q, 8, None, 0, 1e-3, None) == 0
    ts = lib.alpaca_timestep_list
    assert ts(ctypes.byref(d3), p, 8, None, 4, q, None) == 1
    assert ts(ctypes.byref(d3), p, 8, lst, 0, q, None) == 1
    assert ts(ctypes.byref(d3), p, 0, lst, 4, q, None) == 1
    assert ts(ctypes.byref(d3), p, 8, lst, 4, None, None) == 1
    bad = gpu.make_desc(3, [4, 1, 1], [0, 0, 0])
    bad.gamma = 1.0
    assert ts(ctypes.byref(bad), p, 8, lst, 4, q, None) == 1
