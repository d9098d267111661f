"""Multiresolution operators on a device pool of haloed blocks.

Host side of the C ABI's alpaca_mr_* entry points, named after the
reference's classes: ``ThresholdComputer`` (threshold_computer.h:35-97) and
``Multiresolution`` (multiresolution.h:59-188) with ``average``
(Multiresolution::Average), ``average_jump_buffer`` (AverageJumpBuffer),
``prediction`` (Multiresolution::Prediction) and
``child_needs_remeshing`` (ChildNeedsRemeshing<Norm::Linfinity> with
RemeshingDecision).  Where the reference works on one parent/child pair per
call, these take a batch of links -- (child block, parent block, sibling
position, child level) rows of an int32 device tensor -- and run one launch
over all of them.  Blocks are the pool's [nblocks][NE][24][24][24|1] float64
device tensor.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import gpu

# RemeshIdentifier (remesh_identifier.h:29).
NEUTRAL, REFINE, COARSE = 0, 1, 2
# Norm of the wavelet details (norms.h); the reference's default
# CC::NFWA() is L-infinity (compile_time_constants.h:133).
LINF, L1, L2 = 0, 1, 2
MAX_LEVELS = 32

# Compile-time constants of the threshold: STDO, the limiting convergence
# order of space and time (compile_time_constants.h:617), and the refinement
# factor 2^(p+1) of the fifth-order prediction (multiresolution.cpp:412-418).
STDO = 3
REFINE_FACTOR = 32.0


class ThresholdComputer:
    """Level-dependent coarsening threshold (threshold_computer.h:35-97):
    eps_ref = eps_user * 2^(-(STDO + 1)(maximum_level - reference_level)),
    ThresholdOnLevel(l) = eps_ref * 2^(-dim (maximum_level - l))."""

    def __init__(self, dim: int, maximum_level: int, user_reference_level: int,
                 user_reference_epsilon: float, convergence_order: int = STDO):
        self.dim = int(dim)
        self.maximum_level = int(maximum_level)
        self.reference_epsilon = float(user_reference_epsilon) * 2.0 ** (
            -1.0 * (float(convergence_order) + 1.0) *
            (float(maximum_level) - float(user_reference_level)))

    def threshold_on_level(self, level: int) -> float:
        return self.reference_epsilon * 2.0 ** (
            -1.0 * float(self.dim) * (float(self.maximum_level) - float(level)))

    def table(self, nlevels: int | None = None) -> list:
        n = self.maximum_level + 1 if nlevels is None else nlevels
        return [self.threshold_on_level(level) for level in range(n)]


def remeshing_decision(detail: float, eps_coarsen: float) -> int:
    """RemeshingDecision (multiresolution.cpp:408-428) on the host."""
    if detail <= eps_coarsen:
        return COARSE
    if detail >= eps_coarsen * REFINE_FACTOR:
        return REFINE
    return NEUTRAL


# BoundaryLocation (boundary_specifications.h:56-86) in enum order, with the
# per-axis direction each name gives: East/North/Top +1, West/South/Bottom -1.
LOCATIONS = ("East", "West", "North", "South", "Top", "Bottom",
             "BottomNorth", "BottomSouth", "TopNorth", "TopSouth",
             "BottomEast", "BottomWest", "TopEast", "TopWest",
             "NorthEast", "NorthWest", "SouthEast", "SouthWest",
             "EastNorthTop", "EastNorthBottom", "EastSouthTop", "EastSouthBottom",
             "WestNorthTop", "WestNorthBottom", "WestSouthTop", "WestSouthBottom")
_AXIS_WORDS = (("East", "West"), ("North", "South"), ("Top", "Bottom"))


def location_direction(location: int) -> tuple:
    """Per-axis direction (+1, -1 or 0) of a BoundaryLocation, from its name."""
    name = LOCATIONS[location]
    return tuple(1 if hi in name else (-1 if lo in name else 0) for hi, lo in _AXIS_WORDS)


def halo_recv_range(dim: int, location: int) -> tuple:
    """(x_start, x_count, y_start, y_count[, z_start, z_count]) of the halo slab
    at a location: GetStartIndicesHaloRecv and GetHaloSize
    (communication_types.h:76-132,190-197)."""
    d = location_direction(location)
    if dim == 2 and d[2] != 0:
        raise gpu.AlpacaError(f"location {LOCATIONS[location]} has no halo in 2D")
    out = []
    for a in range(dim):
        if d[a] > 0:
            out += [gpu.IC + gpu.HS, gpu.HS]
        elif d[a] < 0:
            out += [0, gpu.HS]
        else:
            out += [gpu.HS, gpu.IC]
    return tuple(out)


def make_links(rows) -> np.ndarray:
    """int32 [n][4] links from (child, parent, position, level) rows."""
    a = np.asarray(rows if isinstance(rows, np.ndarray) else list(rows),
                   dtype=np.int32).reshape(-1, 4)
    return np.ascontiguousarray(a)


def _check_pool(blocks, dim):
    if blocks.dtype != gpu._torch().float64 or not blocks.is_contiguous():
        raise gpu.AlpacaError("blocks must be a contiguous float64 device tensor")
    n = blocks.numel() // gpu.block_doubles(dim)
    if n * gpu.block_doubles(dim) != blocks.numel() or n < 1:
        raise gpu.AlpacaError("blocks is not a whole number of haloed blocks")
    return n


def _check_links(links):
    torch = gpu._torch()
    if links.dtype != torch.int32 or not links.is_contiguous() or links.dim() != 2 \
            or links.shape[1] != 4:
        raise gpu.AlpacaError("links must be a contiguous int32 [n][4] device tensor")
    return links.shape[0]


def _stream(stream):
    return gpu._stream_ptr() if stream is None else stream


class Multiresolution:
    """Batched Multiresolution operators; ``thresholder`` (a ThresholdComputer)
    and ``norm`` (LINF, L1, L2: the template argument of ChildNeedsRemeshing)
    are used by child_needs_remeshing only."""

    def __init__(self, thresholder: ThresholdComputer | None = None, norm: int = LINF):
        if norm not in (LINF, L1, L2):
            raise gpu.AlpacaError(f"unknown norm {norm}")
        self.thresholder = thresholder
        self.norm = norm

    @staticmethod
    def average(blocks, dim, links, stream=None):
        """Averages every link's child into its parent's octant (quadrant)."""
        lib = gpu.load_library()
        nb, nl = _check_pool(blocks, dim), _check_links(links)
        gpu.check(lib.alpaca_mr_average(dim, ctypes.c_void_p(blocks.data_ptr()), nb,
                                        ctypes.c_void_p(links.data_ptr()), nl, _stream(stream)),
                  "mr_average")

    @staticmethod
    def average_jump_buffer(surfaces, dim, links, stream=None):
        """AverageJumpBuffer for every link: surfaces is the pool's float64
        device tensor of [nblocks][6][NE][16][16|1] jump surfaces; the
        children's averaged face values are added into their parents'."""
        torch = gpu._torch()
        lib = gpu.load_library()
        per = 6 * (dim + 2) * gpu.IC * (gpu.IC if dim == 3 else 1)
        if surfaces.dtype != torch.float64 or not surfaces.is_contiguous() \
                or surfaces.numel() % per or surfaces.numel() == 0:
            raise gpu.AlpacaError("surfaces must be a contiguous float64 tensor of whole "
                                  "jump-surface buffers")
        nl = _check_links(links)
        gpu.check(lib.alpaca_mr_average_jump_buffers(dim, ctypes.c_void_p(surfaces.data_ptr()),
                                                     surfaces.numel() // per,
                                                     ctypes.c_void_p(links.data_ptr()), nl,
                                                     _stream(stream)), "mr_average_jump_buffers")

    @staticmethod
    def prediction(blocks, dim, links, rng=None, stream=None):
        """Predicts every link's child cells in rng = (x_start, x_count,
        y_start, y_count[, z_start, z_count]) from its parent (default: the
        whole block, as ChildNeedsRemeshing uses it)."""
        lib = gpu.load_library()
        nb, nl = _check_pool(blocks, dim), _check_links(links)
        if rng is None:
            rng = (0, gpu.TC) * dim
        r = (ctypes.c_int * 6)(*(list(rng) + [0, 1] * (3 - len(rng) // 2)))
        gpu.check(lib.alpaca_mr_prediction(dim, ctypes.c_void_p(blocks.data_ptr()), nb,
                                           ctypes.c_void_p(links.data_ptr()), nl, r,
                                           _stream(stream)), "mr_prediction")

    @staticmethod
    def jump_halo_update(blocks, dim, links, locations, stream=None):
        """Halo update across a level jump (UpdateMaterialJumpNoMpi,
        internal_halo_manager.cpp:231-262): every link's child gets its halo
        slab at locations[i] (int32 device tensor, BoundaryLocation order)
        predicted from its parent."""
        torch = gpu._torch()
        lib = gpu.load_library()
        nb, nl = _check_pool(blocks, dim), _check_links(links)
        if locations.dtype != torch.int32 or not locations.is_contiguous() \
                or locations.numel() != nl:
            raise gpu.AlpacaError("locations must be a contiguous int32 device tensor, one per link")
        gpu.check(lib.alpaca_mr_jump_halo_update(dim, ctypes.c_void_p(blocks.data_ptr()), nb,
                                                 ctypes.c_void_p(links.data_ptr()),
                                                 ctypes.c_void_p(locations.data_ptr()), nl,
                                                 _stream(stream)), "mr_jump_halo_update")

    def child_needs_remeshing(self, blocks, dim, links, stream=None, out=None):
        """(details float64[n], decisions int32[n]) device tensors; decisions
        are RemeshIdentifier values, -1 for a skipped link.  `out` may hand
        in the two tensors to fill."""
        torch = gpu._torch()
        if self.thresholder is None:
            raise gpu.AlpacaError("child_needs_remeshing needs a ThresholdComputer")
        lib = gpu.load_library()
        nb, nl = _check_pool(blocks, dim), _check_links(links)
        table = self.thresholder.table()
        if len(table) > MAX_LEVELS:
            raise gpu.AlpacaError(f"at most {MAX_LEVELS} levels")
        eps = (ctypes.c_double * len(table))(*table)
        if out is None:
            details = torch.empty(nl, dtype=torch.float64, device=blocks.device)
            decisions = torch.empty(nl, dtype=torch.int32, device=blocks.device)
        else:
            details, decisions = out
            if details.dtype != torch.float64 or decisions.dtype != torch.int32 \
                    or details.numel() < nl or decisions.numel() < nl:
                raise gpu.AlpacaError("out: float64 and int32 tensors of at least n entries")
        gpu.check(lib.alpaca_mr_child_needs_remeshing(
            dim, ctypes.c_void_p(blocks.data_ptr()), nb, ctypes.c_void_p(links.data_ptr()), nl,
            self.norm, eps, len(table), ctypes.c_void_p(details.data_ptr()),
            ctypes.c_void_p(decisions.data_ptr()), _stream(stream)), "mr_child_needs_remeshing")
        return details, decisions
