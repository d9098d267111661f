"""Level-wise (local time stepping) integration on a device pool of blocks.

Host side of the alpaca_lts_* entry points, shaped like the reference's
TimeIntegrator<RungeKutta3TVD> (time_integrator.h:240-526): micro time steps
are appended on the finest level, a level that advances with n of them
integrates with their sum (newest first, from 0.0) times the stage
multiplier, and a macro step ends by adding all micro steps to the run time.
The per-node operations -- FillInitialBuffer, PrepareBufferForIntegration,
IntegrateNode (jump conservatives and internal cells) and IntegrateJumpHalos
-- run over a list of pool blocks per launch.  Pools are float64 device
tensors of haloed blocks ([nblocks][NE][24][24][24|1]); jump surfaces are
[nblocks][6][NE][16][16|1] (SurfaceBuffer layout).
"""
from __future__ import annotations

import ctypes

from . import gpu

NUMBER_OF_STAGES = 3


def _list(block_list):
    torch = gpu._torch()
    if block_list.dtype != torch.int32 or not block_list.is_contiguous() or block_list.dim() != 1:
        raise gpu.AlpacaError("block list must be a contiguous int32 device tensor")
    return block_list.shape[0]


def _pool(t, dim):
    torch = gpu._torch()
    per = gpu.block_doubles(dim)
    if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() % per or t.numel() == 0:
        raise gpu.AlpacaError("pool must be a contiguous float64 tensor of whole blocks")
    return t.numel() // per


def _stream(stream):
    return gpu._stream_ptr() if stream is None else stream


class LevelTimeIntegrator:
    """TimeIntegrator<RungeKutta3TVD> bookkeeping plus the device operations."""

    def __init__(self, dim: int, start_time: float = 0.0):
        self.dim = int(dim)
        self.current_run_time = float(start_time)
        self.micro_timestep_sizes: list[float] = []

    # -- time bookkeeping (time_integrator.h:240-260)
    def append_micro_timestep(self, dt: float):
        self.micro_timestep_sizes.append(float(dt))

    def finish_macro_timestep(self):
        s = 0.0
        for dt in self.micro_timestep_sizes:
            s += dt
        self.current_run_time += s
        self.micro_timestep_sizes.clear()

    def timestep(self, stage: int, number_of_timesteps: int, jump: bool = False) -> float:
        """The last n micro steps summed from the newest, times the stage
        multiplier (of the jump conservatives with jump=True)."""
        lib = gpu.load_library()
        m = (ctypes.c_double * max(len(self.micro_timestep_sizes), 1))(*self.micro_timestep_sizes)
        out = ctypes.c_double()
        gpu.check(lib.alpaca_lts_timestep(m, len(self.micro_timestep_sizes), number_of_timesteps,
                                          stage, int(bool(jump)), ctypes.byref(out)),
                  "lts_timestep")
        return out.value

    # -- per-node operations over a list of pool blocks
    def fill_initial_buffer(self, average, initial, block_list, stage: int, stream=None):
        """FillInitialBuffer (time_integrator.h:369-406), single phase: the
        Average buffers of the listed blocks copied to Initial at stage 0."""
        if stage != 0:
            return
        nb = _pool(average, self.dim)
        if _pool(initial, self.dim) != nb:
            raise gpu.AlpacaError("average and initial pools differ in size")
        n = _list(block_list)
        lib = gpu.load_library()
        gpu.check(lib.alpaca_copy_blocks_list(self.dim, ctypes.c_void_p(average.data_ptr()),
                                              ctypes.c_void_p(initial.data_ptr()), nb,
                                              ctypes.c_void_p(block_list.data_ptr()), n,
                                              _stream(stream)), "copy_blocks_list")

    def prepare_buffer_for_integration(self, average, initial, block_list, stage: int, stream=None):
        lib = gpu.load_library()
        nb = _pool(average, self.dim)
        gpu.check(lib.alpaca_lts_prepare_buffer(self.dim, ctypes.c_void_p(average.data_ptr()),
                                                ctypes.c_void_p(initial.data_ptr()), nb,
                                                ctypes.c_void_p(block_list.data_ptr()),
                                                _list(block_list), stage, _stream(stream)),
                  "lts_prepare_buffer")

    def integrate_node(self, average, rhs, jump_cons, jump_flux, block_list, stage: int,
                       number_of_timesteps: int, stream=None):
        """IntegrateNode (time_integrator.h:312-340): the jump conservatives
        with the jump multiplier, then the internal cells rhs = average +
        timestep * rhs."""
        lib = gpu.load_library()
        nb = _pool(average, self.dim)
        n = _list(block_list)
        st = _stream(stream)
        if jump_cons is not None:
            gpu.check(lib.alpaca_lts_integrate_jump_conservatives(
                self.dim, ctypes.c_void_p(jump_cons.data_ptr()), ctypes.c_void_p(jump_flux.data_ptr()),
                nb, ctypes.c_void_p(block_list.data_ptr()), n,
                self.timestep(stage, number_of_timesteps, jump=True), st),
                "lts_integrate_jump_conservatives")
        gpu.check(lib.alpaca_lts_integrate(self.dim, ctypes.c_void_p(average.data_ptr()),
                                           ctypes.c_void_p(rhs.data_ptr()), nb,
                                           ctypes.c_void_p(block_list.data_ptr()), n, None,
                                           self.timestep(stage, number_of_timesteps), st),
                  "lts_integrate")

    def integrate_jump_halos(self, average, rhs, block_list, stage: int, number_of_timesteps: int,
                             start_indices_halo, halo_size, stream=None):
        """IntegrateJumpHalos (time_integrator.h:281-301): the halo range of
        the listed blocks integrated with the conservatives' time step."""
        lib = gpu.load_library()
        nb = _pool(average, self.dim)
        r = (ctypes.c_int * 6)(start_indices_halo[0], halo_size[0], start_indices_halo[1],
                               halo_size[1], start_indices_halo[2], halo_size[2])
        gpu.check(lib.alpaca_lts_integrate(self.dim, ctypes.c_void_p(average.data_ptr()),
                                           ctypes.c_void_p(rhs.data_ptr()), nb,
                                           ctypes.c_void_p(block_list.data_ptr()),
                                           _list(block_list), r,
                                           self.timestep(stage, number_of_timesteps),
                                           _stream(stream)),
                  "lts_integrate (halo)")


def jump_flux_adjustment(dim, cells, jump_cons, entries, one_cell_size, stream=None):
    """JumpFluxAdjustment's exchange on one level's leaves; entries is an
    int32 [n][7] device tensor (leaf, then the non-leaf neighbour per face
    location East, West, North, South, Top, Bottom or -1)."""
    torch = gpu._torch()
    lib = gpu.load_library()
    if entries.dtype != torch.int32 or not entries.is_contiguous() or entries.dim() != 2 \
            or entries.shape[1] != 7:
        raise gpu.AlpacaError("entries must be a contiguous int32 [n][7] device tensor")
    nb = _pool(cells, dim)
    gpu.check(lib.alpaca_lts_jump_flux_adjustment(dim, ctypes.c_void_p(cells.data_ptr()),
                                                  ctypes.c_void_p(jump_cons.data_ptr()), nb,
                                                  ctypes.c_void_p(entries.data_ptr()),
                                                  entries.shape[0], float(one_cell_size),
                                                  _stream(stream)), "lts_jump_flux_adjustment")


def reset_jump_conservatives(dim, jump_cons, block_list, stream=None):
    """ResetJumpConservativeBuffers on the listed blocks."""
    lib = gpu.load_library()
    per = 6 * (dim + 2) * gpu.IC * (gpu.IC if dim == 3 else 1)
    gpu.check(lib.alpaca_lts_reset_jump_conservatives(dim, ctypes.c_void_p(jump_cons.data_ptr()),
                                                      jump_cons.numel() // per,
                                                      ctypes.c_void_p(block_list.data_ptr()),
                                                      _list(block_list), _stream(stream)),
              "lts_reset_jump_conservatives")


def levels_to_update(timestep: int, maximum_level: int) -> list:
    """GetLevels (modular_algorithm_assembler.cpp:2563-2572): the levels that
    advance in micro step `timestep`, in descending order from the finest:
    as many as bits are set in (timestep + 1) ^ timestep (a 16-bit bitset in
    the reference)."""
    height = bin(((timestep + 1) ^ timestep) & 0xFFFF).count("1")
    return list(range(maximum_level, maximum_level - height, -1))
