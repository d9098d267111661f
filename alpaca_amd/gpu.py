"""Python mirror of the C ABI in include/alpaca_gpu.h (ctypes).

PyTorch only provides device memory and the stream; every computation is a
kernel of ``libalpaca_hip.so``.  Loading fails loudly when the library is not
built: there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalpaca_hip.so")

FLUX_HLLC, FLUX_ROE = 0, 1
FP64, MIXED, FP64_FMA = 0, 1, 2
BC_ZERO_GRADIENT, BC_PERIODIC, BC_SYMMETRY = 0, 1, 2
IC, HS, TC = 16, 4, 24

KERNEL_FACE_FLUX, KERNEL_COMBINE, KERNEL_HALO, KERNEL_TIMESTEP, KERNEL_STAGE = 0, 1, 2, 3, 4

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "non-finite state",
          4: "unsupported configuration", 5: "capacity exhausted"}


class CaseDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nb", ctypes.c_int * 3), ("bc", ctypes.c_int * 3),
                ("flux", ctypes.c_int), ("precision", ctypes.c_int), ("gamma", ctypes.c_double),
                ("background_pressure", ctypes.c_double), ("cell_size", ctypes.c_double)]


class CaseState(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("t_end", ctypes.c_double), ("cfl", ctypes.c_double),
                ("dt", ctypes.c_double), ("smax", ctypes.c_uint64), ("steps", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class AlpacaError(RuntimeError):
    pass


_lib = None

_SIGNATURES = {
    "alpaca_status_string": ([ctypes.c_int], ctypes.c_char_p),
    "alpaca_abi_version": ([], ctypes.c_int),
    "alpaca_block_doubles": ([ctypes.c_int], ctypes.c_size_t),
    "alpaca_convective_fluxes": ([ctypes.POINTER(CaseDesc), ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_right_hand_side": ([ctypes.POINTER(CaseDesc), ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_right_hand_side_jump_fluxes": ([ctypes.POINTER(CaseDesc), ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
                                           ctypes.c_int),
    "alpaca_right_hand_side_noalloc": ([ctypes.POINTER(CaseDesc), ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_exact_fp64_ops": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "alpaca_context_create": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "alpaca_context_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "alpaca_context_reserve": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "alpaca_ctx_convective_fluxes": ([ctypes.c_void_p, ctypes.POINTER(CaseDesc), ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_ctx_right_hand_side": ([ctypes.c_void_p, ctypes.POINTER(CaseDesc), ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_ctx_right_hand_side_jump_fluxes": ([ctypes.c_void_p, ctypes.POINTER(CaseDesc),
                                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p], ctypes.c_int),
    "alpaca_ctx_right_hand_side_jump_fluxes_list": ([ctypes.c_void_p, ctypes.POINTER(CaseDesc),
                                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                     ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p], ctypes.c_int),
    "alpaca_swap_blocks_list": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "alpaca_copy_blocks_list": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "alpaca_rk3_stage": ([ctypes.POINTER(CaseDesc), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_fill_halos_table": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_void_p], ctypes.c_int),
    "alpaca_fill_halos_list": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "alpaca_fill_halos": ([ctypes.POINTER(CaseDesc), ctypes.c_void_p, ctypes.c_void_p],
                          ctypes.c_int),
    "alpaca_timestep": ([ctypes.POINTER(CaseDesc), ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p], ctypes.c_int),
    "alpaca_timestep_list": ([ctypes.POINTER(CaseDesc), ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
                             ctypes.c_int),
    "alpaca_finish_step": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "alpaca_case_state_init": ([ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_void_p], ctypes.c_int),
    "alpaca_profile_enable": ([ctypes.c_int], ctypes.c_int),
    "alpaca_profile_read": ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_double)],
                            ctypes.c_int),
    "alpaca_ensemble_bytes": ([ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_size_t),
    "alpaca_ensemble_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "alpaca_ensemble_create_shared": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
                                      ctypes.c_int),
    "alpaca_ensemble_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "alpaca_ensemble_add_case": ([ctypes.c_void_p, ctypes.POINTER(CaseDesc), ctypes.c_void_p,
                                  ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int)],
                                 ctypes.c_int),
    "alpaca_ensemble_remove_case": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "alpaca_ensemble_cancel_case": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "alpaca_ensemble_set_tuning": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "alpaca_ensemble_step": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "alpaca_ensemble_read_states": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int],
                                    ctypes.c_int),
    "alpaca_ensemble_case_blocks": ([ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p),
                                     ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "alpaca_ensemble_copy_case": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
                                  ctypes.c_int),
    "alpaca_ensemble_capacity": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
                                 ctypes.c_int),
    "alpaca_advance": ([ctypes.POINTER(CaseDesc), ctypes.c_void_p, ctypes.c_void_p,
                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
                       ctypes.c_int),
    "alpaca_mr_average": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                           ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "alpaca_mr_average_jump_buffers": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
                                       ctypes.c_int),
    "alpaca_lts_timestep": ([ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                            ctypes.c_int),
    "alpaca_lts_integrate": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                              ctypes.c_double, ctypes.c_void_p], ctypes.c_int),
    "alpaca_lts_integrate_jump_halos": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_void_p], ctypes.c_int),
    "alpaca_lts_integrate_jump_conservatives": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_void_p], ctypes.c_int),
    "alpaca_lts_prepare_buffer": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
                                  ctypes.c_int),
    "alpaca_lts_jump_flux_adjustment": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_void_p], ctypes.c_int),
    "alpaca_lts_reset_jump_conservatives": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
                                            ctypes.c_int),
    "alpaca_mr_prediction": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                              ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p],
                             ctypes.c_int),
    "alpaca_mr_jump_halo_update": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "alpaca_mr_child_needs_remeshing": ([ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
                                        ctypes.c_int),
}


def load_library(path: str = LIB_PATH):
    """Load libalpaca_hip.so and declare the C ABI; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AlpacaError(f"{path} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    for name, (argtypes, restype) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGNATURES)


def check(status: int, what: str = ""):
    if status != 0:
        raise AlpacaError(f"{what}: {STATUS.get(status, status)}")


def make_desc(dim, nb, bc, flux=FLUX_HLLC, precision=FP64, gamma=1.4, b=0.0, cell_size=None,
              length=1.0):
    nb3 = list(nb) + [1] * (3 - len(nb))
    bc3 = list(bc) + [0] * (3 - len(bc))
    if dim == 2:
        nb3[2] = 1
        bc3[2] = 0
    if cell_size is None:
        cell_size = length / (nb3[0] * IC)
    return CaseDesc(dim, (ctypes.c_int * 3)(*nb3), (ctypes.c_int * 3)(*bc3), flux, precision,
                    gamma, b, cell_size)


def _torch():
    import torch
    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def block_doubles(dim):
    return NE(dim) * TC ** dim


def NE(dim):
    return dim + 2


def to_blocks(torch_global, dim, nb):
    """Domain array [NE][NX][NY][NZ] (device) -> haloed blocks (device), halos zero."""
    torch = _torch()
    ne = NE(dim)
    if dim == 2:
        g = torch_global.reshape(ne, nb[0], IC, nb[1], IC)
        g = g.permute(1, 3, 0, 2, 4)
        blocks = torch.zeros((nb[0], nb[1], ne, TC, TC), dtype=torch.float64,
                             device=torch_global.device)
        blocks[..., HS:HS + IC, HS:HS + IC] = g
        return blocks.reshape(nb[0] * nb[1], ne, TC, TC)
    g = torch_global.reshape(ne, nb[0], IC, nb[1], IC, nb[2], IC)
    g = g.permute(1, 3, 5, 0, 2, 4, 6)
    blocks = torch.zeros((nb[0], nb[1], nb[2], ne, TC, TC, TC), dtype=torch.float64,
                         device=torch_global.device)
    blocks[..., HS:HS + IC, HS:HS + IC, HS:HS + IC] = g
    return blocks.reshape(nb[0] * nb[1] * nb[2], ne, TC, TC, TC)


def from_blocks(blocks, dim, nb):
    """Haloed blocks (device) -> domain array [NE][NX][NY][NZ] (device)."""
    ne = NE(dim)
    if dim == 2:
        b = blocks.reshape(nb[0], nb[1], ne, TC, TC)[..., HS:HS + IC, HS:HS + IC]
        return b.permute(2, 0, 3, 1, 4).reshape(ne, nb[0] * IC, nb[1] * IC, 1)
    b = blocks.reshape(nb[0], nb[1], nb[2], ne, TC, TC, TC)[..., HS:HS + IC, HS:HS + IC,
                                                            HS:HS + IC]
    return b.permute(3, 0, 4, 1, 5, 2, 6).reshape(ne, nb[0] * IC, nb[1] * IC, nb[2] * IC)


def grid_neighbors(dim, nb, bc):
    """Neighbour table of a block grid for alpaca_fill_halos_table:
    int32[nblocks][27] (2D: [nblocks][9]), blocks in this package's order
    (bx * ny + by) * nz + bz, slot ((dx+1)*3 + (dy+1))*3 + (dz+1) (2D:
    (dx+1)*3 + (dy+1)); -1 where an offset leaves the domain across a
    non-periodic side, wrapped across periodic ones (bc per direction)."""
    nb3 = [int(x) for x in nb] + [1] * (3 - len(nb))
    if dim == 2:
        nb3[2] = 1
    offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1)
            for dz in ((-1, 0, 1) if dim == 3 else (0,))]
    table = np.full((nb3[0] * nb3[1] * nb3[2], len(offs)), -1, dtype=np.int32)
    for bx in range(nb3[0]):
        for by in range(nb3[1]):
            for bz in range(nb3[2]):
                me = (bx * nb3[1] + by) * nb3[2] + bz
                for slot, off in enumerate(offs):
                    q = [bx + off[0], by + off[1], bz + off[2]]
                    ok = True
                    for a in range(dim):
                        if 0 <= q[a] < nb3[a]:
                            continue
                        if bc[a] == BC_PERIODIC:
                            q[a] %= nb3[a]
                        else:
                            ok = False
                    if ok:
                        table[me, slot] = (q[0] * nb3[1] + q[1]) * nb3[2] + q[2]
    return table


def fill_halos_table(blocks, dim, neighbors, side_bc, stream=None):
    """alpaca_fill_halos_table on device tensors: blocks (float64, haloed
    blocks), neighbors (int32, as grid_neighbors returns); side_bc: the
    condition of each external side (-x, +x, -y, +y, (-z, +z))."""
    lib = load_library()
    bc = (ctypes.c_int * 6)(*(list(side_bc) + [BC_ZERO_GRADIENT] * (6 - len(side_bc))))
    nblocks = neighbors.shape[0]
    st = _stream_ptr() if stream is None else stream
    check(lib.alpaca_fill_halos_table(dim, ctypes.c_void_p(blocks.data_ptr()), nblocks,
                                      ctypes.c_void_p(neighbors.data_ptr()), bc, st),
          "fill_halos_table")


class GpuCase:
    """One uniform case resident in HBM: state u (+ two stage buffers) and
    its device record.  Mirrors the reference's ModularAlgorithmAssembler
    ComputeLoop/Advance for a single level (modular_algorithm_assembler.cpp:
    97-195, 543-910)."""

    def __init__(self, desc: CaseDesc, state0: np.ndarray, cfl: float, t_end: float = 1e300,
                 device="cuda", time0: float = 0.0):
        torch = _torch()
        self.lib = load_library()
        self.desc = desc
        self.dim = desc.dim
        self.nb = [desc.nb[0], desc.nb[1], desc.nb[2] if desc.dim == 3 else 1]
        self.nblocks = self.nb[0] * self.nb[1] * self.nb[2]
        if torch.is_tensor(state0):
            g = state0.to(device=device, dtype=torch.float64).contiguous()
        else:
            g = torch.as_tensor(np.ascontiguousarray(state0), device=device)
        self.u = to_blocks(g, self.dim, self.nb).contiguous()
        self.s1 = torch.zeros_like(self.u)
        self.s2 = torch.zeros_like(self.u)
        self.state = torch.zeros(ctypes.sizeof(CaseState), dtype=torch.uint8, device=device)
        self.cfl = cfl
        self.t_end = t_end
        st = self._stream()
        check(self.lib.alpaca_fill_halos(ctypes.byref(desc), ctypes.c_void_p(self.u.data_ptr()), st),
              "fill_halos")
        check(self.lib.alpaca_case_state_init(ctypes.c_void_p(self.state.data_ptr()), time0,
                                              t_end, cfl, st), "case_state_init")

    def _stream(self):
        return _stream_ptr()

    def advance(self, nsteps: int = 1):
        st = self._stream()
        d = ctypes.byref(self.desc)
        u, s1, s2 = (ctypes.c_void_p(t.data_ptr()) for t in (self.u, self.s1, self.s2))
        stp = ctypes.c_void_p(self.state.data_ptr())
        check(self.lib.alpaca_advance(d, u, s1, s2, stp, int(nsteps), st), "advance")

    def record(self) -> CaseState:
        raw = self.state.cpu().numpy().tobytes()
        return CaseState.from_buffer_copy(raw)

    def run(self, max_steps: int, sync_every: int = 1, on_step=None):
        """ComputeLoop: advance while time < t_end; on_step(record) runs after
        each macro step, where ComputeLoop writes restart snapshots and
        output.  The record is read back every sync_every steps (1 with
        on_step): the device skips the steps of a chunk that follow the end
        time, and dts then holds the time step of each chunk's last step."""
        chunk = 1 if on_step is not None else max(1, int(sync_every))
        steps = 0
        dts = []
        rec = self.record()
        while steps < max_steps and rec.time < self.t_end:
            before = rec.steps
            self.advance(min(chunk, max_steps - steps))
            rec = self.record()
            steps += rec.steps - before
            dts.append(rec.dt)
            if rec.flags & 1:
                raise AlpacaError("non-finite state")
            if on_step is not None:
                on_step(rec)
            if rec.flags & 2:
                break
        return steps, dts

    def download(self) -> np.ndarray:
        return from_blocks(self.u, self.dim, self.nb).cpu().numpy()

    def download_blocks(self) -> np.ndarray:
        """The haloed blocks as the device holds them, [n][NE][24][24][TZ]
        (TZ = 1 in 2D)."""
        tz = TC if self.dim == 3 else 1
        return self.u.cpu().numpy().reshape(self.nblocks, NE(self.dim), TC, TC, tz)


def profile_enable(on: bool = True):
    check(load_library().alpaca_profile_enable(1 if on else 0), "profile_enable")


def profile_read(kernel: int, reset: bool = True):
    """(total_ms, launches, units) of a kernel class since the last reset."""
    ms = ctypes.c_double()
    n = ctypes.c_longlong()
    u = ctypes.c_double()
    check(load_library().alpaca_profile_read(kernel, 1 if reset else 0, ctypes.byref(ms),
                                             ctypes.byref(n), ctypes.byref(u)), "profile_read")
    return ms.value, n.value, u.value
