"""Python mirror of the alpaca_ensemble_* C ABI: many independent uniform
cases resident in one HBM block pool, advanced together by one persistent
kernel per RK stage from a captured hipGraph (see include/alpaca_gpu.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import gpu

CASE_NONFINITE = 1
CASE_UNHEALTHY = 2
CASE_DONE = 4
CASE_CANCELLED = 8
CASE_ACTIVE = 32
CASE_STOP = CASE_NONFINITE | CASE_UNHEALTHY | CASE_DONE | CASE_CANCELLED


def class_bit(flux: int, precision: int) -> int:
    m = {gpu.FP64: 0, gpu.FP64_FMA: 1, gpu.MIXED: 2}[precision]
    return 1 << (3 * flux + m)


class Ensemble:
    """One ensemble; with ``share_with`` (another Ensemble) it admits its
    cases into that ensemble's block pool (alpaca_ensemble_create_shared), so
    a 2D and a 3D ensemble draw on one HBM budget; ``max_blocks`` is then
    ignored.  The pool lives until the last ensemble using it is closed."""

    def __init__(self, dim: int, max_blocks: int, max_cases: int, classes, stream=None,
                 share_with: "Ensemble | None" = None):
        import torch

        self.torch = torch
        self.lib = gpu.load_library()
        self.dim = dim
        self.max_cases = max_cases
        mask = 0
        for flux, prec in classes:
            mask |= class_bit(flux, prec)
        # A stream of its own (default priority): graph capture needs a
        # non-legacy stream.
        self.stream = stream if stream is not None else torch.cuda.Stream()
        h = ctypes.c_void_p()
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        if share_with is not None:
            gpu.check(self.lib.alpaca_ensemble_create_shared(share_with.h, dim, max_cases, mask,
                                                             sp, ctypes.byref(h)),
                      "ensemble_create_shared")
        else:
            gpu.check(self.lib.alpaca_ensemble_create(dim, max_blocks, max_cases, mask, sp,
                                                      ctypes.byref(h)), "ensemble_create")
        self.h = h
        self.desc = {}
        self._staged = []  # admission sources in flight on the stream

    @staticmethod
    def bytes_needed(dim, max_blocks, max_cases):
        return gpu.load_library().alpaca_ensemble_bytes(dim, max_blocks, max_cases)

    def close(self):
        if self.h:
            gpu.check(self.lib.alpaca_ensemble_destroy(self.h), "ensemble_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_case(self, desc: gpu.CaseDesc, state0, t_end: float, cfl: float,
                 sync: bool = True) -> int:
        """Admit a case from its domain conservatives [NE][NX][NY][NZ].  The
        copy into the pool is ordered on the ensemble's stream; with
        sync=False the call returns at once and the staged source is kept
        alive until the next states() (which synchronises the stream)."""
        torch = self.torch
        nb = [desc.nb[0], desc.nb[1], desc.nb[2] if desc.dim == 3 else 1]
        with torch.cuda.stream(self.stream):
            if torch.is_tensor(state0):
                g = state0.to(device="cuda", dtype=torch.float64)
            else:
                g = torch.as_tensor(np.ascontiguousarray(state0), device="cuda")
            blocks = gpu.to_blocks(g, desc.dim, nb).contiguous()
        if torch.is_tensor(state0):
            self.stream.wait_stream(torch.cuda.current_stream())
        cid = ctypes.c_int()
        st = self.lib.alpaca_ensemble_add_case(self.h, ctypes.byref(desc),
                                               ctypes.c_void_p(blocks.data_ptr()), t_end, cfl,
                                               ctypes.byref(cid))
        if st == 5:
            raise MemoryError("ensemble capacity exhausted")
        gpu.check(st, "ensemble_add_case")
        if sync:
            self.stream.synchronize()
        else:
            self._staged.append(blocks)
        self.desc[cid.value] = (desc, nb)
        return cid.value

    def remove_case(self, case_id: int):
        gpu.check(self.lib.alpaca_ensemble_remove_case(self.h, case_id), "ensemble_remove_case")
        self.desc.pop(case_id, None)

    def cancel_case(self, case_id: int):
        gpu.check(self.lib.alpaca_ensemble_cancel_case(self.h, case_id), "ensemble_cancel_case")

    def set_tuning(self, tile_config: int, groups_per_cu: int = 0):
        gpu.check(self.lib.alpaca_ensemble_set_tuning(self.h, tile_config, groups_per_cu),
                  "ensemble_set_tuning")

    def step(self, nsteps: int = 1):
        gpu.check(self.lib.alpaca_ensemble_step(self.h, nsteps), "ensemble_step")

    def states(self, n=None):
        n = self.max_cases if n is None else n
        arr = (gpu.CaseState * n)()
        gpu.check(self.lib.alpaca_ensemble_read_states(self.h, arr, n), "ensemble_read_states")
        self._staged.clear()  # the stream has drained: staged sources are consumed
        return list(arr)

    def case_blocks(self, case_id: int):
        """Copy (torch, device) of a case's current haloed blocks."""
        desc, nb = self.desc[case_id]
        n = nb[0] * nb[1] * nb[2] * gpu.block_doubles(self.dim)
        out = self.torch.empty(n, dtype=self.torch.float64, device="cuda")
        gpu.check(self.lib.alpaca_ensemble_copy_case(self.h, case_id,
                                                     ctypes.c_void_p(out.data_ptr())),
                  "ensemble_copy_case")
        self.stream.synchronize()
        return out

    def download(self, case_id: int) -> np.ndarray:
        desc, nb = self.desc[case_id]
        blocks = self.case_blocks(case_id)
        return gpu.from_blocks(blocks, desc.dim, nb).cpu().numpy()

    def capacity(self):
        fb, lr, fs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        gpu.check(self.lib.alpaca_ensemble_capacity(self.h, ctypes.byref(fb), ctypes.byref(lr),
                                                    ctypes.byref(fs)), "ensemble_capacity")
        return fb.value, lr.value, fs.value

    def run_until_done(self, case_ids, max_steps=100000, poll=8):
        """Step until every listed case has stopped; returns total steps."""
        steps = 0
        while steps < max_steps:
            st = self.states()
            if all(st[c].flags & CASE_STOP for c in case_ids):
                break
            self.step(poll)
            steps += poll
        return steps
