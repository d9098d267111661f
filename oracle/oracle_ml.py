"""ORACLE — test infrastructure only.  CPU backend of the multi-level driver
(alpaca_amd.multilevel.MultiLevelSolver): every cell operation the driver
asks for, done block by block by the C restatement (liboracle.so) on numpy
pools, and the halo update restated the way the reference walks it --
per node and per halo location: same-level copies
(InternalHaloManager::UpdateNoJumpLocal, internal_halo_manager.cpp:439-466,
ranges of communication_types.h:45-132), then the jump halos predicted from
the parent (UpdateMaterialJumpNoMpi, internal_halo_manager.cpp:231-262), then
the external conditions East, West, North, South, Top, Bottom over the full
transverse extent (halo_manager.cpp:112-123, boundary_constants.h:169-253).

Only tests/ use this; the product runs alpaca_amd.multilevel.GpuBackend.
"""
from __future__ import annotations

import numpy as np

import oracle_py

IC, HS, TC = 16, 4, 24
BC_ZERO_GRADIENT, BC_PERIODIC, BC_SYMMETRY = 0, 1, 2

# BoundaryLocation names in enum order (boundary_specifications.h:56-86).
_LOCATIONS = ("East", "West", "North", "South", "Top", "Bottom",
              "BottomNorth", "BottomSouth", "TopNorth", "TopSouth",
              "BottomEast", "BottomWest", "TopEast", "TopWest",
              "NorthEast", "NorthWest", "SouthEast", "SouthWest",
              "EastNorthTop", "EastNorthBottom", "EastSouthTop", "EastSouthBottom",
              "WestNorthTop", "WestNorthBottom", "WestSouthTop", "WestSouthBottom")


def _direction(loc):
    name = _LOCATIONS[loc]
    return tuple(1 if hi in name else (-1 if lo in name else 0)
                 for hi, lo in (("East", "West"), ("North", "South"), ("Top", "Bottom")))


def _rows(links):
    """Link rows (child, parent, position, level) as Python ints, from a list
    of tuples or an int array."""
    return [tuple(int(v) for v in r) for r in links]


def _slot(dim, d):
    return ((d[0] + 1) * 3 + (d[1] + 1)) * 3 + (d[2] + 1) if dim == 3 else (d[0] + 1) * 3 + (d[1] + 1)


def _recv(dim, d):
    """GetStartIndicesHaloRecv / GetHaloSize of a direction, per axis."""
    out = []
    for a in range(dim):
        out.append(slice(IC + HS, TC) if d[a] > 0 else (slice(0, HS) if d[a] < 0 else slice(HS, HS + IC)))
    if dim == 2:
        out.append(slice(0, 1))
    return tuple(out)


def _send(dim, d):
    """The neighbour's cells that fill a recv range (its internal layers
    next to the shared face: GetStartIndicesHaloSend of the opposite side)."""
    out = []
    for a in range(dim):
        out.append(slice(HS, 2 * HS) if d[a] > 0 else (slice(IC, IC + HS) if d[a] < 0 else slice(HS, HS + IC)))
    if dim == 2:
        out.append(slice(0, 1))
    return tuple(out)


class OracleBackend:
    def __init__(self, dim, capacity, flux=0, precision=0, gamma=1.4, background_pressure=0.0):
        if precision != 0:
            raise ValueError("the oracle restates the FP64 path only")
        self.dim = dim
        self.capacity = int(capacity)
        self.flux, self.gamma, self.b = flux, gamma, background_pressure
        tz = TC if dim == 3 else 1
        icz = IC if dim == 3 else 1
        ne = dim + 2
        self.avg = np.zeros((capacity, ne, TC, TC, tz))
        self.rhs = np.zeros_like(self.avg)
        self.init = np.zeros_like(self.avg)
        self.jflux = np.zeros((capacity, 6, ne, IC, icz))
        self.jcons = np.zeros_like(self.jflux)

    def close(self):
        pass

    # -- node data
    def clear(self, slots):
        for slot in slots:
            for t in (self.avg, self.rhs, self.init, self.jflux, self.jcons):
                t[slot] = 0.0

    def _internal(self):
        if self.dim == 3:
            return (slice(None), slice(HS, HS + IC), slice(HS, HS + IC), slice(HS, HS + IC))
        return (slice(None), slice(HS, HS + IC), slice(HS, HS + IC), slice(None))

    def set_internal(self, slots, cons):
        for s, c in zip(slots, cons):
            self.rhs[s][self._internal()] = c

    def get_blocks(self, slots):
        return np.stack([self.avg[s] for s in slots])

    def set_rhs_blocks(self, slots, blocks):
        for s, b in zip(slots, blocks):
            self.rhs[s] = b

    def get_internal(self, slots):
        return np.stack([self.avg[s][self._internal()] for s in slots])

    def swap(self, slots):
        for s in slots:
            a = self.avg[s].copy()
            self.avg[s] = self.rhs[s]
            self.rhs[s] = a

    def fill_initial(self, slots):
        for s in slots:
            self.init[s] = self.avg[s]

    # -- stage operations
    def right_hand_side(self, slots, cell_size):
        for s in slots:
            blk = np.ascontiguousarray(self.avg[s])
            r = oracle_py.block_rhs(self.dim, self.flux, self.gamma, self.b, blk, cell_size)
            self.rhs[s] = 0.0
            self.rhs[s][self._internal()] = r
            self.jflux[s] = oracle_py.block_jump_fluxes(self.dim, self.flux, self.gamma, self.b,
                                                        blk, self.jflux[s])

    def prepare(self, slots, stage):
        if stage == 0:
            return
        for s in slots:
            self.avg[s] = oracle_py.lts_prepare_buffer(self.dim, self.avg[s], self.init[s], stage)

    def lts_timestep(self, micro, n, stage, jump):
        return oracle_py.lts_timestep(micro, n, stage, jump)

    def integrate_node(self, slots, jump_timestep, timestep):
        for s in slots:
            self.jcons[s], self.jflux[s] = oracle_py.lts_integrate_jump(
                self.dim, self.jcons[s], self.jflux[s], jump_timestep)
            self.rhs[s] = oracle_py.lts_integrate(self.dim, self.avg[s], self.rhs[s], None, timestep)

    def integrate_range(self, slots, rng, timestep):
        r = list(rng) + [0, 1] * (3 - len(rng) // 2)
        for s in slots:
            self.rhs[s] = oracle_py.lts_integrate(self.dim, self.avg[s], self.rhs[s], r, timestep)

    def integrate_jump_halos(self, by_loc, timestep):
        """IntegrateJumpHalos (time_integrator.h:281-301) per location of the
        level's jumps (modular_algorithm_assembler.cpp:1084-1096), the halo
        slab from GetStartIndicesHaloRecv / GetHaloSize."""
        for loc, slots in by_loc.items():
            rng = []
            for sl in _recv(self.dim, _direction(loc))[:self.dim]:
                rng += [sl.start, sl.stop - sl.start]
            self.integrate_range(slots, rng, timestep)

    def average(self, links):
        for child, parent, pos, _ in _rows(links):
            self.rhs[parent] = oracle_py.mr_average(self.dim, self.rhs[child], self.rhs[parent], pos)

    def average_jump(self, links):
        for child, parent, pos, _ in _rows(links):
            self.jcons[parent] = oracle_py.mr_average_jump(self.dim, self.jcons[child],
                                                           self.jcons[parent], pos)

    def _predict_range(self, child, parent, pos, rng):
        for e in range(self.dim + 2):
            self.rhs[child][e] = oracle_py.mr_prediction(self.dim, self.rhs[parent][e],
                                                         self.rhs[child][e], pos, rng)

    def halo_update(self, slots, table, jump_links, jump_locs, side_bc, second=None):
        # `second` (the blocks a device pass after the jump update can write)
        # is a device-side launch restriction; this restatement visits every
        # block.
        dim = self.dim
        locs = [loc for loc in range(26) if dim == 3 or _direction(loc)[2] == 0]
        # same-level copies
        for r, s in enumerate(slots):
            for loc in locs:
                d = _direction(loc)
                nb = table[r, _slot(dim, d)]
                if nb >= 0:
                    self.rhs[s][(slice(None),) + _recv(dim, d)] = \
                        self.rhs[nb][(slice(None),) + _send(dim, d)]
        # level jumps: prediction from the parent over the halo range
        for (child, parent, pos, _), loc in zip(_rows(jump_links), jump_locs):
            d = _direction(loc)
            rng = []
            for a in range(3):
                if a >= dim:
                    rng += [0, 1]
                else:
                    rng += [IC + HS, HS] if d[a] > 0 else ([0, HS] if d[a] < 0 else [HS, IC])
            self._predict_range(child, parent, pos, rng)
        # external conditions, East, West, North, South, Top, Bottom
        for r, s in enumerate(slots):
            v = self.rhs[s]
            for a in range(dim):
                for side in (1, 0):
                    d = [0, 0, 0]
                    d[a] = 1 if side else -1
                    if table[r, _slot(dim, d)] != -1:
                        continue
                    kind = side_bc[2 * a + side]
                    if kind == BC_PERIODIC:
                        continue
                    h0 = HS + IC if side else 0
                    for h in range(h0, h0 + HS):
                        if kind == BC_SYMMETRY:
                            src = (2 * (HS + IC) - 1 if side else 2 * HS - 1) - h
                        else:
                            src = HS + IC - 1 if side else HS
                        dst_ix = [slice(None)] * 4
                        src_ix = [slice(None)] * 4
                        dst_ix[1 + a] = h
                        src_ix[1 + a] = src
                        val = v[tuple(src_ix)].copy()
                        if kind == BC_SYMMETRY:
                            val[2 + a] = -1.0 * val[2 + a]
                        v[tuple(dst_ix)] = val

    def jump_flux_adjustment(self, entries, one_cell_size):
        for row in entries:
            leaf = row[0]
            nbs = [None if row[1 + k] < 0 else self.jcons[row[1 + k]] for k in range(6)]
            self.rhs[leaf], self.jcons[leaf] = oracle_py.lts_jump_flux_adjustment(
                self.dim, self.rhs[leaf], self.jcons[leaf], nbs, one_cell_size)

    def reset_jump_conservatives(self, slots):
        for s in slots:
            self.jcons[s] = 0.0

    def reset_jump_all(self, slots):
        for s in slots:
            self.jcons[s] = 0.0
            self.jflux[s] = 0.0

    def predict(self, links, rng):
        r = list(rng) + [0, 1] * (3 - len(rng) // 2)
        for child, parent, pos, _ in _rows(links):
            self._predict_range(child, parent, pos, r)

    def decisions(self, links, thresholds):
        out = []
        for child, parent, pos, level in _rows(links):
            det = oracle_py.mr_detail_linf(self.dim, self.rhs[parent], self.rhs[child], pos)
            out.append(oracle_py.mr_decision(det, thresholds[level]))
        return np.array(out, dtype=np.int32)

    def timestep(self, leaf_slots, cfl, dx_finest, run_time, t_end):
        f = oracle_py.oracle().ora_block_signal
        f.restype = oracle_py.ctypes.c_double
        s = 0.0
        for slot in leaf_slots:
            blk = np.ascontiguousarray(self.avg[slot])
            v = f(self.dim, oracle_py.ctypes.c_double(self.gamma),
                  oracle_py.ctypes.c_double(self.b), oracle_py._dp(blk))
            s = v if s < v else s
        rate = s / dx_finest
        dt = np.finfo(np.float64).max if rate == 0.0 else cfl / rate
        if run_time + dt > t_end:
            dt = max(t_end - run_time, 0.0)
        return dt
