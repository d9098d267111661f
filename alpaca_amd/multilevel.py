"""Multi-level driver: multiresolution with local time stepping.

The host side of ALPACA's default mode for the single-phase path: a tree of
blocks over levels 0..maximum_level, every level's leaves advanced with their
own number of micro steps, parents filled by averaging, halos across level
jumps predicted from the parent, jump fluxes corrected for conservation and
the mesh adapted by the wavelet analysis after each finished level.  It
follows the reference's ModularAlgorithmAssembler -- CreateNewSimulation
(modular_algorithm_assembler.cpp:286-538), Advance (:543-910),
ComputeRightHandSide (:954-1008), Integrate (:1068-1098), SwapBuffers
(:1044-1059), JumpFluxAdjustment (:1347-1647), ComputeTimestepSize
(:1777-1997), Remesh/DetermineRemeshingNodes/RefineNode (:2301-2553),
GetLevels (:2563-2572) -- with the Averager (averager.cpp:62-140), the halo
manager (halo_manager.cpp:59-123, internal_halo_manager.cpp:53-97,
communication_manager.cpp:81-251) and the topology rules it relies on
(topology_manager.cpp:150-260, 363-390).

The driver owns the topology (a dict of nodes keyed by (level, i, j, k)) and
the schedule; every operation on cell data is one batched call of a backend
over the nodes of a level.  ``GpuBackend`` runs them as gfx950 kernels of
libalpaca_hip.so on pools of haloed blocks in HBM.  Buffers per node follow
the reference's Block: Average, RightHandSide and Initial conservatives plus
the jump-flux and jump-conservative surfaces; a node's buffers live at the
same slot of five device pools.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import gpu, lts
from . import multiresolution as mr

PERIODIC = gpu.BC_PERIODIC
RK_STAGES = 3


def halo_locations(dim: int) -> list:
    """CC::HBS(): the halo locations of a block (faces, then the sticks and
    corners the dimension has), as BoundaryLocation indices."""
    return [loc for loc in range(len(mr.LOCATIONS))
            if dim == 3 or mr.location_direction(loc)[2] == 0]


def face_locations(dim: int) -> list:
    """CC::ANBS(): East, West, North, South (, Top, Bottom)."""
    return list(range(2 * dim))


def offset_slot(dim: int, off) -> int:
    """Slot of an offset in a neighbour-table row (alpaca_fill_halos_list)."""
    if dim == 3:
        return ((off[0] + 1) * 3 + (off[1] + 1)) * 3 + (off[2] + 1)
    return (off[0] + 1) * 3 + (off[1] + 1)


class Topology:
    """The tree of nodes: existence, leaves, parents, children and neighbours
    over levels 0..maximum_level (the reference's TopologyManager and the id
    arithmetic of id_information.cpp, with (level, i, j, k) keys instead of
    bit-encoded ids).  Periodic axes wrap at every level."""

    def __init__(self, dim: int, nb0, maximum_level: int, periodic):
        self.dim = dim
        self.nb0 = [int(x) for x in nb0] + [1] * (3 - len(nb0))
        if dim == 2:
            self.nb0[2] = 1
        self.maximum_level = maximum_level
        self.periodic = list(periodic) + [False] * (3 - len(periodic))
        self.leaf = {}  # key -> bool
        # Bumped by every change of the tree (nodes or leaf flags): the
        # solver's per-level tables are rebuilt only when it moved.
        self.version = 0

    def set_leaf(self, key, flag: bool):
        self.leaf[key] = flag
        self.version += 1

    def remove(self, key):
        del self.leaf[key]
        self.version += 1

    def extent(self, level: int, axis: int) -> int:
        if axis >= self.dim:
            return 1
        return self.nb0[axis] << level

    def neighbor(self, key, off):
        """Key of the same-level position at offset `off`, or None outside
        the domain (wrapped across periodic axes)."""
        level = key[0]
        q = [key[1], key[2], key[3]]
        for a in range(self.dim):
            v = q[a] + off[a]
            n = self.extent(level, a)
            if v < 0 or v >= n:
                if not self.periodic[a]:
                    return None
                v %= n
            q[a] = v
        return (level, q[0], q[1], q[2])

    def exists(self, key) -> bool:
        return key in self.leaf

    def is_leaf(self, key) -> bool:
        return self.leaf[key]

    @staticmethod
    def parent(key):
        return (key[0] - 1, key[1] >> 1, key[2] >> 1, key[3] >> 1)

    @staticmethod
    def position(key) -> int:
        """PositionOfNodeAmongSiblings (id & 7): bit 0 x, bit 1 y, bit 2 z."""
        return (key[1] & 1) | ((key[2] & 1) << 1) | ((key[3] & 1) << 2)

    def children(self, key):
        lv, i, j, k = key
        zs = (0, 1) if self.dim == 3 else (0,)
        out = []
        for p in range(1 << self.dim):
            dx, dy, dz = p & 1, (p >> 1) & 1, (p >> 2) & 1
            if self.dim == 2 and dz:
                continue
            out.append((lv + 1, 2 * i + dx, 2 * j + dy, 2 * k + dz if zs != (0,) else 0))
        return out

    def ids_on_level(self, level: int):
        return sorted(k for k in self.leaf if k[0] == level)

    def leaves_on_level(self, level: int):
        return sorted(k for k, lf in self.leaf.items() if lf and k[0] == level)

    def leaves(self):
        return sorted((k for k, lf in self.leaf.items() if lf), key=lambda k: k)

    def current_maximum_level(self) -> int:
        return max(k[0] for k in self.leaf)


class MultiLevelSolver:
    """One single-phase multi-level simulation.

    side_bc: per side (-x, +x, -y, +y[, -z, +z]) ALPACA_BC_*; periodic axes
    must be periodic on both sides.  initial(level, key) returns the
    conservatives [NE][16][16][16|1] of a node's internal cells (prime states
    at that level's cell centres converted as PrimeStatesToConservatives).
    """

    def __init__(self, backend, dim: int, nb0, node_size: float, maximum_level: int, side_bc,
                 initial, cfl: float, t_end: float, epsilon_reference: float,
                 level_of_epsilon_reference: int, flux: int = gpu.FLUX_HLLC,
                 gamma: float = 1.4, background_pressure: float = 0.0):
        for name, want in (("flux", flux), ("gamma", gamma), ("b", background_pressure),
                           ("dim", dim)):
            have = getattr(backend, name, want)
            if have != want:
                raise ValueError(f"the backend's {name} is {have}, the solver's {want}")
        self.be = backend
        self.dim = dim
        self.node_size = float(node_size)
        self.maximum_level = int(maximum_level)
        self.side_bc = list(side_bc)
        periodic = [side_bc[2 * a] == PERIODIC for a in range(dim)]
        for a in range(dim):
            if (side_bc[2 * a] == PERIODIC) != (side_bc[2 * a + 1] == PERIODIC):
                raise ValueError("a periodic axis must be periodic on both sides")
        self.topo = Topology(dim, nb0, maximum_level, periodic)
        self.initial = initial
        self.cfl = float(cfl)
        self.t_end = float(t_end)
        self.flux = flux
        self.gamma = gamma
        self.b = background_pressure
        self.thresholds = mr.ThresholdComputer(dim, maximum_level, level_of_epsilon_reference,
                                               epsilon_reference).table(mr.MAX_LEVELS)
        self.slot = {}
        self.free = list(range(backend.capacity - 1, -1, -1))
        self.micro: list = []  # micro time steps of the running macro step
        self.current_run_time = 0.0
        self.macro_steps = 0
        self.all_levels = list(range(self.maximum_level + 1))
        self.halo_locs = halo_locations(dim)
        self.face_locs = face_locations(dim)
        self._loc_dir = {loc: mr.location_direction(loc) for loc in range(len(mr.LOCATIONS))}
        self._cache = {}

    # ------------------------------------------------------------- helpers
    def cell_size(self, level: int) -> float:
        return self.node_size / float(1 << level) / gpu.IC

    def _new_nodes(self, keys, leaf=True):
        """Nodes with zeroed buffers (one batched clear of their slots)."""
        if len(self.free) < len(keys):
            raise MemoryError("multi-level block pool exhausted")
        slots = []
        for key in keys:
            s = self.free.pop()
            self.slot[key] = s
            self.topo.set_leaf(key, leaf)
            slots.append(s)
        if slots:
            self.be.clear(slots)
        return slots

    def _remove_node(self, key):
        s = self.slot.pop(key)
        self.topo.remove(key)
        self.free.append(s)

    def _slots(self, keys):
        return [self.slot[k] for k in keys]

    def _link(self, key):
        return (self.slot[key], self.slot[self.topo.parent(key)], self.topo.position(key), key[0])

    def _level_links(self, level):
        """The links (child slot, parent slot, position among siblings,
        level) of every node of a level above 0, in the order of _ids, as an
        int32 [n][4] array (from the dense level grids)."""
        def build():
            keys = self._ids(level)
            if level == 0 or not keys:
                return np.zeros((0, 4), dtype=np.int32)
            k = np.asarray(keys, dtype=np.int64)
            slots, _ = self._grid(level)
            pslots, _ = self._grid(level - 1)
            child = slots[k[:, 1], k[:, 2], k[:, 3]]
            parent = pslots[k[:, 1] >> 1, k[:, 2] >> 1, k[:, 3] >> 1]
            pos = (k[:, 1] & 1) | ((k[:, 2] & 1) << 1) | ((k[:, 3] & 1) << 2)
            return np.ascontiguousarray(
                np.stack([child, parent, pos, k[:, 0]], axis=1).astype(np.int32))
        return self._cached("level_links", level, build)

    def _lts_timestep(self, stage, n, jump=False):
        return self.be.lts_timestep(self.micro, n, stage, jump)

    def _cached(self, kind, level, build):
        """A per-level table (node lists, neighbour tables, link lists) built
        once per version of the tree."""
        ent = self._cache.get((kind, level))
        if ent is not None and ent[0] == self.topo.version:
            return ent[1]
        val = build()
        self._cache[(kind, level)] = (self.topo.version, val)
        return val

    def _ids(self, level):
        return self._cached("ids", level, lambda: self.topo.ids_on_level(level))

    def _leaves(self, level):
        return self._cached("leaves", level, lambda: self.topo.leaves_on_level(level))

    # -------------------------------------------------------- halo update
    def _grid(self, level):
        """Dense view of a level: slot of the node at each (i, j, k) (-1 where
        there is none) and its leaf flag, as arrays over the level's extent."""
        def build():
            shape = tuple(self.topo.extent(level, a) for a in range(3))
            slots = np.full(shape, -1, dtype=np.int32)
            leaf = np.zeros(shape, dtype=bool)
            for key in self._ids(level):
                slots[key[1], key[2], key[3]] = self.slot[key]
                leaf[key[1], key[2], key[3]] = self.topo.leaf[key]
            return slots, leaf
        return self._cached("grid", level, build)

    def _neighbors(self, level, keys, offs):
        """For the nodes `keys` of a level and same-level offsets `offs`
        (m x 3): (slot of the neighbour or -1 where no node exists, inside
        the domain, neighbour is a leaf), each an n x m array (the
        reference's neighbour search, topology_manager.cpp:363-390, with
        periodic wrap)."""
        slots, leaf = self._grid(level)
        k = np.array([key[1:] for key in keys], dtype=np.int64).reshape(-1, 1, 3)
        o = np.asarray(offs, dtype=np.int64).reshape(1, -1, 3)
        q = k + o
        ext = np.array(slots.shape, dtype=np.int64)
        inside = np.ones(q.shape[:2], dtype=bool)
        for a in range(3):
            out = (q[..., a] < 0) | (q[..., a] >= ext[a])
            if a < self.dim and self.topo.periodic[a]:
                q[..., a] %= ext[a]
            else:
                inside &= ~out
                q[..., a] = np.clip(q[..., a], 0, ext[a] - 1)
        nb = np.where(inside, slots[q[..., 0], q[..., 1], q[..., 2]], -1)
        nleaf = inside & leaf[q[..., 0], q[..., 1], q[..., 2]]
        return nb, inside, nleaf

    def _neighbor_table(self, keys):
        """Rows for alpaca_fill_halos_list: neighbour slot, -1 outside the
        domain, -2 for a level jump (no node of this level there)."""
        dim = self.dim
        offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1)
                for dz in ((-1, 0, 1) if dim == 3 else (0,))]
        nb, inside, _ = self._neighbors(keys[0][0], keys, offs)
        table = np.where(inside, np.where(nb >= 0, nb, -2), -1).astype(np.int32)
        # columns follow offset_slot's order, which is the order of offs
        return np.ascontiguousarray(table)

    def _level_table(self, level):
        """The neighbour table of a level's nodes (in _ids order), kept until
        the tree changes."""
        return self._cached("table", level, lambda: self._neighbor_table(self._ids(level)))

    def halo_update_on_level(self, level: int, cut_jumps: bool):
        """MaterialHaloUpdateOnLevel for the conservatives (the right-hand-side
        buffers): same-level copies, then (unless cut_jumps) the jump halos
        predicted from the parents, then the external conditions."""
        keys = self._ids(level)
        if not keys:
            return

        def build():
            table = self._level_table(level)
            cols = [offset_slot(self.dim, self._loc_dir[loc]) for loc in self.halo_locs]
            rows, which = np.nonzero(table[:, cols] == -2)
            jump_links = self._level_links(level)[rows] if rows.size else []
            jump_locs = [self.halo_locs[c] for c in which.tolist()]
            slots = self._slots(keys)
            # External-slab cells whose source lies in a jump region (written
            # after the jump update) exist only in blocks with both a jump
            # neighbour (-2) and an offset leaving the domain (-1).
            second = None
            jrows = np.nonzero((table == -2).any(axis=1) & (table == -1).any(axis=1))[0]
            if jrows.size:
                second = ([slots[r] for r in jrows.tolist()], np.ascontiguousarray(table[jrows]))
            return slots, table, jump_links, jump_locs, second

        slots, table, jump_links, jump_locs, second = self._cached("halo", level, build)
        if cut_jumps:
            jump_links, jump_locs = [], []
        self.be.halo_update(slots, table, jump_links, jump_locs, self.side_bc, second)

    def halo_update(self, levels_ascending, cut_jumps=False):
        """HaloManager::MaterialHaloUpdate (halo_manager.cpp:59-76)."""
        levels = list(levels_ascending)
        if cut_jumps and levels:
            self.halo_update_on_level(levels.pop(0), True)
        for level in levels:
            self.halo_update_on_level(level, False)

    # ----------------------------------------------------------- averaging
    def average_material(self, child_levels_descending):
        """Averager::AverageMaterial: every node of each child level (but 0)
        averaged into its parent's right-hand-side buffer, finest first."""
        for level in child_levels_descending:
            if level == 0:
                continue
            links = self._level_links(level)
            if len(links):
                self.be.average(links)

    # ---------------------------------------------------------- stage work
    def compute_right_hand_side(self, levels, stage):
        for level in levels:
            slots = self._cached("leaf_slots", level, lambda: self._slots(self._leaves(level)))
            if not slots:
                continue
            if stage == 0:
                self.be.fill_initial(slots)
            self.be.right_hand_side(slots, self.cell_size(level))
            self.be.prepare(slots, stage)

    def integrate(self, levels, stage):
        for level in levels:
            n = 1 << (self.maximum_level - level)
            slots = self._cached("leaf_slots", level, lambda: self._slots(self._leaves(level)))
            if slots:
                self.be.integrate_node(slots, self._lts_timestep(stage, n, True),
                                       self._lts_timestep(stage, n))

            def build():
                keys = self._ids(level)
                by_loc = {}
                if not keys:
                    return by_loc
                # a jump at a location: no node of this level there (-2)
                cols = [offset_slot(self.dim, self._loc_dir[loc]) for loc in self.halo_locs]
                missing = self._level_table(level)[:, cols] == -2
                slots = self._slots(keys)
                for c, loc in enumerate(self.halo_locs):
                    rows = np.nonzero(missing[:, c])[0]
                    if rows.size:
                        by_loc[loc] = [slots[r] for r in rows.tolist()]
                return by_loc

            by_loc = self._cached("jump_halos", level, build)
            if by_loc:
                self.be.integrate_jump_halos(by_loc, self._lts_timestep(stage, n))

    def swap(self, levels):
        slots = [s for lv in levels for s in self._cached("slots", lv,
                                                            lambda: self._slots(self._ids(lv)))]
        if slots:
            self.be.swap(slots)

    def jump_flux_adjustment(self, finished_levels_descending):
        for level in finished_levels_descending:
            if level == 0:
                continue
            links = self._level_links(level)
            if len(links):
                self.be.average_jump(links)
        for level in finished_levels_descending[1:]:
            leaves = self._leaves(level)
            if not leaves:
                continue

            def build():
                entries = np.full((len(leaves), 7), -1, dtype=np.int32)
                entries[:, 0] = self._slots(leaves)
                nb, inside, nleaf = self._neighbors(level, leaves,
                                                    [self._loc_dir[loc] for loc in self.face_locs])
                parent_of_finer = inside & (nb >= 0) & ~nleaf
                for c, loc in enumerate(self.face_locs):
                    entries[:, 1 + loc] = np.where(parent_of_finer[:, c], nb[:, c], -1)
                return entries

            entries = self._cached("jfa", level, build)
            self.be.jump_flux_adjustment(entries, 1.0 / self.cell_size(level))
        slots = [s for lv in finished_levels_descending
                 for s in self._cached("slots", lv, lambda: self._slots(self._ids(lv)))]
        if slots:
            self.be.reset_jump_conservatives(slots)

    # ------------------------------------------------------------ remeshing
    def determine_remeshing_nodes(self, parent_levels):
        """The children of every node of the parent levels, per parent (in
        key order) and position: leaf children get a remeshing decision (one
        batched call), the others count as neutral; children to refine, and
        the whole family of a parent whose children are all to be coarsened
        (DetermineRemeshingNodes, modular_algorithm_assembler.cpp:2426-2532)."""
        remove, refine = [], []
        fams = []  # (child keys, leaf flags) per parent level, in family order
        links = []
        for lp in parent_levels:
            lk = self._level_links(lp + 1)
            if not len(lk):
                continue
            k = np.asarray(self._ids(lp + 1), dtype=np.int64)
            order = np.lexsort((lk[:, 2], k[:, 3] >> 1, k[:, 2] >> 1, k[:, 1] >> 1))
            k = k[order]
            _, leafgrid = self._grid(lp + 1)
            leaf = leafgrid[k[:, 1], k[:, 2], k[:, 3]]
            fams.append((k, leaf))
            links.append(lk[order][leaf])
        links = np.concatenate(links) if links else np.zeros((0, 4), dtype=np.int32)
        decisions = np.asarray(self.be.decisions(links, self.thresholds)) if len(links) else []
        off = 0
        for k, leaf in fams:
            dec = np.full(len(k), mr.NEUTRAL, dtype=np.int64)
            n = int(leaf.sum())
            dec[leaf] = decisions[off:off + n]
            off += n
            refine += [tuple(x) for x in k[dec == mr.REFINE].tolist()]
            parent = k[:, 1:] >> 1
            start = np.ones(len(k), dtype=bool)
            start[1:] = np.any(parent[1:] != parent[:-1], axis=1)
            starts = np.nonzero(start)[0]
            all_coarse = np.logical_and.reduceat(dec == mr.COARSE, starts)
            fam = np.cumsum(start) - 1
            remove += [tuple(x) for x in k[all_coarse[fam]].tolist()]
        return remove, refine

    def _coarsen(self, removed):
        parents = sorted({self.topo.parent(k) for k in removed if self.topo.parent(k)[0] > 0})
        for pkey in parents:
            for c in self.topo.children(pkey):
                if self.topo.exists(c):
                    self._remove_node(c)
            self.topo.set_leaf(pkey, True)

    def refine_node(self, key):
        """RefineNode: children created, their internal cells predicted from
        the node's right-hand-side buffer."""
        self.topo.set_leaf(key, False)
        kids = self.topo.children(key)
        self._new_nodes(kids)
        return [self._link(c) for c in kids]

    def remesh(self, levels_ascending):
        cur_max = self.topo.current_maximum_level()
        parent_levels = [lv for lv in levels_ascending if lv < cur_max]
        remove, refine = self.determine_remeshing_nodes(parent_levels)
        refine = [k for k in refine if self.topo.is_leaf(k) and k[0] != self.maximum_level]
        if refine:
            # RefineNode of every listed leaf at once: the same nodes and slots
            # as one call per leaf, one batched clear and one prediction.
            for key in refine:
                self.topo.set_leaf(key, False)
            kids = [c for key in refine for c in self.topo.children(key)]
            self._new_nodes(kids)
            self.be.predict([self._link(c) for c in kids], (gpu.HS, gpu.IC) * self.dim)
        self.halo_update(levels_ascending[1:])
        self._coarsen([k for k in remove if k[0] != 1])

    # ------------------------------------------------------------ time step
    def compute_timestep(self) -> float:
        run_time = self.current_run_time
        for dt in self.micro:
            run_time += dt
        dx_finest = self.cell_size(self.maximum_level)
        slots = self._cached("all_leaf_slots", -1, lambda: self._slots(self.topo.leaves()))
        return self.be.timestep(slots, self.cfl, dx_finest, run_time, self.t_end)

    # ------------------------------------------------------- initialisation
    def create_new_simulation(self):
        for level in self.all_levels:
            if level > 0:
                kids = []
                for pkey in self.topo.ids_on_level(level - 1):
                    self.topo.set_leaf(pkey, False)
                    kids += self.topo.children(pkey)
                self._new_nodes(kids)
            else:
                nb0 = self.topo.nb0
                self._new_nodes([(0, i, j, k) for i in range(nb0[0]) for j in range(nb0[1])
                                 for k in range(nb0[2])])
            keys = self.topo.ids_on_level(level)
            self.be.set_internal(self._slots(keys),
                                 np.stack([self.initial(level, key) for key in keys]))
            self.halo_update_on_level(level, False)
            if level > 1:
                remove, _ = self.determine_remeshing_nodes([level - 1])
                self._coarsen(remove)
        self.swap(self.all_levels)

    # -------------------------------------------------------------- advance
    def advance(self):
        """One macro step: 2^maximum_level micro steps on the finest level."""
        lmax = self.maximum_level
        levels_desc = list(reversed(self.all_levels))
        for timestep in range(1 << lmax):
            self.micro.append(self.compute_timestep())
            for stage in range(RK_STAGES):
                self.compute_right_hand_side(levels_desc, stage)
                self.average_material(levels_desc)
                self.halo_update(self.all_levels)
                levels_desc = lts.levels_to_update(timestep, lmax)
                levels_asc = list(reversed(levels_desc))
                self.integrate(levels_desc, stage)
                self.average_material(levels_desc[:-1])
                last = stage == RK_STAGES - 1
                if last:
                    self.jump_flux_adjustment(levels_desc)
                self.halo_update(levels_asc, cut_jumps=True)
                if last:
                    self.remesh(levels_asc)
                self.swap(levels_desc)
        # ComputeLoop: ResetAllJumpBuffers, then FinishMacroTimestep.
        self.be.reset_jump_all(list(self.slot.values()))
        s = 0.0
        for dt in self.micro:
            s += dt
        self.current_run_time += s
        last_dt = self.micro[-1]
        self.micro = []
        self.macro_steps += 1
        return last_dt

    def run(self, max_macro_steps: int = 1 << 30):
        """ComputeLoop (modular_algorithm_assembler.cpp:97-195): macro steps
        until t_end, or until the last micro step falls below DBL_EPSILON."""
        while self.current_run_time < self.t_end and self.macro_steps < max_macro_steps:
            last = self.advance()
            if 0.0 < last < 2.220446049250313e-16:
                break
        return self.macro_steps

    # --------------------------------------------------------------- restart
    def write_restart(self, path):
        """Every node of every level with its haloed Average buffer and its
        reference node id (RestartManager::WriteRestartFile,
        restart_manager.cpp:300-470); call between macro steps."""
        from . import restart

        keys = sorted(self.slot)
        ids = [restart.node_id(*k) for k in keys]
        blocks = self.be.get_blocks(self._slots(keys))
        return restart.write_nodes(path, self.current_run_time, self.dim, ids, blocks, self.gamma,
                                   self.b, self.maximum_level)

    def restore(self, path):
        """RestoreSimulation + FinalizeSimulationRestart (restart_manager.cpp:
        91-290, modular_algorithm_assembler.cpp:247-284): the nodes of the
        file created with their conservatives in the right-hand-side buffers,
        leaves where no child exists, then the swap of all levels."""
        from . import restart

        if self.slot:
            raise ValueError("restore needs a solver that holds no nodes yet")
        t, _, ids, cons = restart.read_nodes(path, self.dim, self.maximum_level)
        keys = [restart.decode_node_id(int(i)) for i in ids]
        keys = [(lv, i, j, k if self.dim == 3 else 0) for lv, i, j, k in keys]
        self._new_nodes(keys)
        for key in keys:
            self.topo.set_leaf(key, not any(self.topo.exists(c) for c in self.topo.children(key)))
        self.be.set_rhs_blocks(self._slots(keys), cons)
        self.swap(self.all_levels)
        self.current_run_time = float(t)
        self.micro = []

    # ---------------------------------------------------------------- output
    def leaf_data(self):
        """(keys, internal conservatives [n][NE][16][16][16|1]) of all leaves
        from the Average buffers."""
        keys = self.topo.leaves()
        return keys, self.be.get_internal(self._slots(keys))

    def leaf_cells(self):
        """Cell centres, volumes and conservatives of every leaf cell, as the
        finest-level output mesh lays them out (standard_finest_level_mesh_
        generator.cpp:160-300): a leaf cell of level l spans dx_l per axis in
        2D plus node_size / 2^l along z."""
        keys, cons = self.leaf_data()
        dim = self.dim
        icz = gpu.IC if dim == 3 else 1
        centres, vols, vals = [], [], []
        for key, c in zip(keys, cons):
            lv = key[0]
            dx = self.cell_size(lv)
            ii, jj, kk = np.meshgrid(np.arange(gpu.IC), np.arange(gpu.IC), np.arange(icz),
                                     indexing="ij")
            x = (key[1] * gpu.IC + ii + 0.5) * dx
            y = (key[2] * gpu.IC + jj + 0.5) * dx
            z = (key[3] * icz + kk + 0.5) * dx
            centres.append(np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1))
            dz = dx if dim == 3 else self.node_size / float(1 << lv)
            vols.append(np.full(x.size, dx * dx * dz))
            vals.append(c.reshape(dim + 2, -1).T)
        return np.concatenate(centres), np.concatenate(vols), np.concatenate(vals)


class GpuBackend:
    """The driver's cell operations as gfx950 kernels on device pools."""

    def __init__(self, dim: int, capacity: int, flux: int = gpu.FLUX_HLLC,
                 precision: int = gpu.FP64, gamma: float = 1.4, background_pressure: float = 0.0):
        import torch

        self.torch = torch
        self.lib = gpu.load_library()
        self.dim = dim
        self.capacity = int(capacity)
        self.flux, self.precision = flux, precision
        self.gamma, self.b = gamma, background_pressure
        tz = gpu.TC if dim == 3 else 1
        icz = gpu.IC if dim == 3 else 1
        shape = (self.capacity, dim + 2, gpu.TC, gpu.TC, tz)
        self.avg = torch.zeros(shape, dtype=torch.float64, device="cuda")
        self.rhs = torch.zeros_like(self.avg)
        self.init = torch.zeros_like(self.avg)
        sshape = (self.capacity, 6, dim + 2, gpu.IC, icz)
        self.jflux = torch.zeros(sshape, dtype=torch.float64, device="cuda")
        self.jcons = torch.zeros_like(self.jflux)
        self.state = torch.zeros(ctypes.sizeof(gpu.CaseState), dtype=torch.uint8, device="cuda")
        # Every operation is queued on the stream current at construction (the
        # context's stream too); torch work of the backend runs under it.
        self.stream = torch.cuda.current_stream()
        self.st = ctypes.c_void_p(self.stream.cuda_stream)
        self.ctx = ctypes.c_void_p()
        gpu.check(self.lib.alpaca_context_create(self.st, ctypes.byref(self.ctx)),
                  "context_create")
        self.lti = lts.LevelTimeIntegrator(dim)
        self.mrop = mr.Multiresolution()
        self._memo = {}

    def close(self):
        if self.ctx:
            gpu.check(self.lib.alpaca_context_destroy(self.ctx), "context_destroy")
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers
    def _device(self, obj, kind, make):
        """Device copy of a host table the solver keeps between calls (its
        per-level tables live until the tree changes): uploaded once per
        object.  The memo holds the object too, so its id cannot be reused
        while the entry exists."""
        key = (id(obj), kind)
        hit = self._memo.get(key)
        if hit is not None and hit[0] is obj:
            return hit[1]
        if len(self._memo) > 4096:
            self._memo.clear()
        t = self._upload(make(obj))
        self._memo[key] = (obj, t)
        return t

    def _upload(self, arr):
        """Host array to the device without waiting for the queued work: a
        pinned staging copy, then an asynchronous copy on the backend's
        stream (torch keeps the staging block until that copy is done)."""
        host = self.torch.from_numpy(np.ascontiguousarray(arr))
        with self.torch.cuda.stream(self.stream):
            return host.pin_memory().to("cuda", non_blocking=True)

    def _idx(self, slots):
        return self._device(slots, "i64", lambda o: np.asarray(o, dtype=np.int64))

    def _list(self, slots):
        return self._device(slots, "i32", lambda o: np.asarray(o, dtype=np.int32))

    def _links(self, links):
        return self._device(links, "links", mr.make_links)

    def _p(self, t):
        return ctypes.c_void_p(t.data_ptr())

    def _desc(self, n, cell_size):
        return gpu.make_desc(self.dim, [n, 1, 1], [0, 0, 0], flux=self.flux,
                             precision=self.precision, gamma=self.gamma, b=self.b,
                             cell_size=cell_size)

    # -- node data
    def clear(self, slots):
        idx = self._upload(np.asarray(slots, dtype=np.int64))
        with self.torch.cuda.stream(self.stream):
            for t in (self.avg, self.rhs, self.init, self.jflux, self.jcons):
                t.index_fill_(0, idx, 0.0)

    def set_internal(self, slots, cons):
        """Internal cells of the right-hand-side buffers (ImposeInitialCondition)."""
        h, i = gpu.HS, gpu.IC
        src = self._upload(cons)
        idx = self._idx(slots)
        with self.torch.cuda.stream(self.stream):
            if self.dim == 3:
                self.rhs[idx, :, h:h + i, h:h + i, h:h + i] = src
            else:
                self.rhs[idx, :, h:h + i, h:h + i, :] = src

    def get_blocks(self, slots):
        with self.torch.cuda.stream(self.stream):
            return self.avg[self._idx(slots)].cpu().numpy()

    def set_rhs_blocks(self, slots, blocks):
        src = self._upload(blocks)
        with self.torch.cuda.stream(self.stream):
            self.rhs.index_copy_(0, self._idx(slots),
                                 src.reshape((len(slots),) + tuple(self.rhs.shape[1:])))

    def get_internal(self, slots):
        h, i = gpu.HS, gpu.IC
        with self.torch.cuda.stream(self.stream):
            a = self.avg[self._idx(slots)]
            if self.dim == 3:
                return a[:, :, h:h + i, h:h + i, h:h + i].cpu().numpy()
            return a[:, :, h:h + i, h:h + i, :].cpu().numpy()

    def swap(self, slots):
        """SwapBuffers as a one-way copy of the new state (RightHandSide) into
        the Average buffers: what the exchange leaves in the RightHandSide
        buffers is never read before it is overwritten (a leaf's by its next
        right-hand side, any other node's internal cells by the averaging of
        its children and its halo cells by the halo update that follows), so
        the copy gives the reference's results at half the traffic of the
        exchange.  The CPU backend exchanges, and the two backends' states are
        compared bit for bit."""
        lst = self._list(slots)
        gpu.check(self.lib.alpaca_copy_blocks_list(self.dim, self._p(self.rhs), self._p(self.avg),
                                                   self.capacity, self._p(lst), len(slots),
                                                   self.st), "copy_blocks_list")

    def fill_initial(self, slots):
        self.lti.fill_initial_buffer(self.avg, self.init, self._list(slots), 0, self.st)

    # -- stage operations
    def right_hand_side(self, slots, cell_size):
        lst = self._list(slots)
        n = len(slots)
        desc = self._desc(n, cell_size)
        gpu.check(self.lib.alpaca_ctx_right_hand_side_jump_fluxes_list(
            self.ctx, ctypes.byref(desc), self.capacity, self._p(lst), n, self._p(self.avg),
            self._p(self.rhs), self._p(self.jflux)), "rhs")

    def prepare(self, slots, stage):
        if stage > 0:
            self.lti.prepare_buffer_for_integration(self.avg, self.init, self._list(slots), stage,
                                                    self.st)

    def lts_timestep(self, micro, n, stage, jump):
        self.lti.micro_timestep_sizes = list(micro)
        return self.lti.timestep(stage, n, jump)

    def integrate_node(self, slots, jump_timestep, timestep):
        lst = self._list(slots)
        nb = self.capacity
        gpu.check(self.lib.alpaca_lts_integrate_jump_conservatives(
            self.dim, self._p(self.jcons), self._p(self.jflux), nb, self._p(lst), len(slots),
            jump_timestep, self.st), "lts_integrate_jump_conservatives")
        gpu.check(self.lib.alpaca_lts_integrate(self.dim, self._p(self.avg), self._p(self.rhs), nb,
                                                self._p(lst), len(slots), None, timestep,
                                                self.st), "lts_integrate")

    def integrate_jump_halos(self, by_loc, timestep):
        """IntegrateJumpHalos for every (node, location) of a level at once;
        by_loc maps a BoundaryLocation to the pool blocks with a jump there."""
        ent = self._device(by_loc, "entries", lambda o: np.array(
            [(s, loc) for loc, slots in o.items() for s in slots], dtype=np.int32).reshape(-1, 2))
        gpu.check(self.lib.alpaca_lts_integrate_jump_halos(
            self.dim, self._p(self.avg), self._p(self.rhs), self.capacity, self._p(ent),
            ent.shape[0], timestep, self.st), "lts_integrate_jump_halos")

    def average(self, links):
        self.mrop.average(self.rhs, self.dim, self._links(links), self.st)

    def average_jump(self, links):
        self.mrop.average_jump_buffer(self.jcons, self.dim, self._links(links), self.st)

    def halo_update(self, slots, table, jump_links, jump_locs, side_bc, second=None):
        """second: (slots, table rows) of the blocks the pass after the jump
        update can write (None: no such block)."""
        lst = self._list(slots)
        tab = self._device(table, "table", lambda o: np.ascontiguousarray(o, dtype=np.int32))
        bc = (ctypes.c_int * 6)(*(list(side_bc) + [0] * (6 - len(side_bc))))
        st = self.st
        gpu.check(self.lib.alpaca_fill_halos_list(self.dim, self._p(self.rhs), self.capacity,
                                                  self._p(lst), len(slots), self._p(tab), bc, 0,
                                                  st), "fill_halos_list")
        if len(jump_links):
            locs = self._device(jump_locs, "i32", lambda o: np.asarray(o, dtype=np.int32))
            self.mrop.jump_halo_update(self.rhs, self.dim, self._links(jump_links), locs, self.st)
        if second is not None:
            lst2 = self._list(second[0])
            tab2 = self._device(second[1], "table",
                                lambda o: np.ascontiguousarray(o, dtype=np.int32))
            gpu.check(self.lib.alpaca_fill_halos_list(self.dim, self._p(self.rhs), self.capacity,
                                                      self._p(lst2), len(second[0]),
                                                      self._p(tab2), bc, 1, st),
                      "fill_halos_list (after jumps)")

    def jump_flux_adjustment(self, entries, one_cell_size):
        ent = self._device(entries, "table", lambda o: np.ascontiguousarray(o, dtype=np.int32))
        lts.jump_flux_adjustment(self.dim, self.rhs, self.jcons, ent, one_cell_size, self.st)

    def reset_jump_conservatives(self, slots):
        lts.reset_jump_conservatives(self.dim, self.jcons, self._list(slots), self.st)

    def reset_jump_all(self, slots):
        lst = self._list(slots)
        lts.reset_jump_conservatives(self.dim, self.jcons, lst, self.st)
        lts.reset_jump_conservatives(self.dim, self.jflux, lst, self.st)

    def predict(self, links, rng):
        self.mrop.prediction(self.rhs, self.dim, self._links(links), rng, self.st)

    def decisions(self, links, thresholds):
        m = mr.Multiresolution(norm=mr.LINF)
        lk = self._links(links)
        n = lk.shape[0]
        det = self.torch.empty(n, dtype=self.torch.float64, device="cuda")
        dec = self.torch.empty(n, dtype=self.torch.int32, device="cuda")
        eps = (ctypes.c_double * len(thresholds))(*thresholds)
        gpu.check(self.lib.alpaca_mr_child_needs_remeshing(
            self.dim, self._p(self.rhs), self.capacity, self._p(lk), n, m.norm, eps,
            len(thresholds), self._p(det), self._p(dec), self.st),
            "mr_child_needs_remeshing")
        with self.torch.cuda.stream(self.stream):
            return dec.cpu().numpy()

    def timestep(self, leaf_slots, cfl, dx_finest, run_time, t_end):
        lst = self._list(leaf_slots)
        desc = self._desc(len(leaf_slots), dx_finest)
        st = self.st
        sp = self._p(self.state)
        gpu.check(self.lib.alpaca_case_state_init(sp, run_time, t_end, cfl, st), "state_init")
        gpu.check(self.lib.alpaca_timestep_list(ctypes.byref(desc), self._p(self.avg),
                                                self.capacity, self._p(lst), len(leaf_slots), sp,
                                                st), "timestep")
        with self.torch.cuda.stream(self.stream):
            raw = self.state.cpu().numpy().tobytes()
        rec = gpu.CaseState.from_buffer_copy(raw)
        return rec.dt
