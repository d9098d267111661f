"""Restart snapshots in the reference's HDF5 layout (restart_manager.cpp:
300-480 writes, 91-290 reads) and the snapshot schedule of
InputOutputManager::WriteRestartFile / RestoreSimulationFromSnapshot
(input_output_manager.cpp:399-520).

One file holds

  /node_info_data/NodeIds                  [nodes]  unsigned long long
  /node_info_data/NumberOfMaterials        [nodes]  unsigned short (1)
  /node_info_data/Materials                [blocks] unsigned short (0)
  /node_info_data/NumberOfInterfaceBlocks  [nodes]  unsigned short (0)
  /node_cell_data/Conservatives  [blocks][NE][TCX][TCY][TCZ] double (average buffer)
  /node_cell_data/PrimeStates    [blocks][NE][TCX][TCY][TCZ] double (rho, p, u, v, (w))
  /node_cell_data/Levelset       [0][TCX][TCY][TCZ] double (no interface blocks)
  /node_cell_data/InterfaceTags  [0][TCX][TCY][TCZ] char
  /simulation_data  attributes Time, Dimensions, InternalCells, HaloSize,
                    MaximumLevel (unsigned int), Length/Velocity/Density/
                    TemperatureReference (double)

with each block's haloed buffer (24 cells a side, 4 of them halo; TCZ = 1 in
2D) as the device keeps it.  The reference writes its nodes in hash-map order
and places each block by its id when reading; so does read() here.

A level-0 node id encodes the block position: IdSeed() = 0x1400000 plus the
block indices' bits interleaved x, y, z from bit 0 (the neighbour-id
arithmetic of id_information.cpp steps these interleaved counters; the ids of
topology_manager.cpp:109-140's creation loop pin the encoding).
"""
from __future__ import annotations

import ctypes
import os
import time as _time

import numpy as np

from . import cases, h5output
from .h5output import H5P_DEFAULT, H5S_SCALAR, _lib, _ok, _type, _write_dataset

IC, HS, TC = 16, 4, 24
ID_SEED = 0x1400000
POSITION_BITS = 21          # 7 interleaved (x, y, z) triples below the seed's bits
MAX_BLOCKS_PER_DIRECTION = 1 << (POSITION_BITS // 3)


class RestartError(RuntimeError):
    pass


def level0_id(bx: int, by: int, bz: int) -> int:
    if not all(0 <= b < MAX_BLOCKS_PER_DIRECTION for b in (bx, by, bz)):
        raise RestartError(f"block ({bx}, {by}, {bz}) outside the level-0 id range")
    v = 0
    for k in range(POSITION_BITS // 3):
        v |= (((bx >> k) & 1) << (3 * k) | ((by >> k) & 1) << (3 * k + 1)
              | ((bz >> k) & 1) << (3 * k + 2))
    return ID_SEED + v


def decode_level0_id(node_id: int):
    v = int(node_id) - ID_SEED
    if v < 0 or v >> POSITION_BITS:
        raise RestartError(f"node id {int(node_id):#x} is not a level-0 id")
    b = [0, 0, 0]
    for k in range(POSITION_BITS // 3):
        for d in range(3):
            b[d] |= ((v >> (3 * k + d)) & 1) << k
    return tuple(b)


def node_id(level: int, bx: int, by: int, bz: int) -> int:
    """Id of the node at block position (bx, by, bz) of `level`: its level-0
    ancestor's id followed by one sibling position (bit 0 x, bit 1 y, bit 2 z)
    per level, IdsOfChildren = (parent << 3) + position (id_information.h:
    150-192)."""
    nid = level0_id(bx >> level, by >> level, bz >> level)
    for m in range(level - 1, -1, -1):
        pos = ((bx >> m) & 1) | (((by >> m) & 1) << 1) | (((bz >> m) & 1) << 2)
        nid = (nid << 3) + pos
    return nid


def level_of_node(nid: int) -> int:
    """LevelOfNode (id_information.h:59-67): the shadow bits cut, three bits
    per level above the seed's 0b1010."""
    w = int(nid) >> POSITION_BITS
    level = 0
    while w > 10:
        w >>= 3
        level += 1
    return level


def decode_node_id(nid: int):
    """(level, bx, by, bz) of a node id."""
    level = level_of_node(nid)
    b = list(decode_level0_id(int(nid) >> (3 * level)))
    for m in range(level - 1, -1, -1):
        pos = (int(nid) >> (3 * m)) & 7
        for d in range(3):
            b[d] = (b[d] << 1) | ((pos >> d) & 1)
    return (level, b[0], b[1], b[2])


def creation_order(nb):
    """(package block index, node id) in the reference's level-0 creation
    order: x fastest, then y, then z (nb: 2 or 3 block counts)."""
    nb3 = [int(x) for x in nb] + [1] * (3 - len(nb))
    out = []
    for bz in range(nb3[2]):
        for by in range(nb3[1]):
            for bx in range(nb3[0]):
                out.append(((bx * nb3[1] + by) * nb3[2] + bz, level0_id(bx, by, bz)))
    return out


def _nb3(dim, nb):
    nb3 = [int(x) for x in nb] + [1] * (3 - len(nb))
    if dim == 2:
        nb3[2] = 1
    return nb3


def blocks_from_domain(cons, dim, nb):
    """Domain array [NE][NX][NY][NZ] -> haloed blocks [n][NE][24][24][TZ]
    (package block order, halos zero)."""
    nb3 = _nb3(dim, nb)
    ne = dim + 2
    tz, iz, hz = (TC, IC, HS) if dim == 3 else (1, 1, 0)
    g = np.asarray(cons, dtype=np.float64).reshape(ne, nb3[0], IC, nb3[1], IC, nb3[2], iz)
    out = np.zeros((nb3[0], nb3[1], nb3[2], ne, TC, TC, tz))
    out[..., HS:HS + IC, HS:HS + IC, hz:hz + iz] = g.transpose(1, 3, 5, 0, 2, 4, 6)
    return out.reshape(-1, ne, TC, TC, tz)


def domain_from_blocks(blocks, dim, nb):
    """Haloed blocks [n][NE][24][24][TZ] (package order) -> domain array."""
    nb3 = _nb3(dim, nb)
    ne = dim + 2
    tz, iz, hz = (TC, IC, HS) if dim == 3 else (1, 1, 0)
    b = np.asarray(blocks, dtype=np.float64).reshape(nb3[0], nb3[1], nb3[2], ne, TC, TC, tz)
    b = b[..., HS:HS + IC, HS:HS + IC, hz:hz + iz]
    return np.ascontiguousarray(b.transpose(3, 0, 4, 1, 5, 2, 6).reshape(
        ne, nb3[0] * IC, nb3[1] * IC, nb3[2] * iz))


def _attr(loc, name, value, h5type, ctype):
    lib = _lib()
    s = _ok(lib.H5Screate(H5S_SCALAR), "H5Screate")
    try:
        a = _ok(lib.H5Acreate2(loc, name.encode(), h5type, s, H5P_DEFAULT, H5P_DEFAULT),
                f"H5Acreate2 {name}")
        v = ctype(value)
        try:
            _ok(lib.H5Awrite(a, h5type, ctypes.byref(v)), f"H5Awrite {name}")
        finally:
            lib.H5Aclose(a)
    finally:
        lib.H5Sclose(s)


def _group(fid, name):
    return _ok(_lib().H5Gcreate2(fid, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               f"H5Gcreate2 {name}")


def write(path, time, dim, nb, blocks, gamma, b=0.0, order=None):
    """Write a restart file.  blocks: haloed blocks [n][NE][24][24][TZ] in the
    package's block order (bx * ny + by) * nz + bz; order: the (block index,
    node id) sequence to write, default the reference's creation order.
    Returns the path."""
    nb3 = _nb3(dim, nb)
    tz = TC if dim == 3 else 1
    blocks = np.asarray(blocks, dtype=np.float64).reshape(-1, dim + 2, TC, TC, tz)
    if blocks.shape[0] != nb3[0] * nb3[1] * nb3[2]:
        raise RestartError(f"{blocks.shape[0]} blocks for a {nb3} grid")
    if order is None:
        order = creation_order(nb3)
    idx = [k for k, _ in order]
    ids = np.array([i for _, i in order], dtype=np.uint64)
    return write_nodes(path, time, dim, ids, blocks[idx], gamma, b, 0)


def write_nodes(path, time, dim, ids, blocks, gamma, b=0.0, maximum_level=0):
    """Write a restart file of the given nodes (ids) and their haloed Average
    buffers [n][NE][24][24][TZ] (RestartManager::WriteRestartFile writes every
    node of every level, restart_manager.cpp:406-431)."""
    lib = _lib()
    tz = TC if dim == 3 else 1
    ids = np.asarray(ids, dtype=np.uint64)
    cons = np.ascontiguousarray(np.asarray(blocks, dtype=np.float64).reshape(-1, dim + 2, TC, TC, tz))
    idx = range(cons.shape[0])
    # halo cells that no halo update reaches (edges and corners in 3D) hold
    # zeros on the device; their prime states are written as zeros
    with np.errstate(divide="ignore", invalid="ignore"):
        rho, vel, p = cases.conservatives_to_primes(dim, gamma, b, np.moveaxis(cons, 1, 0))
        prim = np.stack([rho, p] + vel, axis=1)
    prim = np.ascontiguousarray(np.where((cons[:, :1] != 0.0), prim, 0.0))
    n = len(idx)
    u16, u32 = _type("H5T_NATIVE_USHORT_g"), _type("H5T_NATIVE_UINT_g")
    u64, f64 = _type("H5T_NATIVE_ULLONG_g"), _type("H5T_NATIVE_DOUBLE_g")
    i8 = _type("H5T_NATIVE_SCHAR_g")
    fid = _ok(lib.H5Fcreate(str(path).encode(), h5output.H5F_ACC_TRUNC, H5P_DEFAULT,
                            H5P_DEFAULT), "H5Fcreate")
    try:
        g = _group(fid, "node_info_data")
        _write_dataset(g, "NodeIds", ids, u64)
        _write_dataset(g, "NumberOfMaterials", np.ones(n, dtype=np.uint16), u16)
        _write_dataset(g, "Materials", np.zeros(n, dtype=np.uint16), u16)
        _write_dataset(g, "NumberOfInterfaceBlocks", np.zeros(n, dtype=np.uint16), u16)
        lib.H5Gclose(g)
        g = _group(fid, "node_cell_data")
        _write_dataset(g, "Conservatives", cons, f64)
        _write_dataset(g, "PrimeStates", prim, f64)
        _write_dataset(g, "Levelset", np.zeros((0, TC, TC, tz)), f64)
        _write_dataset(g, "InterfaceTags", np.zeros((0, TC, TC, tz), dtype=np.int8), i8)
        lib.H5Gclose(g)
        g = _group(fid, "simulation_data")
        _attr(g, "Time", time, f64, ctypes.c_double)
        _attr(g, "Dimensions", dim, u32, ctypes.c_uint)
        _attr(g, "InternalCells", IC, u32, ctypes.c_uint)
        _attr(g, "HaloSize", HS, u32, ctypes.c_uint)
        _attr(g, "MaximumLevel", int(maximum_level), u32, ctypes.c_uint)
        for name in ("LengthReference", "VelocityReference", "DensityReference",
                     "TemperatureReference"):
            _attr(g, name, 1.0, f64, ctypes.c_double)
        lib.H5Gclose(g)
    finally:
        lib.H5Fclose(fid)
    return str(path)


def _read(fid, path, dtype, h5name):
    lib = _lib()
    d = _ok(lib.H5Dopen2(fid, path.encode(), H5P_DEFAULT), f"H5Dopen2 {path}")
    try:
        s = lib.H5Dget_space(d)
        nd = lib.H5Sget_simple_extent_ndims(s)
        dims = (h5output.hsize_t * nd)()
        lib.H5Sget_simple_extent_dims(s, dims, None)
        lib.H5Sclose(s)
        out = np.empty(tuple(int(x) for x in dims), dtype=dtype)
        if out.size:
            _ok(lib.H5Dread(d, _type(h5name), h5output.H5S_ALL, h5output.H5S_ALL, H5P_DEFAULT,
                            out.ctypes.data_as(ctypes.c_void_p)), f"H5Dread {path}")
        return out
    finally:
        lib.H5Dclose(d)


def _read_attr(fid, name, h5name, ctype):
    lib = _lib()
    a = _ok(lib.H5Aopen_by_name(fid, b"simulation_data", name.encode(), H5P_DEFAULT,
                                H5P_DEFAULT), f"H5Aopen_by_name {name}")
    v = ctype()
    try:
        _ok(lib.H5Aread(a, _type(h5name), ctypes.byref(v)), f"H5Aread {name}")
    finally:
        lib.H5Aclose(a)
    return v.value


def read_nodes(path, dim=None, maximum_level=0):
    """Read a restart file of any level: (time, dim, ids, blocks [n][NE][24][24][TZ]
    in file order).  Refuses another dimension (when given), block size, halo
    width, maximum level or non-unit references, several materials and
    interface blocks (RestoreSimulation's checks, restart_manager.cpp:91-170)."""
    lib = _lib()
    fid = _ok(lib.H5Fopen(str(path).encode(), h5output.H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    try:
        u32, f64 = "H5T_NATIVE_UINT_g", "H5T_NATIVE_DOUBLE_g"
        fdim = _read_attr(fid, "Dimensions", u32, ctypes.c_uint)
        if dim is not None and fdim != dim:
            raise RestartError(f"restart file is {fdim}D, the case is {dim}D")
        if _read_attr(fid, "InternalCells", u32, ctypes.c_uint) != IC:
            raise RestartError("restart file has another number of internal cells")
        if _read_attr(fid, "HaloSize", u32, ctypes.c_uint) != HS:
            raise RestartError("restart file has another halo size")
        if _read_attr(fid, "MaximumLevel", u32, ctypes.c_uint) != maximum_level:
            raise RestartError("restart file has another maximum level")
        for name in ("LengthReference", "VelocityReference", "DensityReference",
                     "TemperatureReference"):
            if _read_attr(fid, name, f64, ctypes.c_double) != 1.0:
                raise RestartError(f"restart file {name} is not 1")
        t = _read_attr(fid, "Time", f64, ctypes.c_double)
        ids = _read(fid, "/node_info_data/NodeIds", np.uint64, "H5T_NATIVE_ULLONG_g")
        nmat = _read(fid, "/node_info_data/NumberOfMaterials", np.uint16, "H5T_NATIVE_USHORT_g")
        nint = _read(fid, "/node_info_data/NumberOfInterfaceBlocks", np.uint16,
                     "H5T_NATIVE_USHORT_g")
        if np.any(nmat != 1) or np.any(nint != 0):
            raise RestartError("restart file holds several materials or interface blocks")
        cons = _read(fid, "/node_cell_data/Conservatives", np.float64, "H5T_NATIVE_DOUBLE_g")
    finally:
        lib.H5Fclose(fid)
    tz = TC if fdim == 3 else 1
    if cons.shape[1:] != (fdim + 2, TC, TC, tz) or cons.shape[0] != ids.shape[0]:
        raise RestartError(f"Conservatives has shape {cons.shape}")
    return t, int(fdim), ids, cons


def read(path, dim=None):
    """Read a level-0 restart file: (time, dim, nb, blocks [n][NE][24][24][TZ]
    in the package's block order).  Besides read_nodes' checks it refuses an
    incomplete level-0 grid."""
    t, fdim, ids, cons = read_nodes(path, dim, 0)
    pos = [decode_level0_id(i) for i in ids]
    nb3 = [max(p[d] for p in pos) + 1 for d in range(3)]
    if len(set(pos)) != len(pos) or len(pos) != nb3[0] * nb3[1] * nb3[2]:
        raise RestartError("restart file does not hold a complete level-0 block grid")
    blocks = np.empty_like(cons)
    for k, (bx, by, bz) in enumerate(pos):
        blocks[(bx * nb3[1] + by) * nb3[2] + bz] = cons[k]
    return t, int(fdim), nb3[:fdim], blocks


def snapshot_name(t, time_naming_factor=1.0):
    """restart_<t>.h5 with t as std::to_string prints a double (%f)."""
    return "restart_%f.h5" % (t * time_naming_factor)


class SnapshotSchedule:
    """When to write restart snapshots (InputOutputManager::WriteRestartFile):
    on the first time stamp not above the current time (all stamps up to it
    are then consumed), every interval_s wall seconds, or when forced.
    Interval snapshots beyond `keep` are deleted oldest first; stamp
    snapshots are kept.  A link latest_restart_snapshot names the newest."""

    def __init__(self, folder, stamps=(), interval_s=0, keep=2, time_naming_factor=1.0,
                 restored_time=None):
        self.folder = folder
        self.stamps = sorted(stamps)
        if restored_time is not None:   # RestoreSimulationFromSnapshot drops stamps <= t
            self.stamps = [s for s in self.stamps if s > restored_time]
        self.interval_s = int(interval_s)
        self.keep = int(keep)
        self.factor = time_naming_factor
        self.last_wall = _time.time()
        self.interval_files = []
        self.written = []

    def due(self, t, force=False):
        by_interval = False
        if self.interval_s > 0 and _time.time() - self.last_wall >= self.interval_s:
            by_interval = True
        by_stamp = bool(self.stamps) and self.stamps[0] <= t
        return by_interval, by_stamp, by_interval or by_stamp or force

    def maybe_write(self, t, write_file, force=False):
        """Call after each macro step with the current time; write_file(path)
        writes the snapshot.  Returns the path written or None."""
        by_interval, by_stamp, go = self.due(t, force)
        if not go:
            return None
        if by_interval:
            self.last_wall = _time.time()
        self.stamps = [s for s in self.stamps if s > t]
        os.makedirs(self.folder, exist_ok=True)
        path = os.path.join(self.folder, snapshot_name(t, self.factor))
        write_file(path)
        link = os.path.join(self.folder, "latest_restart_snapshot")
        if os.path.lexists(link):
            os.remove(link)
        os.symlink(os.path.basename(path), link)
        if not by_stamp:
            if self.keep > 0 and len(self.interval_files) == self.keep:
                old = self.interval_files.pop(0)
                if os.path.exists(old) and old != path:
                    os.remove(old)
            self.interval_files.append(path)
        self.written.append(path)
        return path
