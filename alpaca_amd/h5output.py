"""Simulation output in the reference's HDF5 + XDMF layout.

The reference's OutputWriter (output_writer.cpp:143-240) writes one HDF5 file
per output time with

  /metadata            attribute ``time`` (double)
  /mesh_topology/cell_vertex_IDs           [cells][8]  unsigned long long
  /mesh_topology/cell_vertex_coordinates   [vertices][3] double
  /cell_data/density   [cells][1][1] double
  /cell_data/velocity  [cells][3][1] double
  /cell_data/pressure  [cells][1][1] double

on the finest-level mesh (StandardFinestLevelMeshGenerator,
standard_finest_level_mesh_generator.cpp:17-32, 91-190): vertices on the
(nx+1)(ny+1)(nz+1) lattice, x fastest, at ``index * cell_size``; cells block
by block, k outer and i fastest inside a block, each a hexahedron of the 8
lattice vertices in the order (x,y,z) = 000, 100, 110, 010, 001, 101, 111,
011.  2D runs are one cell thick (two vertex layers, z = 0 and the cell
size).  Beside it an XDMF file (xdmf_utilities.cpp) makes the data readable
by ParaView and the reference's post-processing (e.g. sod_analysis.py, which
reads the mesh, density, velocity, pressure and time above).

HDF5 is reached through its C library with ctypes (no Python binding is
installed); ``ALPACA_HDF5_LIB`` names the shared library if it is not found
on the default paths.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

IC = 16
_LIB = None
_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so",
               "/usr/lib/x86_64-linux-gnu/libhdf5.so")
H5F_ACC_RDONLY, H5F_ACC_TRUNC, H5P_DEFAULT, H5S_ALL, H5S_SCALAR = 0, 2, 0, 0, 0
hid_t = ctypes.c_int64
hsize_t = ctypes.c_ulonglong


class Hdf5Error(RuntimeError):
    pass


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    paths = [os.environ.get("ALPACA_HDF5_LIB"), ctypes.util.find_library("hdf5"), *_CANDIDATES]
    for p in paths:
        if not p:
            continue
        try:
            lib = ctypes.CDLL(p)
        except OSError:
            continue
        break
    else:
        raise Hdf5Error("libhdf5 not found (set ALPACA_HDF5_LIB)")
    sig = {
        "H5open": ([], ctypes.c_int),
        "H5Fcreate": ([ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t], hid_t),
        "H5Fopen": ([ctypes.c_char_p, ctypes.c_uint, hid_t], hid_t),
        "H5Fclose": ([hid_t], ctypes.c_int),
        "H5Gcreate2": ([hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t], hid_t),
        "H5Gclose": ([hid_t], ctypes.c_int),
        "H5Screate_simple": ([ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.c_void_p], hid_t),
        "H5Screate": ([ctypes.c_int], hid_t),
        "H5Sclose": ([hid_t], ctypes.c_int),
        "H5Sget_simple_extent_ndims": ([hid_t], ctypes.c_int),
        "H5Sget_simple_extent_dims": ([hid_t, ctypes.POINTER(hsize_t), ctypes.c_void_p],
                                      ctypes.c_int),
        "H5Dcreate2": ([hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Dopen2": ([hid_t, ctypes.c_char_p, hid_t], hid_t),
        "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p], ctypes.c_int),
        "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p], ctypes.c_int),
        "H5Dget_space": ([hid_t], hid_t),
        "H5Dclose": ([hid_t], ctypes.c_int),
        "H5Acreate2": ([hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Aopen_by_name": ([hid_t, ctypes.c_char_p, ctypes.c_char_p, hid_t, hid_t], hid_t),
        "H5Awrite": ([hid_t, hid_t, ctypes.c_void_p], ctypes.c_int),
        "H5Aread": ([hid_t, hid_t, ctypes.c_void_p], ctypes.c_int),
        "H5Aclose": ([hid_t], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if lib.H5open() < 0:
        raise Hdf5Error("H5open failed")
    _LIB = lib
    return lib


def _type(name):
    return hid_t.in_dll(_lib(), name).value


def _ok(v, what):
    if v < 0:
        raise Hdf5Error(f"{what} failed")
    return v


def _dims(shape):
    return (hsize_t * len(shape))(*shape)


def _write_dataset(loc, name, array, h5type):
    lib = _lib()
    a = np.ascontiguousarray(array)
    space = _ok(lib.H5Screate_simple(a.ndim, _dims(a.shape), None), "H5Screate_simple")
    try:
        dset = _ok(lib.H5Dcreate2(loc, name.encode(), h5type, space, H5P_DEFAULT, H5P_DEFAULT,
                                  H5P_DEFAULT), f"H5Dcreate2 {name}")
        try:
            _ok(lib.H5Dwrite(dset, h5type, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                             a.ctypes.data_as(ctypes.c_void_p)), f"H5Dwrite {name}")
        finally:
            lib.H5Dclose(dset)
    finally:
        lib.H5Sclose(space)


def finest_level_mesh(dim, nb, cell_size):
    """(vertex_ids [cells][8] uint64, vertex_coordinates [vertices][3]) of a
    uniform block grid nb (blocks per direction), cells block by block in the
    block order (bx * nb[1] + by) * nb[2] + bz, k outer / i fastest inside."""
    nbx, nby = nb[0], nb[1]
    nbz = nb[2] if dim == 3 else 1
    icz = IC if dim == 3 else 1
    vx, vy, vz = nbx * IC + 1, nby * IC + 1, nbz * icz + 1
    k, j, i = np.meshgrid(np.arange(icz), np.arange(IC), np.arange(IC), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    ids = []
    for bx in range(nbx):
        for by in range(nby):
            for bz in range(nbz):
                x = bx * IC + i
                y = by * IC + j
                z = bz * icz + k
                corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                           (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
                ids.append(np.stack([(x + a) + (y + b) * vx + (z + c) * vx * vy
                                     for a, b, c in corners], axis=1))
    vertex_ids = np.concatenate(ids).astype(np.uint64)
    g = np.arange(vx * vy * vz, dtype=np.int64)
    coords = np.stack([(g % vx) * cell_size, ((g // vx) % vy) * cell_size,
                       (g // (vx * vy)) * cell_size], axis=1).astype(np.float64)
    return vertex_ids, coords


def _block_cell_order(field, dim, nb):
    """Domain array [NX][NY][NZ] -> cells block by block, k outer, i fastest."""
    nbx, nby = nb[0], nb[1]
    nbz = nb[2] if dim == 3 else 1
    icz = IC if dim == 3 else 1
    f = field.reshape(nbx, IC, nby, IC, nbz, icz)        # bx i by j bz k
    return f.transpose(0, 2, 4, 5, 3, 1).reshape(-1)     # bx by bz k j i


def write(path_without_extension, time, dim, nb, cell_size, density, velocity, pressure,
          series=None):
    """Write ``<path>.h5`` and ``<path>.xdmf`` for one output time, and append
    the time's grid to ``<series>.xdmf`` when a time-series file is given.
    density, pressure: [NX][NY][NZ]; velocity: [dim][NX][NY][NZ]."""
    vertex_ids, coords = finest_level_mesh(dim, nb, cell_size)
    ncells = vertex_ids.shape[0]
    rho = _block_cell_order(np.asarray(density, dtype=np.float64), dim, nb)
    p = _block_cell_order(np.asarray(pressure, dtype=np.float64), dim, nb)
    vel = np.zeros((ncells, 3, 1))
    for d in range(dim):
        vel[:, d, 0] = _block_cell_order(np.asarray(velocity[d], dtype=np.float64), dim, nb)
    return write_cells(path_without_extension, time, vertex_ids, coords, rho, vel, p, series)


def multilevel_mesh(dim, nb0, node_size, leaf_keys, current_maximum_level):
    """(vertex_ids, vertex_coordinates) of the leaves of a multi-level tree on
    the finest-level vertex lattice (StandardFinestLevelMeshGenerator::
    DoComputeVertexIDs, standard_finest_level_mesh_generator.cpp:214-300): a
    leaf (level, i, j, k) spans level_factor = 2^(current max - level)
    lattice steps per cell, its cells k outer and i fastest; lattice index n
    lies at n * dx_finest (2D: the cell of a level-l leaf is one of its own
    cell sizes thick)."""
    res = 1 << current_maximum_level
    dx_f = node_size / res / IC
    icz = IC if dim == 3 else 1
    vx = nb0[0] * res * IC + 1
    vy = nb0[1] * res * IC + 1
    vz = (nb0[2] * res * IC if dim == 3 else res) + 1
    k, j, i = np.meshgrid(np.arange(icz), np.arange(IC), np.arange(IC), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1),
               (0, 1, 1)]
    ids = []
    for level, bx, by, bz in leaf_keys:
        lf = res >> level
        x = (bx * IC + i) * lf
        y = (by * IC + j) * lf
        z = (bz * icz + k) * lf
        ids.append(np.stack([(x + a * lf) + (y + b * lf) * vx + (z + c * lf) * vx * vy
                             for a, b, c in corners], axis=1))
    vertex_ids = np.concatenate(ids).astype(np.uint64)
    g = np.arange(vx * vy * vz, dtype=np.int64)
    coords = np.stack([(g % vx) * dx_f, ((g // vx) % vy) * dx_f, (g // (vx * vy)) * dx_f],
                      axis=1).astype(np.float64)
    return vertex_ids, coords


def write_multilevel(path_without_extension, time, solver, series=None):
    """Output of a multi-level run (alpaca_amd.multilevel.MultiLevelSolver):
    every leaf's cells with density, velocity and pressure from its Average
    buffer (ConservativesToPrimeStates, prime_state_handler.h:279-370)."""
    keys, cons = solver.leaf_data()
    dim = solver.dim
    gamma, b = solver.gamma, solver.b
    vertex_ids, coords = multilevel_mesh(dim, solver.topo.nb0, solver.node_size, keys,
                                         solver.topo.current_maximum_level())
    # leaf blocks [n][NE][16][16][icz] -> cells per leaf, k outer and i fastest
    c = np.asarray(cons, dtype=np.float64).transpose(0, 1, 4, 3, 2).reshape(len(keys), dim + 2, -1)
    c = c.transpose(1, 0, 2).reshape(dim + 2, -1)
    rho = c[0]
    one_rho = 1.0 / rho
    vel = np.zeros((rho.size, 3, 1))
    for d in range(dim):
        vel[:, d, 0] = c[2 + d] * one_rho
    q = sum(c[2 + d] * c[2 + d] for d in range(dim))
    p = -gamma * b + (gamma - 1.0) * (c[1] - 0.5 * q / rho)
    return write_cells(path_without_extension, time, vertex_ids, coords, rho, vel, p, series)


def write_cells(path_without_extension, time, vertex_ids, coords, rho, vel, p, series=None):
    """The HDF5 file and its XDMF description from cell-ordered data:
    vertex_ids [cells][8], coords [vertices][3], rho and p [cells],
    vel [cells][3][1]."""
    lib = _lib()
    h5 = path_without_extension + ".h5"
    ncells = vertex_ids.shape[0]
    f64, u64 = _type("H5T_NATIVE_DOUBLE_g"), _type("H5T_NATIVE_ULLONG_g")
    fid = _ok(lib.H5Fcreate(h5.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    try:
        g = _ok(lib.H5Gcreate2(fid, b"metadata", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                "H5Gcreate2")
        s = _ok(lib.H5Screate(H5S_SCALAR), "H5Screate")
        a = _ok(lib.H5Acreate2(g, b"time", f64, s, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
        t = ctypes.c_double(time)
        _ok(lib.H5Awrite(a, f64, ctypes.byref(t)), "H5Awrite")
        lib.H5Aclose(a)
        lib.H5Sclose(s)
        lib.H5Gclose(g)
        g = _ok(lib.H5Gcreate2(fid, b"mesh_topology", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                "H5Gcreate2")
        _write_dataset(g, "cell_vertex_IDs", vertex_ids, u64)
        _write_dataset(g, "cell_vertex_coordinates", coords, f64)
        lib.H5Gclose(g)
        g = _ok(lib.H5Gcreate2(fid, b"cell_data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                "H5Gcreate2")
        _write_dataset(g, "density", np.ascontiguousarray(rho).reshape(ncells, 1, 1), f64)
        _write_dataset(g, "velocity", np.ascontiguousarray(vel), f64)
        _write_dataset(g, "pressure", np.ascontiguousarray(p).reshape(ncells, 1, 1), f64)
        lib.H5Gclose(g)
    finally:
        lib.H5Fclose(fid)
    block = spatial_xdmf(time, os.path.basename(h5), ncells, coords.shape[0])
    with open(path_without_extension + ".xdmf", "w") as fh:
        fh.write(xdmf_header("TimeStep") + block + xdmf_footer())
    if series is not None:
        with open(series + ".xdmf", "a") as fh:
            fh.write(block)
    return h5


def xdmf_header(name):
    """XdmfUtilities::HeaderInformation (xdmf_utilities.cpp:220-227)."""
    return ('<?xml version="1.0" ?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n<Xdmf Version="3.0">\n'
            f' <Domain>\n  <Grid Name="{name}" GridType="Collection" CollectionType="Temporal">\n')


def xdmf_footer():
    """XdmfUtilities::FooterInformation (xdmf_utilities.cpp:234-238)."""
    return "  </Grid>\n </Domain>\n</Xdmf>"


def spatial_xdmf(time, short, ncells, nv):
    """One time's grid (OutputWriter::XdmfSpatialDataInformation,
    output_writer.cpp:210-260): time, hexahedron topology, vertex geometry,
    density / velocity / pressure cell attributes in the HDF5 file `short`."""
    ind = lambda n: " " * n  # noqa: E731

    def item(name, dims):
        return (f'<DataItem Format="HDF" NumberType="Float" Precision="8" Dimensions="{ncells}'
                f'{dims}"> {short}:/cell_data/{name} </DataItem>\n')

    return (f'{ind(4)}<Grid Name="SpatialData_{time:.6e}" GridType="Uniform">\n'
            f'{ind(6)}<Time TimeType="Single" Value="{time:.6e}" />\n\n'
            f'{ind(6)}<Topology TopologyType="Hexahedron" NumberOfElements="{ncells}">\n'
            f'{ind(8)}<DataItem NumberType="Int" Format="HDF" Dimensions="{ncells} 8"> '
            f'{short}:/mesh_topology/cell_vertex_IDs </DataItem>\n{ind(6)}</Topology>\n'
            f'{ind(6)}<Geometry name="geometry" GeometryType="XYZ" NumberOfElements="{nv}">\n'
            f'{ind(8)}<DataItem Format="HDF" NumberType="Float" Precision="8" Dimensions="{nv} 3">'
            f' {short}:/mesh_topology/cell_vertex_coordinates </DataItem>\n{ind(6)}</Geometry>\n\n'
            f'{ind(6)}<Attribute Name="density" AttributeType="Scalar" Center="Cell">\n'
            f'{ind(8)}{item("density", "")}{ind(6)}</Attribute>\n'
            f'{ind(6)}<Attribute Name="velocity" AttributeType="Vector" Center="Cell">\n'
            f'{ind(8)}{item("velocity", " 3")}{ind(6)}</Attribute>\n'
            f'{ind(6)}<Attribute Name="pressure" AttributeType="Scalar" Center="Cell">\n'
            f'{ind(8)}{item("pressure", "")}{ind(6)}</Attribute>\n'
            f'{ind(4)}</Grid>\n')


def output_times(kind, interval, stamps, start, end):
    """ComputeOutputTimes (instantiation_input_output_manager.cpp:41-160):
    kind is off, interval, stamps or intervalstamps.  Interval times are
    start + k * interval (summed one interval at a time) below the end; stamps
    strictly inside (start, end); then the end time itself."""
    import math

    if kind == "off":
        return []
    times = []
    if kind in ("interval", "intervalstamps"):
        n = math.ceil(math.nextafter((end - start) / interval, 0.0))
        if n != 0:
            n -= 1
        t = start
        for _ in range(n):
            t += interval
            times.append(t)
    if kind in ("stamps", "intervalstamps"):
        times = sorted(times + [x for x in stamps if start < x < end])
    times.append(end)
    return times


class OutputSeries:
    """Standard output during a run (InputOutputManager::WriteFullOutput,
    input_output_manager.cpp:229-300): data_<t>.h5 / .xdmf in `folder` when
    the first pending output time is reached (all passed times are consumed)
    or when forced, each also appended to data_time_series.xdmf, which
    begin() opens and end() closes."""

    def __init__(self, folder, times, time_naming_factor=1.0, restored_time=None):
        self.folder = folder
        self.times = sorted(times)
        if restored_time is not None:     # RemoveTimeStamps(..., restart time)
            self.times = [t for t in self.times if t > restored_time]
        self.factor = time_naming_factor
        self.series = os.path.join(folder, "data_time_series")
        self.written = []

    def begin(self):
        os.makedirs(self.folder, exist_ok=True)
        with open(self.series + ".xdmf", "w") as fh:
            fh.write(xdmf_header("TimeSeries"))

    def maybe_write(self, t, write_file, force=False):
        """write_file(path_without_extension, series) writes one output."""
        if not (force or (self.times and self.times[0] <= t)):
            return None
        self.times = [x for x in self.times if x > t]
        path = os.path.join(self.folder, "data_%f" % (t * self.factor))
        write_file(path, self.series)
        self.written.append(path + ".h5")
        return path + ".h5"

    def end(self):
        with open(self.series + ".xdmf", "a") as fh:
            fh.write(xdmf_footer())


def read_dataset(h5, path):
    """Read a dataset of an HDF5 file (double or uint64) into numpy."""
    lib = _lib()
    fid = _ok(lib.H5Fopen(h5.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    try:
        d = _ok(lib.H5Dopen2(fid, path.encode(), H5P_DEFAULT), f"H5Dopen2 {path}")
        try:
            s = lib.H5Dget_space(d)
            nd = lib.H5Sget_simple_extent_ndims(s)
            dims = (hsize_t * nd)()
            lib.H5Sget_simple_extent_dims(s, dims, None)
            lib.H5Sclose(s)
            shape = tuple(int(x) for x in dims)
            ids = path.endswith("cell_vertex_IDs")
            out = np.empty(shape, dtype=np.uint64 if ids else np.float64)
            t = _type("H5T_NATIVE_ULLONG_g" if ids else "H5T_NATIVE_DOUBLE_g")
            _ok(lib.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                            out.ctypes.data_as(ctypes.c_void_p)), "H5Dread")
            return out
        finally:
            lib.H5Dclose(d)
    finally:
        lib.H5Fclose(fid)


def read_time(h5):
    lib = _lib()
    fid = _ok(lib.H5Fopen(h5.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    try:
        a = _ok(lib.H5Aopen_by_name(fid, b"metadata", b"time", H5P_DEFAULT, H5P_DEFAULT),
                "H5Aopen_by_name")
        t = ctypes.c_double()
        lib.H5Aread(a, _type("H5T_NATIVE_DOUBLE_g"), ctypes.byref(t))
        lib.H5Aclose(a)
        return t.value
    finally:
        lib.H5Fclose(fid)
