"""Restart files in the reference's layout and the snapshot schedule (CPU).

The node-id encoding is pinned by the reference's own TopologyManager loop
and id arithmetic (golden/level0_ids.json, written by make_golden.py --ids
from oracle/_ref, and live against oracle/_ref where that is built).  The
file layout follows restart_manager.cpp:300-480; a file written here is read
back with the HDF5 C library dataset by dataset (names, shapes, types and
attributes the reference's RestoreSimulation opens), and a run continued from
a snapshot is bit-identical to the uninterrupted run (oracle trajectory).
"""
import ctypes
import hashlib
import json
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle_py
from alpaca_amd import cases, gpu, h5output, inputfile, restart

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
INPUTS = os.path.join(GOLDEN, "inputs")

try:
    h5output._lib()
    HAVE_HDF5 = True
except h5output.Hdf5Error:
    HAVE_HDF5 = False

needs_hdf5 = pytest.mark.skipif(not HAVE_HDF5, reason="libhdf5 not available")


def _ids(nb):
    return np.array([i for _, i in restart.creation_order(nb)], dtype=np.uint64)


def test_level0_ids_match_reference_fixture():
    meta = json.load(open(os.path.join(GOLDEN, "level0_ids.json")))
    for key, ids in meta["full"].items():
        nb = [int(x) for x in key.split(",")]
        assert _ids(nb).tolist() == ids, key
    for key, digest in meta["sha256"].items():
        nb = [int(x) for x in key.split(",")]
        assert hashlib.sha256(_ids(nb).tobytes()).hexdigest() == digest, key


@pytest.mark.skipif(not oracle_py.ref_available(3), reason="oracle/_ref not built")
def test_level0_ids_match_reference_live():
    rng = np.random.default_rng(11)
    for _ in range(12):
        nb = [int(x) for x in rng.integers(1, 12, 3)]
        assert np.array_equal(_ids(nb), oracle_py.ref_level0_ids(nb)), nb


def test_id_decode_inverts_encode():
    rng = np.random.default_rng(5)
    for b in rng.integers(0, restart.MAX_BLOCKS_PER_DIRECTION, (200, 3)):
        b = tuple(int(x) for x in b)
        assert restart.decode_level0_id(restart.level0_id(*b)) == b
    assert restart.level0_id(0, 0, 0) == restart.ID_SEED
    # a level-1 child id (parent id shifted by three bits) is not a level-0 id
    with pytest.raises(restart.RestartError):
        restart.decode_level0_id(restart.ID_SEED << 3)
    with pytest.raises(restart.RestartError):
        restart.level0_id(restart.MAX_BLOCKS_PER_DIRECTION, 0, 0)


def test_blocks_domain_round_trip():
    rng = np.random.default_rng(2)
    for dim, nb in ((2, [3, 2]), (3, [2, 1, 3])):
        shape = (dim + 2, nb[0] * 16, nb[1] * 16, nb[2] * 16 if dim == 3 else 1)
        a = rng.standard_normal(shape)
        b = restart.blocks_from_domain(a, dim, nb)
        assert b.shape == (np.prod(nb), dim + 2, 24, 24, 24 if dim == 3 else 1)
        assert np.array_equal(restart.domain_from_blocks(b, dim, nb), a)
        # block (bx, by, bz) sits at index (bx * ny + by) * nz + bz, interior at [4, 20)
        bx, by = nb[0] - 1, nb[1] - 1
        k = (bx * nb[1] + by) * (nb[2] if dim == 3 else 1)
        assert b[k, 1, 4, 4, 4 if dim == 3 else 0] == a[1, bx * 16, by * 16, 0]


def _state(dim, nb, seed=0):
    rng = np.random.default_rng(seed)
    n = [nb[d] * 16 for d in range(dim)] + ([1] if dim == 2 else [])
    rho = rng.uniform(0.5, 2.0, n)
    p = rng.uniform(0.5, 2.0, n)
    vel = [rng.uniform(-1, 1, n) for _ in range(dim)]
    return cases.primes_to_conservatives(dim, 1.4, 0.0, rho, vel, p)


def _raw(path, name, dtype, h5name):
    lib = h5output._lib()
    fid = lib.H5Fopen(path.encode(), h5output.H5F_ACC_RDONLY, h5output.H5P_DEFAULT)
    try:
        return restart._read(fid, name, dtype, h5name)
    finally:
        lib.H5Fclose(fid)


def _attr(path, name, h5name, ctype):
    lib = h5output._lib()
    fid = lib.H5Fopen(path.encode(), h5output.H5F_ACC_RDONLY, h5output.H5P_DEFAULT)
    try:
        return restart._read_attr(fid, name, h5name, ctype)
    finally:
        lib.H5Fclose(fid)


@needs_hdf5
@pytest.mark.parametrize("dim,nb", [(2, [3, 2]), (3, [2, 2, 3])])
def test_write_layout_and_round_trip(tmp_path, dim, nb):
    cons = _state(dim, nb)
    blocks = restart.blocks_from_domain(cons, dim, nb)
    blocks[:, :, :4] = 7.0                  # halo content travels as it is
    path = restart.write(str(tmp_path / "r.h5"), 0.125, dim, nb, blocks, 1.4)
    nblk = int(np.prod(nb))
    tz = 24 if dim == 3 else 1
    ids = _raw(path, "/node_info_data/NodeIds", np.uint64, "H5T_NATIVE_ULLONG_g")
    assert np.array_equal(ids, _ids(nb))
    for name in ("NumberOfMaterials", "Materials", "NumberOfInterfaceBlocks"):
        v = _raw(path, f"/node_info_data/{name}", np.uint16, "H5T_NATIVE_USHORT_g")
        assert v.shape == (nblk,)
        assert np.all(v == (1 if name == "NumberOfMaterials" else 0))
    c = _raw(path, "/node_cell_data/Conservatives", np.float64, "H5T_NATIVE_DOUBLE_g")
    assert c.shape == (nblk, dim + 2, 24, 24, tz)
    order = [k for k, _ in restart.creation_order(nb + [1] * (3 - len(nb)))]
    assert np.array_equal(c, blocks[order])
    pr = _raw(path, "/node_cell_data/PrimeStates", np.float64, "H5T_NATIVE_DOUBLE_g")
    assert pr.shape == (nblk, dim + 2, 24, 24, tz)
    # PrimeStates are density, pressure, velocities (field_details.h:201-212)
    # on every cell that holds a state; halo cells left empty hold zeros
    full = c[:, 0] != 0.0
    assert np.all(np.moveaxis(pr, 1, 0)[:, ~full] == 0.0)
    cf = np.moveaxis(c, 1, 0)[:, full]
    rho, vel, p = cases.conservatives_to_primes(dim, 1.4, 0.0, cf)
    pf = np.moveaxis(pr, 1, 0)[:, full]
    assert np.array_equal(pf[0], rho) and np.array_equal(pf[1], p)
    for d in range(dim):
        assert np.array_equal(pf[2 + d], vel[d])
    inner = (slice(None), slice(4, 20), slice(4, 20), slice(4, 20) if dim == 3 else slice(0, 1))
    x = np.moveaxis(pr, 1, 0)[(slice(None),) + inner]
    back = cases.primes_to_conservatives(dim, 1.4, 0.0, x[0], [x[2 + d] for d in range(dim)],
                                         x[1])
    assert np.allclose(back, np.moveaxis(c, 1, 0)[(slice(None),) + inner], rtol=1e-13)
    assert _raw(path, "/node_cell_data/Levelset", np.float64, "H5T_NATIVE_DOUBLE_g").shape == \
        (0, 24, 24, tz)
    assert _raw(path, "/node_cell_data/InterfaceTags", np.int8, "H5T_NATIVE_SCHAR_g").shape == \
        (0, 24, 24, tz)
    u32, f64 = "H5T_NATIVE_UINT_g", "H5T_NATIVE_DOUBLE_g"
    assert _attr(path, "Time", f64, ctypes.c_double) == 0.125
    assert _attr(path, "Dimensions", u32, ctypes.c_uint) == dim
    assert _attr(path, "InternalCells", u32, ctypes.c_uint) == 16
    assert _attr(path, "HaloSize", u32, ctypes.c_uint) == 4
    assert _attr(path, "MaximumLevel", u32, ctypes.c_uint) == 0
    for name in ("LengthReference", "VelocityReference", "DensityReference",
                 "TemperatureReference"):
        assert _attr(path, name, f64, ctypes.c_double) == 1.0
    t, fdim, fnb, got = restart.read(path, dim)
    assert (t, fdim, fnb) == (0.125, dim, nb)
    assert np.array_equal(got, blocks)


@needs_hdf5
def test_read_places_blocks_by_id(tmp_path):
    """The reference writes nodes in hash-map order; reading must not depend
    on the order of the file."""
    dim, nb = 3, [3, 2, 2]
    blocks = restart.blocks_from_domain(_state(dim, nb, 4), dim, nb)
    order = restart.creation_order(nb)
    perm = np.random.default_rng(9).permutation(len(order))
    path = restart.write(str(tmp_path / "p.h5"), 1.5, dim, nb, blocks, 1.4,
                         order=[order[i] for i in perm])
    ids = _raw(path, "/node_info_data/NodeIds", np.uint64, "H5T_NATIVE_ULLONG_g")
    assert not np.array_equal(ids, _ids(nb))
    t, _, fnb, got = restart.read(path)
    assert fnb == nb and np.array_equal(got, blocks)


@needs_hdf5
def test_read_refuses_what_cannot_continue(tmp_path):
    dim, nb = 2, [2, 2]
    blocks = restart.blocks_from_domain(_state(dim, nb), dim, nb)
    path = restart.write(str(tmp_path / "a.h5"), 0.0, dim, nb, blocks, 1.4)
    with pytest.raises(restart.RestartError):
        restart.read(path, 3)
    order = restart.creation_order([2, 2, 1])[:3]     # one block missing
    path = restart.write(str(tmp_path / "b.h5"), 0.0, dim, nb, blocks, 1.4, order=order)
    with pytest.raises(restart.RestartError):
        restart.read(path)


@needs_hdf5
@pytest.mark.parametrize("dim,nb,bc,flux", [(2, [2, 2], [1, 0], gpu.FLUX_HLLC),
                                            (3, [1, 2, 1], [0, 1, 2], gpu.FLUX_ROE)])
def test_continuation_from_snapshot_is_exact(tmp_path, dim, nb, bc, flux):
    """N steps, snapshot, read, M steps == N + M steps, bit for bit (oracle)."""
    st0 = _state(dim, nb, 8)
    n, m = 3, 4
    full, full_dts, _ = oracle_py.run(dim, nb, bc, flux, 0.5, 1e9, n + m, st0)
    part, dts1, t1 = oracle_py.run(dim, nb, bc, flux, 0.5, 1e9, n, st0)
    path = restart.write(str(tmp_path / "s.h5"), t1, dim, nb,
                         restart.blocks_from_domain(part, dim, nb), 1.4)
    t, _, fnb, blocks = restart.read(path, dim)
    assert t == t1
    rest, dts2, _ = oracle_py.run(dim, nb, bc, flux, 0.5, 1e9, m,
                                  restart.domain_from_blocks(blocks, dim, fnb))
    assert np.array_equal(np.concatenate([dts1, dts2]), full_dts)
    assert np.array_equal(rest, full)


def test_snapshot_names_follow_to_string():
    assert restart.snapshot_name(0.0) == "restart_0.000000.h5"
    assert restart.snapshot_name(0.1000000123) == "restart_0.100000.h5"
    assert restart.snapshot_name(2.5, 1000.0) == "restart_2500.000000.h5"


def test_snapshot_schedule(tmp_path, monkeypatch):
    clock = [100.0]
    monkeypatch.setattr(restart._time, "time", lambda: clock[0])
    written = []

    def writer(path):
        written.append(os.path.basename(path))
        open(path, "w").close()

    s = restart.SnapshotSchedule(str(tmp_path), stamps=[0.0, 0.1, 0.3, 0.5], interval_s=10,
                                 keep=2)
    assert s.maybe_write(0.0, writer)                       # stamp 0 at initialisation
    assert s.maybe_write(0.05, writer) is None
    assert s.maybe_write(0.31, writer)                      # consumes 0.1 and 0.3
    assert s.stamps == [0.5]
    for k in range(3):                                      # three interval snapshots
        clock[0] += 10
        assert s.maybe_write(0.32 + 0.01 * k, writer)
    assert s.maybe_write(0.4, writer, force=True)
    assert written == ["restart_0.000000.h5", "restart_0.310000.h5", "restart_0.320000.h5",
                       "restart_0.330000.h5", "restart_0.340000.h5", "restart_0.400000.h5"]
    left = sorted(f for f in os.listdir(tmp_path) if f.endswith(".h5"))
    # stamp snapshots stay; interval/forced ones beyond two are removed oldest first
    assert left == ["restart_0.000000.h5", "restart_0.310000.h5", "restart_0.340000.h5",
                    "restart_0.400000.h5"]
    assert os.readlink(tmp_path / "latest_restart_snapshot") == "restart_0.400000.h5"
    r = restart.SnapshotSchedule(str(tmp_path), stamps=[0.1, 0.3, 0.5], restored_time=0.3)
    assert r.stamps == [0.5]


def _with_restart_section(tmp_path, mode="Off", filename="", stamps=None, interval=None):
    tree = ET.parse(os.path.join(INPUTS, "DetailedSod_2D_SodX.xml"))
    root = tree.getroot()
    rs = root.find("restart")
    rs.find("restore/mode").text = mode
    rs.find("restore/fileName").text = filename
    snap = rs.find("snapshots")
    for c in list(snap):
        snap.remove(c)
    if stamps is not None:
        ET.SubElement(snap, "type").text = "Stamps"
        st = ET.SubElement(snap, "stamps")
        for k, t in enumerate(stamps):
            ET.SubElement(st, f"ts{k + 1}").text = str(t)
    elif interval is not None:
        ET.SubElement(snap, "type").text = "Interval"
        ET.SubElement(snap, "interval").text = str(interval)
        ET.SubElement(snap, "intervalsToKeep").text = "3"
    else:
        ET.SubElement(snap, "type").text = "Off"
    path = str(tmp_path / "case.xml")
    tree.write(path)
    return path


def test_input_restart_section(tmp_path):
    case = inputfile.read_input(os.path.join(INPUTS, "DetailedSod_2D_SodX.xml"))
    assert case.restore_mode == "off" and case.snapshot_stamps == []
    assert case.snapshot_interval == 3600 and case.snapshots_to_keep == 2
    p = _with_restart_section(tmp_path, "Soft", "x.h5", stamps=[0.05, 0.1, 0.5])
    case = inputfile.read_input(p)
    assert case.restore_mode == "soft" and case.restore_file == "x.h5"
    # stamps beyond the end time drop out and the end time is added
    assert case.snapshot_stamps == [0.05, 0.1, 0.2]
    assert inputfile.restore_path(case) is None              # soft: missing file, fresh start
    p = _with_restart_section(tmp_path, "Forced", str(tmp_path / "missing.h5"))
    with pytest.raises(inputfile.InputError):
        inputfile.restore_path(inputfile.read_input(p))
    p = _with_restart_section(tmp_path, "Off", "", stamps=[0.1, 0.05])
    with pytest.raises(inputfile.InputError):
        inputfile.read_input(p)
    p = _with_restart_section(tmp_path, "Sometimes", "x.h5")
    with pytest.raises(inputfile.InputError):
        inputfile.read_input(p)
