"""HDF5 + XDMF output in the reference's layout (CPU).

Pinned by the reference's own consumer: the file is read back the way
python/alpacapy/post_analysis/sod_analysis.py:206-238 reads ALPACA output
(cell centres as the mean of the 8 vertices by cell_vertex_IDs, cell volume
from the vertex extent, density/velocity/pressure columns, metadata time),
and the DetailedSod level-0 norms computed from that reading equal the
testsuite's reference values (2D_DetailedSod_errors_reference.csv rows 2, 20).
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import exact_riemann
import oracle_py
from alpaca_amd import gpu, h5output, inputfile

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")

try:
    h5output._lib()
    HAVE_HDF5 = True
except h5output.Hdf5Error:
    HAVE_HDF5 = False

pytestmark = pytest.mark.skipif(not HAVE_HDF5, reason="libhdf5 not available")


def _read_like_sod_analysis(h5):
    """sod_analysis.py:219-238, with the HDF5 reads done through the C library."""
    cell_vertices = h5output.read_dataset(h5, "/mesh_topology/cell_vertex_IDs").astype(np.int64)
    vertex_coordinates = h5output.read_dataset(h5, "/mesh_topology/cell_vertex_coordinates")
    density = h5output.read_dataset(h5, "/cell_data/density")[:, 0, 0]
    velocity = h5output.read_dataset(h5, "/cell_data/velocity")[:, :, 0]
    pressure = h5output.read_dataset(h5, "/cell_data/pressure")[:, 0, 0]
    time = np.float64(h5output.read_time(h5))
    ordered = vertex_coordinates[cell_vertices]
    centers = np.mean(ordered, axis=1)
    longest = np.argmax(np.max(centers, axis=0))
    delta = np.max(ordered, axis=1) - np.min(ordered, axis=1)
    volume = np.prod(delta, axis=1)
    return centers[:, longest], density, velocity[:, longest], pressure, volume, time


@pytest.mark.parametrize("flux,rho_l1,p_l1", [(gpu.FLUX_ROE, 1.00420788e-04, 1.13784806e-04),
                                              (gpu.FLUX_HLLC, 1.02588746e-04, 1.12368183e-04)])
def test_sod_output_read_by_reference_analysis(tmp_path, flux, rho_l1, p_l1):
    case = inputfile.read_input(os.path.join(INPUTS, "DetailedSod_2D_SodX.xml"))
    desc = inputfile.case_desc(case, 0, flux)
    st0 = inputfile.initial_state(case, 0)
    st, _, tf = oracle_py.run(2, [4, 1], case.bc[:2], flux, case.cfl, case.end_time, 100000, st0,
                              cell_size=desc.cell_size)
    pr = inputfile.primes(case, st)
    h5 = h5output.write(str(tmp_path / "sod"), tf, 2, [4, 1, 1], desc.cell_size, pr["density"],
                        pr["velocity"], pr["pressure"])
    x, rho, u, p, vol, t = _read_like_sod_analysis(h5)
    assert t == tf
    assert np.allclose(vol, (1.0 / 64) ** 3, rtol=1e-14)
    ex = exact_riemann.sample((x - 0.5) / t, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    # sod_analysis compares |q / q_exact - 1| summed with the cell volume
    got_rho = np.sum(np.abs(rho / ex[0] - 1.0) * vol)
    got_p = np.sum(np.abs(p / ex[2] - 1.0) * vol)
    assert abs(got_rho / rho_l1 - 1.0) < 1e-6
    assert abs(got_p / p_l1 - 1.0) < 1e-6
    # the velocity column is the x velocity of every cell
    assert np.allclose(np.sort(u), np.sort(pr["velocity"][0].ravel()), rtol=0, atol=0)


def test_3d_layout_and_xdmf(tmp_path):
    dim, nb, cs = 3, [2, 1, 3], 0.125
    rng = np.random.default_rng(3)
    shape = (nb[0] * 16, nb[1] * 16, nb[2] * 16)
    rho = rng.uniform(1, 2, shape)
    vel = rng.uniform(-1, 1, (3,) + shape)
    p = rng.uniform(1, 2, shape)
    h5 = h5output.write(str(tmp_path / "out"), 0.25, dim, nb, cs, rho, vel, p)
    ids = h5output.read_dataset(h5, "/mesh_topology/cell_vertex_IDs").astype(np.int64)
    xyz = h5output.read_dataset(h5, "/mesh_topology/cell_vertex_coordinates")
    n = np.prod(shape)
    assert ids.shape == (n, 8) and xyz.shape == ((shape[0] + 1) * (shape[1] + 1) * (shape[2] + 1), 3)
    centers = xyz[ids].mean(axis=1)
    idx = np.rint(centers / cs - 0.5).astype(np.int64)
    got = h5output.read_dataset(h5, "/cell_data/density")[:, 0, 0]
    assert np.array_equal(got, rho[idx[:, 0], idx[:, 1], idx[:, 2]])
    gv = h5output.read_dataset(h5, "/cell_data/velocity")
    assert gv.shape == (n, 3, 1)
    for d in range(3):
        assert np.array_equal(gv[:, d, 0], vel[d][idx[:, 0], idx[:, 1], idx[:, 2]])
    # every cell is a unit hexahedron: vertex 0 at the minimum corner, 6 at the maximum
    assert np.allclose(xyz[ids[:, 6]] - xyz[ids[:, 0]], cs)
    # first block, first row: i fastest (standard_finest_level_mesh_generator.cpp:145-157)
    assert idx[1].tolist() == [1, 0, 0] and idx[16].tolist() == [0, 1, 0]
    assert h5output.read_time(h5) == 0.25
    root = ET.parse(str(tmp_path / "out.xdmf")).getroot()
    grid = root.find("Domain/Grid/Grid")
    assert grid.find("Topology").get("NumberOfElements") == str(n)
    names = [a.get("Name") for a in grid.findall("Attribute")]
    assert names == ["density", "velocity", "pressure"]
    item = grid.find("Attribute/DataItem").text.strip()
    assert item == "out.h5:/cell_data/density"


def test_output_times_follow_compute_output_times():
    """ComputeOutputTimes (instantiation_input_output_manager.cpp:41-160):
    interval times summed one interval at a time and kept below the end,
    stamps strictly inside (start, end), the end time always last."""
    ot = h5output.output_times
    assert ot("off", 0.05, [], 0.0, 0.2) == []
    assert ot("interval", 0.05, [], 0.0, 0.2) == [0.05, 0.1, 0.05 + 0.05 + 0.05, 0.2]
    assert ot("interval", 0.3, [], 0.0, 1.0) == [0.3, 0.6, 0.3 + 0.3 + 0.3, 1.0]
    assert ot("interval", 5.0, [], 0.0, 1.0) == [1.0]
    assert ot("stamps", 0.0, [0.5, 0.0, 1.0, 0.25, 2.0], 0.0, 1.0) == [0.25, 0.5, 1.0]
    assert ot("intervalstamps", 0.4, [0.1], 0.0, 1.0) == [0.1, 0.4, 0.8, 1.0]


def test_output_series_schedule_and_time_series_file(tmp_path):
    rng = np.random.default_rng(1)
    dim, nb, cs = 2, [2, 1, 1], 1.0 / 32
    rho = rng.uniform(1, 2, (32, 16, 1))
    vel = rng.uniform(-1, 1, (2, 32, 16, 1))
    p = rng.uniform(1, 2, (32, 16, 1))
    series = h5output.OutputSeries(str(tmp_path / "domain"), [0.05, 0.1, 0.2])
    series.begin()

    def writer(t):
        return lambda path, ts: h5output.write(path, t, dim, nb, cs, rho, vel, p, series=ts)

    for t, force in ((0.0, True), (0.03, False), (0.12, False), (0.2, False), (0.21, False)):
        series.maybe_write(t, writer(t), force)
    series.end()
    names = [os.path.basename(f) for f in series.written]
    # forced at the start; 0.12 consumes 0.05 and 0.1; then the end time
    assert names == ["data_0.000000.h5", "data_0.120000.h5", "data_0.200000.h5"]
    root = ET.parse(str(tmp_path / "domain" / "data_time_series.xdmf")).getroot()
    grids = root.findall("Domain/Grid/Grid")
    assert [g.find("Time").get("Value") for g in grids] == ["0.000000e+00", "1.200000e-01",
                                                           "2.000000e-01"]
    assert grids[1].find("Attribute/DataItem").text.strip() == "data_0.120000.h5:/cell_data/density"
    # a restart at 0.1 drops the times not after it
    assert h5output.OutputSeries("x", [0.05, 0.1, 0.2], restored_time=0.1).times == [0.2]
