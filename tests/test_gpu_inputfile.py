"""The reference's own input files driven through the GPU path.

Each input (copies of testsuite/InputFiles/**, see golden/inputs/SOURCES.txt)
is read by alpaca_amd.inputfile, advanced to its end time on the GPU in FP64
and compared with the oracle advancing the same initial state with the same
sides: the same bits in every cell, the same step count and the same time
steps.  DetailedSod is also checked against the testsuite's level-0 error
norms (2D_DetailedSod_errors_reference.csv rows 2 and 20).

Sod's decimal data reach FP64 divisions whose quotient lies within ~2^-105
of a rounding midpoint by a divisor just below a power of two (0.35 / 0.125
with round-off); the exact form rounds those as x86 does, so these runs are
bit-exact too.
"""
import os

import numpy as np
import pytest

import exact_riemann
import oracle_py
from alpaca_amd import gpu, inputfile

pytestmark = pytest.mark.gpu

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _oracle(case, level, flux):
    desc = inputfile.case_desc(case, level, flux)
    st0 = inputfile.initial_state(case, level)
    nb = [desc.nb[d] for d in range(case.dim)]
    return oracle_py.run(case.dim, nb, case.bc[:case.dim], flux, case.cfl, case.end_time,
                         100000, st0, gamma=case.gamma, b=case.background_pressure,
                         cell_size=desc.cell_size)


@pytest.mark.parametrize("name,dim,flux", [
    ("DetailedSod_2D_SodX.xml", 2, gpu.FLUX_ROE), ("DetailedSod_2D_SodX.xml", 2, gpu.FLUX_HLLC),
    ("SinglePhase_2D_ToroImplosion.xml", 2, gpu.FLUX_HLLC),
    ("Symmetry_3D_FullImplosion.xml", 3, gpu.FLUX_HLLC),
    ("Parallelization_3D_GaussPuls.xml", 3, gpu.FLUX_ROE)])
def test_input_file_runs_match_oracle(torch, name, dim, flux):
    print(f"[input file] {name} dim={dim} flux={flux}", flush=True)
    case = inputfile.read_input(os.path.join(INPUTS, name), dim=dim)
    res = inputfile.run(case, 0, flux)
    want, want_dts, tf = _oracle(case, 0, flux)
    got = res["conservatives"]
    scale = np.maximum(np.abs(want).reshape(want.shape[0], -1).max(axis=1), 1e-300)
    err = float((np.abs(got - want).reshape(want.shape[0], -1).max(axis=1) / scale).max())
    print(f"  steps={res['steps']} t={res['time']} max rel err {err:.3e} "
          f"bit-exact {np.array_equal(got, want)}", flush=True)
    assert res["steps"] == len(want_dts)
    assert np.array_equal(np.asarray(res["dts"]), np.asarray(want_dts))
    assert res["time"] == tf and tf == case.end_time
    assert np.array_equal(got, want), f"max rel err {err:.3e}"
    if name.startswith("DetailedSod"):
        rho_l1, p_l1 = {gpu.FLUX_ROE: (1.00420788e-04, 1.13784806e-04),
                        gpu.FLUX_HLLC: (1.02588746e-04, 1.12368183e-04)}[flux]
        x = (np.arange(64) + 0.5) / 64
        ex = exact_riemann.sample((x - 0.5) / tf, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
        vol = (1.0 / 64) ** 3
        got_rho = np.sum(np.abs(res["density"][:, :, 0] / ex[0][:, None] - 1.0) * vol)
        got_p = np.sum(np.abs(res["pressure"][:, :, 0] / ex[2][:, None] - 1.0) * vol)
        print(f"  DensityL1 {got_rho:.8e} PressureL1 {got_p:.8e}", flush=True)
        assert abs(got_rho / rho_l1 - 1.0) < 1e-6
        assert abs(got_p / p_l1 - 1.0) < 1e-6


def test_input_file_command_line(torch, tmp_path):
    out = tmp_path / "sod.npz"
    rc = inputfile.main([os.path.join(INPUTS, "DetailedSod_2D_SodY.xml"), "--flux", "roe",
                         "--out", str(out), "--output-folder", str(tmp_path / "run1")])
    assert rc == 0
    data = np.load(out)
    assert float(data["time"]) == 0.2 and data["density"].shape == (16, 64, 1)
    from alpaca_amd import h5output

    try:
        h5output._lib()
    except h5output.Hdf5Error:
        return
    rc = inputfile.main([os.path.join(INPUTS, "DetailedSod_2D_SodY.xml"), "--flux", "roe",
                         "--out", str(tmp_path / "sod.h5"), "--output-folder",
                         str(tmp_path / "run2")])
    assert rc == 0
    rho = h5output.read_dataset(str(tmp_path / "sod.h5"), "/cell_data/density")
    assert rho.shape == (16 * 64, 1, 1)
    assert np.array_equal(np.sort(rho.ravel()), np.sort(data["density"].ravel()))


def test_run_time_output_in_reference_layout(torch, tmp_path):
    """SodX asks for standard output every 0.05: the run writes the initial
    state, one file after each output time is passed and the end state, each
    in the reference's HDF5 + XDMF layout, plus the time-series collection."""
    import xml.etree.ElementTree as ET

    from alpaca_amd import h5output

    try:
        h5output._lib()
    except h5output.Hdf5Error:
        pytest.skip("libhdf5 not available")
    case = inputfile.read_input(os.path.join(INPUTS, "DetailedSod_2D_SodX.xml"))
    assert case.output_times == [0.05, 0.1, 0.05 + 0.05 + 0.05, 0.2]
    res = inputfile.run(case, 0, gpu.FLUX_ROE, output_folder=str(tmp_path / "domain"))
    files = [os.path.basename(f) for f in res["output_files"]]
    print(f"[run-time output] {files}", flush=True)
    assert len(files) == 5 and files[0] == "data_0.000000.h5" and files[-1] == "data_0.200000.h5"
    times = [h5output.read_time(f) for f in res["output_files"]]
    assert times[0] == 0.0 and times[-1] == 0.2 and times == sorted(times)
    for t, want in zip(times[1:4], (0.05, 0.1, 0.15)):
        assert want <= t < want + 0.01
    rho = h5output.read_dataset(res["output_files"][-1], "/cell_data/density")[:, 0, 0]
    assert np.array_equal(np.sort(rho), np.sort(res["density"].ravel()))
    root = ET.parse(str(tmp_path / "domain" / "data_time_series.xdmf")).getroot()
    assert len(root.findall("Domain/Grid/Grid")) == 5
