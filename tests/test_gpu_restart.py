"""Restart snapshots from the GPU path, and runs continued from them.

An input file asks for snapshots at time stamps; the run writes them in the
reference's restart layout; a second run restores from a snapshot (as the
reference's <restore><mode> Forced does) and finishes.  The restored run must
be bit-identical to the uninterrupted one, with the same time steps, and the
snapshot's interior must equal the oracle advanced to the same step count to
the FP64 bound (max relative error <= 1e-10).
"""
import json
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle_py
from alpaca_amd import gpu, h5output, inputfile, restart

pytestmark = pytest.mark.gpu

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs")


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    try:
        h5output._lib()
    except h5output.Hdf5Error:
        pytest.skip("libhdf5 not available")
    return _t


def _input_with_stamps(tmp_path, name, stamps):
    tree = ET.parse(os.path.join(INPUTS, name))
    snap = tree.getroot().find("restart/snapshots")
    for c in list(snap):
        snap.remove(c)
    ET.SubElement(snap, "type").text = "Stamps"
    st = ET.SubElement(snap, "stamps")
    for k, t in enumerate(stamps):
        ET.SubElement(st, f"ts{k + 1}").text = repr(t)
    path = str(tmp_path / name)
    tree.write(path)
    return path


@pytest.mark.parametrize("name,dim,flux,stamp", [
    ("DetailedSod_2D_SodX.xml", 2, gpu.FLUX_HLLC, 0.1),
    ("Parallelization_3D_GaussPuls.xml", 3, gpu.FLUX_ROE, None)])
def test_restart_continuation_is_exact(torch, tmp_path, name, dim, flux, stamp):
    print(f"[restart] {name} dim={dim} flux={flux}", flush=True)
    case0 = inputfile.read_input(os.path.join(INPUTS, name), dim=dim)
    if stamp is None:
        stamp = 0.5 * case0.end_time
    path = _input_with_stamps(tmp_path, name, [0.0, stamp])
    case = inputfile.read_input(path, dim=dim)
    folder = str(tmp_path / "restart")
    full = inputfile.run(case, 0, flux, restart_folder=folder)
    files = full["restart_files"]
    names = [os.path.basename(f) for f in files]
    print(f"  {full['steps']} steps, snapshots {names}", flush=True)
    # stamps 0, `stamp` and the end time; each written after the step that passes it
    assert len(files) == 3 and names[0] == "restart_0.000000.h5"
    assert names[2] == restart.snapshot_name(full["time"])
    assert os.readlink(os.path.join(folder, "latest_restart_snapshot")) == names[2]
    t_snap, fdim, fnb, blocks = restart.read(files[1], dim)
    assert fdim == dim and t_snap >= stamp
    k = int(np.searchsorted(np.cumsum(full["dts"]), t_snap * (1 - 1e-14)))
    assert abs(np.sum(full["dts"][:k + 1]) - t_snap) <= 1e-12 * t_snap
    nsnap = k + 1
    # the end-time snapshot holds the final state
    _, _, _, last = restart.read(files[2], dim)
    assert np.array_equal(restart.domain_from_blocks(last, dim, fnb), full["conservatives"])
    # the snapshot equals the oracle after the same number of steps
    desc = inputfile.case_desc(case, 0, flux)
    want, want_dts, _ = oracle_py.run(dim, fnb, case.bc[:dim], flux, case.cfl, case.end_time,
                                      nsnap, inputfile.initial_state(case, 0),
                                      gamma=case.gamma, b=case.background_pressure,
                                      cell_size=desc.cell_size)
    got = restart.domain_from_blocks(blocks, dim, fnb)
    scale = np.maximum(np.abs(want).reshape(want.shape[0], -1).max(axis=1), 1e-300)
    err = float((np.abs(got - want).reshape(want.shape[0], -1).max(axis=1) / scale).max())
    print(f"  snapshot at step {nsnap}, t={t_snap}: max rel err vs oracle {err:.3e}", flush=True)
    assert err <= 1e-10
    # restore from the middle snapshot and finish: identical to the uninterrupted run
    res = inputfile.run(case, 0, flux, restore=files[1])
    assert res["start_time"] == t_snap
    assert res["steps"] == full["steps"] - nsnap
    assert np.array_equal(res["dts"], full["dts"][nsnap:])
    assert res["time"] == full["time"]
    assert np.array_equal(res["conservatives"], full["conservatives"])


def test_restore_from_input_file_and_command_line(torch, tmp_path):
    name = "DetailedSod_2D_SodY.xml"
    path = _input_with_stamps(tmp_path, name, [0.05])
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        rc = inputfile.main([path, "--flux", "roe", "--out", str(tmp_path / "a.npz")])
        assert rc == 0
        snap = os.path.join("DetailedSod_2D_SodY", "restart")
        files = sorted(f for f in os.listdir(snap) if f.endswith(".h5"))
        assert len(files) == 2, files
        mid = [f for f in files if f != restart.snapshot_name(0.2)][0]
        # the input file's <restore> section names the snapshot (mode Forced)
        tree = ET.parse(path)
        tree.getroot().find("restart/restore/mode").text = "Forced"
        tree.getroot().find("restart/restore/fileName").text = os.path.join(snap, mid)
        tree.write(path)
        rc = inputfile.main([path, "--flux", "roe", "--output-folder", str(tmp_path / "r2"),
                             "--out", str(tmp_path / "b.npz")])
        assert rc == 0
        a, b = np.load(tmp_path / "a.npz"), np.load(tmp_path / "b.npz")
        assert float(b["start_time"]) > 0.05 and float(b["time"]) == 0.2
        assert np.array_equal(a["conservatives"], b["conservatives"])
    finally:
        os.chdir(cwd)
    print(json.dumps({"snapshots": files}), flush=True)
