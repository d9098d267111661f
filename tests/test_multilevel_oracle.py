"""The multi-level driver (multiresolution + local time stepping) on the CPU
backend (the C restatement, block by block): pinned by the reference
testsuite's DetailedSod norms at maximum level 2 (every remeshing decision,
jump flux and level-jump halo enters them), by conservation under periodic
sides (a wrong jump-flux correction breaks it) and by the topology rules."""
import numpy as np
import pytest

from alpaca_amd import gpu, lts, multilevel
import ml_cases
import oracle_ml


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_detailed_sod_level2_norms(flux):
    """2D SodX at maximumLevel 2 (finest 256x64 cells): density and pressure
    L1 norms equal 2D_DetailedSod_errors_reference.csv rows 3 (Roe) and 21
    (HLLC) to 1e-6 relative (the suite's gate is 3%)."""
    be = oracle_ml.OracleBackend(2, ml_cases.capacity(2, 4, 2), flux=flux)
    sol = ml_cases.sod_solver(be, flux, 2)
    sol.create_new_simulation()
    sol.run()
    assert sol.current_run_time == 0.2
    rho, p = ml_cases.sod_norms(sol)
    want = ml_cases.SOD_CSV[(flux, 2)]
    print(f"flux {flux}: rho L1 {rho:.9e} (csv {want[0]:.8e}), p L1 {p:.9e} (csv {want[1]:.8e})")
    assert abs(rho / want[0] - 1.0) < 1e-6
    assert abs(p / want[1] - 1.0) < 1e-6
    # the mesh is adaptive: the finest level covers only the waves
    assert 0 < len(sol.topo.leaves_on_level(2)) < 64


def test_level0_driver_equals_uniform_oracle():
    """With maximumLevel 0 the driver is the uniform path: the oracle's
    uniform run gives the same bits."""
    import oracle_py
    be = oracle_ml.OracleBackend(2, 4)
    sol = ml_cases.sod_solver(be, gpu.FLUX_HLLC, 0)
    sol.create_new_simulation()
    sol.run()
    keys, got = sol.leaf_data()
    st0 = np.concatenate([ml_cases.sod_initial(0, (0, i, 0, 0)) for i in range(4)], axis=1)
    want, _, tf = oracle_py.run(2, [4, 1], [2, 2], 0, 0.6, 0.2, 100000, st0, cell_size=1.0 / 64)
    assert tf == sol.current_run_time
    assert np.array_equal(np.concatenate(list(got), axis=1), want)


@pytest.mark.parametrize("dim,lmax", [(2, 2), (3, 1)])
def test_periodic_conservation(dim, lmax):
    be = oracle_ml.OracleBackend(dim, ml_cases.capacity(dim, 4 if dim == 2 else 2, lmax))
    sol = ml_cases.blob_solver(be, dim, lmax)
    sol.create_new_simulation()
    s0, a0 = ml_cases.conserved_sums(sol)
    for _ in range(3 if dim == 2 else 2):
        sol.advance()
    s1, _ = ml_cases.conserved_sums(sol)
    drift = np.abs(s1 - s0)[:2] / a0[:2]  # mass, energy
    print(f"{dim}D drift {drift}")
    assert np.all(drift <= 1e-13)
    assert len(sol.topo.leaves_on_level(lmax)) > 0


def test_topology_rules():
    t = multilevel.Topology(2, [4, 1], 3, [True, False])
    assert t.neighbor((1, 7, 0, 0), (1, 0, 0)) == (1, 0, 0, 0)  # periodic x wraps
    assert t.neighbor((1, 3, 1, 0), (0, 1, 0)) is None          # y is not periodic
    assert t.neighbor((2, 0, 3, 0), (-1, 1, 0)) is None
    kids = t.children((1, 3, 1, 0))
    assert kids == [(2, 6, 2, 0), (2, 7, 2, 0), (2, 6, 3, 0), (2, 7, 3, 0)]
    assert [t.position(k) for k in kids] == [0, 1, 2, 3]
    assert all(t.parent(k) == (1, 3, 1, 0) for k in kids)
    assert multilevel.halo_locations(2) == [0, 1, 2, 3, 14, 15, 16, 17]
    assert len(multilevel.halo_locations(3)) == 26
    # GetLevels: level l advances 2^l times per macro step, all at the last
    counts = {lv: 0 for lv in range(4)}
    for ts in range(8):
        for lv in lts.levels_to_update(ts, 3):
            counts[lv] += 1
    assert counts == {0: 1, 1: 2, 2: 4, 3: 8}
    assert lts.levels_to_update(7, 3) == [3, 2, 1, 0]


def test_multilevel_output_read_by_reference_analysis(tmp_path):
    """A maximum-level-2 SodX run written on the finest-level mesh (leaves of
    levels 1 and 2) and read back with sod_analysis.py's recipe (cell centres
    and volumes from the vertices): the testsuite's norms (CSV row 21)."""
    from alpaca_amd import h5output
    from test_h5output import HAVE_HDF5, _read_like_sod_analysis
    import exact_riemann
    if not HAVE_HDF5:
        pytest.skip("libhdf5 not available")
    be = oracle_ml.OracleBackend(2, ml_cases.capacity(2, 4, 2), flux=gpu.FLUX_HLLC)
    sol = ml_cases.sod_solver(be, gpu.FLUX_HLLC, 2)
    sol.create_new_simulation()
    sol.run()
    h5 = h5output.write_multilevel(str(tmp_path / "sod_ml"), sol.current_run_time, sol)
    x, rho, u, p, vol, t = _read_like_sod_analysis(h5)
    assert t == 0.2 and len(rho) == 256 * len(sol.topo.leaves())
    ex = exact_riemann.sample((x - 0.5) / t, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    got_rho = np.sum(np.abs(rho / ex[0] - 1.0) * vol)
    got_p = np.sum(np.abs(p / ex[2] - 1.0) * vol)
    want = ml_cases.SOD_CSV[(gpu.FLUX_HLLC, 2)]
    assert abs(got_rho / want[0] - 1.0) < 1e-6 and abs(got_p / want[1] - 1.0) < 1e-6
    # the vertex extents give each leaf cell its own level's size
    levels = {round(v ** (1 / 3) * 64 * 4) for v in vol}
    assert levels <= {1, 2}


def test_node_ids_follow_reference_arithmetic():
    """Node ids of every level: level-0 ids as the reference's creation loop
    makes them, children (parent << 3) + position (IdsOfChildren), LevelOfNode
    counts the appended triples, and decoding inverts encoding."""
    from alpaca_amd import restart
    rng = np.random.default_rng(5)
    for _ in range(200):
        lv = int(rng.integers(0, 6))
        ijk = [int(rng.integers(0, 3 << lv)) for _ in range(3)]
        nid = restart.node_id(lv, *ijk)
        assert restart.level_of_node(nid) == lv
        assert restart.decode_node_id(nid) == (lv, *ijk)
        if lv > 0:
            parent = restart.node_id(lv - 1, ijk[0] >> 1, ijk[1] >> 1, ijk[2] >> 1)
            pos = (ijk[0] & 1) | ((ijk[1] & 1) << 1) | ((ijk[2] & 1) << 2)
            assert nid == (parent << 3) + pos and nid >> 3 == parent and nid & 7 == pos
        else:
            assert nid == restart.level0_id(*ijk)


def test_multilevel_restart_continuation_bitexact(tmp_path):
    """N macro steps, a snapshot of every node, M more; restoring the snapshot
    into a fresh solver and running M steps gives the same tree and bits."""
    from test_h5output import HAVE_HDF5
    if not HAVE_HDF5:
        pytest.skip("libhdf5 not available")
    cap = ml_cases.capacity(2, 4, 2)
    sol = ml_cases.sod_solver(oracle_ml.OracleBackend(2, cap), gpu.FLUX_HLLC, 2)
    sol.create_new_simulation()
    for _ in range(12):
        sol.advance()
    path = sol.write_restart(str(tmp_path / "restart_ml.h5"))
    for _ in range(6):
        sol.advance()
    again = ml_cases.sod_solver(oracle_ml.OracleBackend(2, cap), gpu.FLUX_HLLC, 2)
    again.restore(path)
    assert sorted(again.slot) == sorted(k for k in again.slot)
    for _ in range(6):
        again.advance()
    k1, c1 = sol.leaf_data()
    k2, c2 = again.leaf_data()
    assert k1 == k2 and np.array_equal(c1, c2)
    assert again.current_run_time == sol.current_run_time


def test_solver_refuses_a_backend_of_another_flux_or_gas():
    """The backend computes the right-hand side, so its flux family and gas
    must be the solver's: a mismatch is refused instead of silently running
    the backend's."""
    cap = ml_cases.capacity(2, 4, 2)
    with pytest.raises(ValueError, match="flux"):
        ml_cases.sod_solver(oracle_ml.OracleBackend(2, cap), gpu.FLUX_ROE, 2)
    with pytest.raises(ValueError, match="gamma"):
        multilevel.MultiLevelSolver(oracle_ml.OracleBackend(2, cap, gamma=1.67), 2, [4, 1],
                                    ml_cases.SOD_NODE, 2, [gpu.BC_SYMMETRY] * 4,
                                    ml_cases.sod_initial, 0.6, 0.2, 1e-3, 2)
