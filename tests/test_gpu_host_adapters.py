"""The C++ host adapters (alpaca_amd/csrc/host/alpaca_host.h), exercised
through their command-line driver, against the CPU oracle: the
ConvectiveTermSolver-shaped per-block call, the uniform-level
ComputeLoop/Advance built on the RungeKutta3TVD-shaped stage calls, and the
Multiresolution-shaped per-pair calls."""
import os
import struct
import subprocess

import numpy as np
import pytest

from alpaca_amd import cases, gpu
import oracle_py

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "alpaca_amd", "alpaca_selftest")


def _header(dim, nb, bc, flux, precision, max_steps, gamma, b, cell_size, cfl, end_time):
    nb3 = list(nb) + [1] * (3 - len(nb))
    bc3 = list(bc) + [0] * (3 - len(bc))
    return struct.pack("<10i5d", dim, *nb3, *bc3, flux, precision, max_steps, gamma, b,
                       cell_size, cfl, end_time)


def _run(mode, payload, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("alpaca_selftest is not built (run __graft_entry__.build())")
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    inp.write_bytes(payload)
    res = subprocess.run([EXE, mode, str(inp), str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    return out.read_bytes()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", [0, 1])
def test_convective_term_solver_adapter(tmp_path, dim, flux):
    print(f"[host fluxes] dim={dim} flux={flux}", flush=True)
    blk = cases.random_block(dim, 21 + flux)
    raw = _run("fluxes", _header(dim, [1, 1, 1], [0, 0, 0], flux, 0, 0, 1.4, 0.0, 1.0 / 64,
                                 0.0, 0.0) + blk.tobytes(), tmp_path)
    got = np.frombuffer(raw, dtype=np.float64)
    want = oracle_py.block_fluxes(dim, flux, 1.4, 0.0, blk, 1.0 / 64).ravel()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("flux", [0, 1])
def test_uniform_solver_adapter_2d(tmp_path, flux):
    print(f"[host solver 2d] flux={flux}", flush=True)
    nb = [4, 3]
    st0 = cases.four_quadrant_2d(4, 3, seed=9)
    raw = _run("run", _header(2, nb, [0, 0], flux, 0, 10000, 1.4, 0.0, 1.0 / 64, 0.6, 0.05) +
               st0.tobytes(), tmp_path)
    steps, t = struct.unpack("<id", raw[:12])
    got = np.frombuffer(raw[12:], dtype=np.float64).reshape(st0.shape)
    want, dts, tf = oracle_py.run(2, nb, [0, 0], flux, 0.6, 0.05, 10000, st0,
                                  cell_size=1.0 / 64)
    assert steps == len(dts) and t == tf
    assert np.array_equal(got, want)


def test_uniform_solver_adapter_3d_periodic(tmp_path):
    print("[host solver 3d]", flush=True)
    st0 = cases.blobs_shock_3d(2, seed=6)
    raw = _run("run", _header(3, [2, 2, 2], [1, 1, 1], 0, 0, 10, 1.4, 0.0, 1.0 / 32, 0.5,
                              1e300) + st0.tobytes(), tmp_path)
    steps, _ = struct.unpack("<id", raw[:12])
    got = np.frombuffer(raw[12:], dtype=np.float64).reshape(st0.shape)
    want, _, _ = oracle_py.run(3, [2, 2, 2], [1, 1, 1], 0, 0.5, 1e300, 10, st0)
    assert steps == 10
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("norm", [0, 1, 2])
def test_multiresolution_adapter(tmp_path, dim, norm):
    """GpuMultiresolution's reference-shaped calls (Average, Prediction over
    the block and over the internal cells, ChildNeedsRemeshing,
    AverageJumpBuffer) on one parent/child pair against the oracle."""
    import mr_states

    pos = (3 * norm + dim) % (2 ** dim)
    print(f"[host mr] dim={dim} norm={norm} position={pos}", flush=True)
    parent, child = mr_states.pair(dim, "smooth", pos, seed=norm)
    icz = 16 if dim == 3 else 1
    rng = np.random.default_rng(5 + norm)
    ps = rng.standard_normal((6, dim + 2, 16, icz))
    cs = rng.standard_normal((6, dim + 2, 16, icz))
    eps = 1.0e-3
    raw = _run("mr", struct.pack("<4id", dim, pos, norm, 0, eps) + parent.tobytes() +
               child.tobytes() + ps.tobytes() + cs.tobytes(), tmp_path)
    got = np.frombuffer(raw, dtype=np.float64)
    nb = parent.size
    avg, whole, inner = got[:nb], got[nb:2 * nb], got[2 * nb:3 * nb]
    detail, decision = got[3 * nb], int(got[3 * nb + 1])
    ns = 6 * (dim + 2) * 16 * icz
    surf = got[3 * nb + 2:3 * nb + 2 + ns]
    halos = got[3 * nb + 2 + ns:]
    assert np.array_equal(avg, oracle_py.mr_average(dim, child, parent, pos).ravel())
    rng_inner = (4, 16, 4, 16) + ((4, 16) if dim == 3 else (0, 1))
    want_whole = np.stack([oracle_py.mr_prediction(dim, parent[e], child[e], pos)
                           for e in range(dim + 2)])
    want_inner = np.stack([oracle_py.mr_prediction(dim, parent[e], child[e], pos, rng_inner)
                           for e in range(dim + 2)])
    assert np.array_equal(whole, want_whole.ravel())
    assert np.array_equal(inner, want_inner.ravel())
    want = oracle_py.mr_detail(dim, parent, child, pos, norm)
    assert abs(detail - want) <= 1e-13 * want
    assert decision == oracle_py.mr_decision(want, eps)
    assert np.array_equal(surf, oracle_py.mr_average_jump(dim, cs, ps, pos).ravel())
    # JumpHaloUpdateLinks over every location of the dimension
    from alpaca_amd import multiresolution as mr
    want_halo = child.copy()
    for loc in range(26):
        if dim == 2 and mr.location_direction(loc)[2] != 0:
            continue
        r = mr.halo_recv_range(dim, loc)
        for e in range(dim + 2):
            want_halo[e] = oracle_py.mr_prediction(dim, parent[e], want_halo[e], pos, r)
    assert np.array_equal(halos, want_halo.ravel())


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_level_time_integrator_adapter(tmp_path, dim, stage):
    """GpuLevelTimeIntegrator (C++): PrepareBufferForIntegration,
    IntegrateNode and IntegrateJumpHalos on one block with summed micro time
    steps, against the oracle's restatement of the reference's
    RungeKutta3TVD."""
    print(f"[host lts] dim={dim} stage={stage}", flush=True)
    tz, icz = (24, 16) if dim == 3 else (1, 1)
    rng = np.random.default_rng(17 * dim + stage)
    micro = rng.uniform(1e-4, 1e-3, 4)
    n = 2
    avg, rhs, init = (rng.standard_normal((dim + 2, 24, 24, tz)) for _ in range(3))
    jc, jf = (rng.standard_normal((6, dim + 2, 16, icz)) for _ in range(2))
    raw = _run("lts", struct.pack("<4i", dim, stage, n, len(micro)) + micro.tobytes() +
               avg.tobytes() + rhs.tobytes() + init.tobytes() + jc.tobytes() + jf.tobytes(),
               tmp_path)
    got = np.frombuffer(raw, dtype=np.float64)
    nb, ns = avg.size, jc.size
    g_avg, g_rhs = got[:nb], got[nb:2 * nb]
    g_jc, g_jf, g_rhs2 = got[2 * nb:2 * nb + ns], got[2 * nb + ns:2 * nb + 2 * ns], got[2 * nb + 2 * ns:]
    w_avg = oracle_py.lts_prepare_buffer(dim, avg, init, stage)
    w_rhs = oracle_py.lts_integrate(dim, w_avg, rhs, None, oracle_py.lts_timestep(micro, n, stage))
    k = 2 * dim
    w_jc, w_jf = jc.copy(), jf.copy()
    c, f = oracle_py.lts_integrate_jump(dim, jc[:k], jf[:k], oracle_py.lts_timestep(micro, n, stage, True))
    w_jc[:k], w_jf[:k] = c, f
    halo = (0, 4, 4, 16, 4 if dim == 3 else 0, 16 if dim == 3 else 1)
    w_rhs2 = oracle_py.lts_integrate(dim, w_avg, w_rhs, halo, oracle_py.lts_timestep(micro, n, stage))
    assert np.array_equal(g_avg, w_avg.ravel())
    assert np.array_equal(g_rhs, w_rhs.ravel())
    assert np.array_equal(g_jc, w_jc.ravel()) and np.array_equal(g_jf, w_jf.ravel())
    assert np.array_equal(g_rhs2, w_rhs2.ravel())
