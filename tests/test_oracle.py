"""CPU tests of the oracle (no GPU).

The oracle (oracle/alpaca_oracle.c) is pinned by:
  * golden vectors written by tests/golden/make_golden.py from a harness that
    links the reference's own translation units (bit-exact comparisons);
  * the reference's unit-test known answers (WENO5, stiffened-gas EOS);
  * the reference testsuite's DetailedSod level-0 error norms;
  * closed forms and invariants: L.R = I, the Roe property A(U~)(U_R - U_L) =
    F(U_R) - F(U_L), HLLC consistency and upwinding, conservation under
    periodic sides, the CFL step of a uniform state, x/y symmetry.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from alpaca_amd import cases
import exact_riemann
import oracle_py

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def vec():
    return np.load(os.path.join(GOLDEN, "golden_vectors.npz"))


@pytest.fixture(scope="module")
def runs():
    with open(os.path.join(GOLDEN, "golden_runs.json")) as fh:
        return json.load(fh)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


# ------------------------------------------------------------ golden vectors
def test_weno5_golden(vec):
    got = np.array([oracle_py.weno5(v) for v in vec["weno_in"]])
    assert np.array_equal(got, vec["weno_out"])


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("tag,gamma,b", [("ideal", 1.4, 0.0), ("stiff", 1.6, 2.0)])
def test_eos_golden(vec, dim, tag, gamma, b):
    cons = vec[f"eos{dim}_{tag}_cons"]
    prim = vec[f"eos{dim}_{tag}_prim"]
    got = np.array([oracle_py.eos(dim, gamma, b, c, p) for c, p in zip(cons, prim)])
    assert np.array_equal(got, vec[f"eos{dim}_{tag}_out"])


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_golden(vec, dim):
    cl, cr = vec[f"hllc{dim}_left"], vec[f"hllc{dim}_right"]
    for d in range(dim):
        got = np.array([oracle_py.hllc(dim, d, 1.4, 0.0, a, b) for a, b in zip(cl, cr)])
        assert np.array_equal(got, vec[f"hllc{dim}_dir{d}"]), f"direction {d}"


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flux", [0, 1])
def test_block_fluxes_and_rhs_golden(vec, runs, dim, flux):
    for seed, smooth in ((11, False), (12, True)):
        key = f"block{dim}_flux{flux}_seed{seed}"
        meta = runs["blocks"][key]
        blk = cases.random_block(dim, seed, smooth=smooth)
        f = oracle_py.block_fluxes(dim, flux, 1.4, 0.0, blk, 1.0 / 128)
        r = oracle_py.block_rhs(dim, flux, 1.4, 0.0, blk, 1.0 / 128)
        assert sha(f) == meta["fluxes_sha256"]
        assert sha(r) == meta["rhs_sha256"]
        if dim == 2:
            assert np.array_equal(f, vec[key + "_fluxes"])
            assert np.array_equal(r, vec[key + "_rhs"])
        else:
            assert np.array_equal(f.ravel()[::97], vec[key + "_fluxes_sample"])
            assert np.array_equal(r.ravel()[::13], vec[key + "_rhs_sample"])


def _initial_state(run):
    ic = run["ic"]
    if ic["kind"] == "four_quadrant_2d":
        return cases.four_quadrant_2d(run["nb"][0], run["nb"][1], seed=ic["seed"])
    if ic["kind"] == "blobs_shock_3d":
        return cases.blobs_shock_3d(run["nb"][0], seed=ic["seed"])
    return sod_state()


def sod_state(axis=0):
    """Sod tube on [0,1]x[0,1/4] along x (axis 0) or its transpose along y."""
    n_long, n_short = 64, 16
    x = (np.arange(n_long) + 0.5) / 64
    X = np.repeat(x[:, None], n_short, axis=1)
    left = X < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    st = cases.primes_to_conservatives(2, 1.4, 0.0, rho, (z, z), p)[..., None]
    if axis == 1:
        st = np.ascontiguousarray(st.transpose(0, 2, 1, 3))
    return st


def test_trajectories_golden(runs):
    """Every golden trajectory (2D four-quadrant to t_end, 3D periodic 100
    steps, Sod) is reproduced bit for bit: step count, every dt, final time
    and the SHA-256 of the final state."""
    for run in runs["runs"]:
        st0 = _initial_state(run)
        st, dts, tf = oracle_py.run(run["dim"], run["nb"], run["bc"], run["flux"],
                                    run["cfl"], run["t_end"], run["max_steps"], st0,
                                    cell_size=run.get("cell_size"))
        assert len(dts) == run["steps"], run["name"]
        assert dts.tolist() == run["dts"], run["name"]
        assert tf == run["t_final"], run["name"]
        assert sha(st) == run["sha256"], run["name"]


# ------------------------------------------- reference unit-test known answers
def test_weno5_reference_unit_tests():
    """test_stencils.cpp:144-149: unity within 1e-16, step within 1e-11."""
    out = oracle_py.weno5(np.ones(6))
    assert abs(out[0] - 1.0) <= 1e-16 and abs(out[1] - 1.0) <= 1e-16
    out = oracle_py.weno5(np.array([0, 0, 0, 1, 1, 1.0]))
    assert abs(out[0]) <= 1e-11 and abs(out[1] - 1.0) <= 1e-11


def _approx_eps(a, b):
    return abs(a - b) <= np.finfo(float).eps * (abs(b) + 1.0)


def test_stiffened_gas_reference_known_answers():
    """test_equations_of_state.cpp:159-182 (gamma 1.6, B 2; 3D build) and
    test_generic_stiffened_gas.cpp:101-115 (energy 15.577)."""
    cons = np.array([1.2, 0.36, 2.7, 3.05, 4.5])  # mass, mx, my, mz, energy
    prim = np.array([2.4, 0.75, 1.9, 4.003, 3.75])  # rho, u, v, w, p
    out = oracle_py.eos(3, 1.6, 2.0, cons, prim)
    assert _approx_eps(out[0], -4.6805249999999994)
    assert _approx_eps(out[1], -0.1504374999999995)
    assert _approx_eps(out[2], 1.957890020745122)
    assert _approx_eps(out[3], 35.819144133333332)
    assert _approx_eps(out[4], 2.8958333333333335)
    out = oracle_py.eos(3, 1.4, 3.0, np.array([1, 0, 0, 0, 1.0]),
                        np.array([1.1, 0.1, 0.2, 0.3, 2.0]))
    assert abs(out[3] - 15.577) <= 1e-15 + 1e-15 * 15.577


# ------------------------------------------------ reference testsuite pin
@pytest.mark.parametrize("flux,rho_l1,p_l1", [(1, 1.00420788e-04, 1.13784806e-04),
                                              (0, 1.02588746e-04, 1.12368183e-04)])
def test_detailed_sod_level0_norms(flux, rho_l1, p_l1):
    """2D SodX at level 0 (4x1 blocks, 64x16 cells, CFL 0.6, t = 0.2):
    relative error norms sum |q/q_exact - 1| V against the exact solution
    (mathematical_operations.py norms, cell volume dx^3 as the 2D output
    writes it) match 2D_DetailedSod_errors_reference.csv rows 2 (Roe flux
    splitting) and 20 (FV-HLLC).  The testsuite passes at 3%; the oracle
    agrees to 1e-6."""
    st, dts, tf = oracle_py.run(2, [4, 1], [0, 0], flux, 0.6, 0.2, 100000, sod_state(),
                                cell_size=1.0 / 64)
    assert tf == 0.2
    x = (np.arange(64) + 0.5) / 64
    ex = exact_riemann.sample((x - 0.5) / tf, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    r = st[0, :, :, 0]
    p = 0.4 * (st[1, :, :, 0] - 0.5 * (st[2, :, :, 0] ** 2 + st[3, :, :, 0] ** 2) / r)
    vol = (1.0 / 64) ** 3
    got_rho = np.sum(np.abs(r / ex[0][:, None] - 1.0) * vol)
    got_p = np.sum(np.abs(p / ex[2][:, None] - 1.0) * vol)
    assert abs(got_rho / rho_l1 - 1.0) < 1e-6
    assert abs(got_p / p_l1 - 1.0) < 1e-6


@pytest.mark.parametrize("flux", [0, 1])
def test_sod_x_y_symmetry(flux):
    """SodY equals the transposed SodX (CSV rows 2/5 and 20/23 coincide)."""
    sx, _, _ = oracle_py.run(2, [4, 1], [0, 0], flux, 0.6, 0.2, 100000, sod_state(0),
                             cell_size=1.0 / 64)
    sy, _, _ = oracle_py.run(2, [1, 4], [0, 0], flux, 0.6, 0.2, 100000, sod_state(1),
                             cell_size=1.0 / 64)
    sy_t = sy.transpose(0, 2, 1, 3)[[0, 1, 3, 2]]
    assert np.max(np.abs(sx - sy_t)) <= 1e-14


# ---------------------------------------------------- closed forms/invariants
def _rand_state(rng, dim):
    rho = rng.uniform(0.2, 2.0)
    p = rng.uniform(0.1, 2.0)
    vel = [rng.uniform(-1.0, 1.0) for _ in range(dim)]
    c = cases.primes_to_conservatives(dim, 1.4, 0.0, np.array(rho), [np.array(v) for v in vel],
                                      np.array(p))
    return c.astype(float).ravel(), rho, p, vel


def _euler_flux(dim, d, c, gamma=1.4):
    rho = c[0]
    m = c[2:2 + dim]
    u = m / rho
    p = (gamma - 1.0) * (c[1] - 0.5 * np.dot(m, m) / rho)
    f = np.empty(dim + 2)
    f[0] = m[d]
    f[1] = u[d] * (c[1] + p)
    f[2:] = m * u[d]
    f[2 + d] += p
    return f


@pytest.mark.parametrize("dim", [2, 3])
def test_roe_eigensystem_invariants(dim):
    """L.R = I, and the Roe matrix A = R diag(u-c, u.., u+c) L maps the jump
    of the conservatives onto the jump of the Euler fluxes (Roe's property U)
    for every direction."""
    rng = np.random.default_rng(7)
    ne = dim + 2
    for _ in range(20):
        ct, *_ = _rand_state(rng, dim)
        cn, *_ = _rand_state(rng, dim)
        for d in range(dim):
            ok, L, R, lam = oracle_py.roe_eigensystem(dim, d, 1.4, 0.0, ct, cn)
            assert ok
            assert np.allclose(L @ R, np.eye(ne), atol=1e-12)
            mp = 2 + d
            u = 0.5 * (R[mp, 0] + R[mp, ne - 1])
            c = 0.5 * (R[mp, ne - 1] - R[mp, 0])
            lam_signed = np.array([u - c] + [u] * (ne - 2) + [u + c])
            assert np.allclose(lam, np.abs(lam_signed), rtol=0, atol=1e-14)
            A = R @ np.diag(lam_signed) @ L
            dF = _euler_flux(dim, d, cn) - _euler_flux(dim, d, ct)
            assert np.allclose(A @ (cn - ct), dF, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_consistency_and_upwinding(dim):
    """HLLC(U, U) = F(U); supersonic states to the right (S_L >= 0) give
    F(U_L), to the left (S_R <= 0) give F(U_R) (Toro ch. 10.4)."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        c, rho, p, vel = _rand_state(rng, dim)
        for d in range(dim):
            f = oracle_py.hllc(dim, d, 1.4, 0.0, c, c)
            assert np.allclose(f, _euler_flux(dim, d, c), rtol=1e-13, atol=1e-13)
    for d in range(dim):
        vel = [0.0] * dim
        vel[d] = 5.0  # Mach > 3 to the right
        cl = cases.primes_to_conservatives(dim, 1.4, 0.0, np.array(1.0),
                                           [np.array(v) for v in vel], np.array(1.0)).ravel()
        cr = cases.primes_to_conservatives(dim, 1.4, 0.0, np.array(0.5),
                                           [np.array(v) for v in vel], np.array(0.4)).ravel()
        assert np.allclose(oracle_py.hllc(dim, d, 1.4, 0.0, cl, cr), _euler_flux(dim, d, cl),
                           rtol=1e-13, atol=1e-13)
        vel[d] = -5.0
        cl = cases.primes_to_conservatives(dim, 1.4, 0.0, np.array(1.0),
                                           [np.array(v) for v in vel], np.array(1.0)).ravel()
        cr = cases.primes_to_conservatives(dim, 1.4, 0.0, np.array(0.5),
                                           [np.array(v) for v in vel], np.array(0.4)).ravel()
        assert np.allclose(oracle_py.hllc(dim, d, 1.4, 0.0, cl, cr), _euler_flux(dim, d, cr),
                           rtol=1e-13, atol=1e-13)


def test_weno5_fifth_order_on_smooth_data():
    """Upwind-left reconstruction of cell averages of sin(x) converges at
    fifth order (error ratio ~32 per halving)."""
    errs = []
    for h in (0.2, 0.1, 0.05):
        xs = (np.arange(6) - 2) * h
        avg = (np.cos(xs - h / 2) - np.cos(xs + h / 2)) / h  # averages of sin over cells
        exact_face = np.sin(h / 2)
        errs.append(abs(oracle_py.weno5(avg)[0] - exact_face))
    assert errs[0] / errs[1] > 20 and errs[1] / errs[2] > 25


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_state_timestep_and_steady(dim):
    """A uniform state is steady, and its CFL step is the closed form
    dt = CFL / (sum_d(|u_d| + c) / dx) (modular_algorithm_assembler.cpp:
    1840-1982)."""
    nb = [2, 2, 2] if dim == 3 else [2, 2]
    n = [16 * b for b in nb]
    rho, p = 0.9, 1.3
    vel = [0.3, -0.2, 0.1][:dim]
    shp = tuple(n)
    c = cases.primes_to_conservatives(dim, 1.4, 0.0, np.full(shp, rho),
                                      [np.full(shp, v) for v in vel], np.full(shp, p))
    st0 = c if dim == 3 else c[..., None]
    st, dts, _ = oracle_py.run(dim, nb, [0] * dim, 0, 0.6, 1e300, 1, st0)
    dx = 1.0 / n[0]
    cs = np.sqrt(1.4 * p / rho)
    expected = 0.6 / (sum(abs(v) + cs for v in vel) / dx)
    assert dts[0] == pytest.approx(expected, rel=1e-14)
    assert np.max(np.abs(st - st0)) <= 1e-13 * np.max(np.abs(st0))


def test_periodic_conservation_3d():
    """Sums of mass, momentum and energy are conserved to rounding under
    periodic sides (telescoping flux differences)."""
    st0 = cases.blobs_shock_3d(2, seed=3)
    for flux in (0, 1):
        st, _, _ = oracle_py.run(3, [2, 2, 2], [1, 1, 1], flux, 0.5, 1e300, 20, st0)
        s0 = st0.sum(axis=(1, 2, 3))
        s1 = st.sum(axis=(1, 2, 3))
        scale = np.abs(st0).sum(axis=(1, 2, 3))
        scale[2:] = np.abs(st0[2:]).sum() + np.abs(st[2:]).sum()
        assert np.all(np.abs(s1 - s0) / scale <= 1e-13)


@pytest.mark.skipif(not oracle_py.ref_available(3), reason="reference harness not built here")
@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_equals_reference_harness_on_fresh_blocks(dim):
    """Where the reference harness is built: fresh random blocks agree bit for
    bit (beyond the committed fixtures)."""
    for seed in (101, 102):
        blk = cases.random_block(dim, seed)
        for flux in (0, 1):
            a = oracle_py.block_rhs(dim, flux, 1.4, 0.0, blk, 1.0 / 64)
            b = oracle_py.block_rhs(dim, flux, 1.4, 0.0, blk, 1.0 / 64, which="ref")
            assert np.array_equal(a, b)


def test_symmetry_equals_mirrored_periodic_domain():
    """Symmetry sides on [0,1] equal the right half of the mirror-extended
    domain [-1,1] with periodic sides (mirror: rho, p even; u odd in x): the
    mirror image of every halo cell is then an internal cell.  A wrong mirror
    offset or a missing sign flip of the normal momentum breaks this."""
    nbx, nby = 3, 2
    st = cases.four_quadrant_2d(nbx, nby, seed=11)
    mirrored = st[:, ::-1, :, :].copy()
    mirrored[2] *= -1.0  # x momentum is odd under the mirror
    ext = np.concatenate([mirrored, st], axis=1)
    a, da, ta = oracle_py.run(2, [nbx, nby], [2, 2], 0, 0.6, 0.15, 10000, st,
                              cell_size=1.0 / (16 * nbx))
    b, db, tb = oracle_py.run(2, [2 * nbx, nby], [1, 2], 0, 0.6, 0.15, 10000, ext,
                              cell_size=1.0 / (16 * nbx))
    assert len(da) == len(db)
    assert np.max(np.abs(np.array(da) - np.array(db))) <= 1e-14 * max(da)
    half = b[:, 16 * nbx:]
    scale = np.abs(a).reshape(4, -1).max(axis=1)[:, None, None, None]
    assert np.max(np.abs(half - a) / scale) <= 1e-12
    # and the left half is the mirror image
    left = b[:, :16 * nbx][:, ::-1]
    assert np.max(np.abs(left[[0, 1, 3]] - a[[0, 1, 3]]) / scale[[0, 1, 3]]) <= 1e-12
    assert np.max(np.abs(-left[2] - a[2]) / scale[2]) <= 1e-12
