"""Cases of the multi-level driver tests: the reference's DetailedSod SodX
input (testsuite/InputFiles/DetailedSod/2D/SodX.xml: 4x1 level-0 nodes of
size 0.25, symmetry sides, epsilon 1e-3 at level 2, CFL 0.6, t_end 0.2) and
a periodic blob case, with the norms of the testsuite's Sod analysis."""
import numpy as np

from alpaca_amd import cases, gpu, multilevel
import exact_riemann

SOD_NODE = 0.25
# 2D_DetailedSod_errors_reference.csv: (DensityL1, PressureL1) by (flux, level)
SOD_CSV = {(gpu.FLUX_ROE, 0): (1.00420788e-04, 1.13784806e-04),
           (gpu.FLUX_ROE, 2): (7.19817501e-06, 6.43090634e-06),
           (gpu.FLUX_ROE, 4): (4.16316344e-07, 3.16522804e-07),
           (gpu.FLUX_HLLC, 0): (1.02588746e-04, 1.12368183e-04),
           (gpu.FLUX_HLLC, 2): (7.06139876e-06, 6.24712396e-06),
           (gpu.FLUX_HLLC, 4): (4.07856324e-07, 3.10301306e-07)}


def capacity(dim, nb0_count, lmax):
    return nb0_count * sum((1 << dim) ** lv for lv in range(lmax + 1))


def sod_initial(level, key):
    """SodX's initial condition at a level-`level` node's cell centres."""
    dx = SOD_NODE / (1 << level) / gpu.IC
    x = (key[1] * gpu.IC + np.arange(gpu.IC) + 0.5) * dx
    X = np.repeat(x[:, None], gpu.IC, axis=1)
    left = X < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    return cases.primes_to_conservatives(2, 1.4, 0.0, rho, (z, z), p)[..., None]


def sod_solver(backend, flux, lmax):
    return multilevel.MultiLevelSolver(backend, 2, [4, 1], SOD_NODE, lmax,
                                       [gpu.BC_SYMMETRY] * 4, sod_initial, 0.6, 0.2, 1e-3, 2,
                                       flux=flux)


def sod_norms(sol):
    """DensityL1, PressureL1 of the leaves against the exact solution: sum
    |q / q_exact - 1| V over every leaf cell (sod_analysis.py with
    mathematical_operations.py), V = dx^3 of the leaf's level, as the
    level-0 pin of these norms fixes it."""
    centres, _, v = sol.leaf_cells()
    t = sol.current_run_time
    ex = exact_riemann.sample((centres[:, 0] - 0.5) / t, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    keys = sol.topo.leaves()
    dx = np.repeat([sol.cell_size(k[0]) for k in keys], gpu.IC * gpu.IC)
    vol = dx ** 3
    r = v[:, 0]
    p = 0.4 * (v[:, 1] - 0.5 * (v[:, 2] ** 2 + v[:, 3] ** 2) / r)
    return float(np.sum(np.abs(r / ex[0] - 1.0) * vol)), float(np.sum(np.abs(p / ex[2] - 1.0) * vol))


def blob_initial(dim, node_size):
    """Periodic case: a dense, high-pressure blob in a uniform flow."""
    return cases.blob_node_initial(dim, node_size)


def blob_solver(backend, dim, lmax, flux=gpu.FLUX_HLLC):
    nb0 = [2, 2] if dim == 2 else [2, 1, 1]
    return multilevel.MultiLevelSolver(backend, dim, nb0, 0.5, lmax, [gpu.BC_PERIODIC] * (2 * dim),
                                       blob_initial(dim, 0.5), 0.6 if dim == 2 else 0.5, 10.0,
                                       1e-3, lmax, flux=flux)


def conserved_sums(sol):
    """Per equation: sum over leaf cells of q dx^dim, and of |q| dx^dim."""
    keys, c = sol.leaf_data()
    dim = sol.dim
    tot = np.zeros(dim + 2)
    ab = np.zeros(dim + 2)
    for k, cc in zip(keys, c):
        v = sol.cell_size(k[0]) ** dim
        tot += cc.reshape(dim + 2, -1).sum(axis=1) * v
        ab += np.abs(cc).reshape(dim + 2, -1).sum(axis=1) * v
    return tot, ab
