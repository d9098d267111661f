"""The multi-level driver on the GPU backend (gfx950 kernels on HBM pools)
against the same driver on the CPU oracle backend: the same tree and the
same bits in every leaf, the testsuite's DetailedSod norms at maximum level
2, and conservation under periodic sides (2D and 3D)."""
import numpy as np
import pytest

from alpaca_amd import gpu, multilevel
import ml_cases
import oracle_ml

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _pair(make, dim, cap, flux):
    g = make(multilevel.GpuBackend(dim, cap, flux=flux))
    o = make(oracle_ml.OracleBackend(dim, cap, flux=flux))
    for sol in (g, o):
        sol.create_new_simulation()
    return g, o


def _same(g, o, what):
    kg, cg = g.leaf_data()
    ko, co = o.leaf_data()
    assert kg == ko, f"{what}: trees differ"
    assert np.array_equal(cg, co), f"{what}: max diff {np.abs(cg - co).max()}"
    assert g.current_run_time == o.current_run_time


@pytest.mark.parametrize("flux", [gpu.FLUX_HLLC, gpu.FLUX_ROE])
def test_sod_level2_gpu_equals_oracle_and_csv(torch, flux):
    print(f"[multilevel sod] flux={flux}", flush=True)
    cap = ml_cases.capacity(2, 4, 2)
    g, o = _pair(lambda be: ml_cases.sod_solver(be, flux, 2), 2, cap, flux)
    _same(g, o, "initial")
    while g.current_run_time < g.t_end:
        g.advance()
        o.advance()
        _same(g, o, f"macro step {g.macro_steps}")
    rho, p = ml_cases.sod_norms(g)
    want = ml_cases.SOD_CSV[(flux, 2)]
    print(f"  rho L1 {rho:.9e} (csv {want[0]:.8e}) p L1 {p:.9e} (csv {want[1]:.8e}), "
          f"{g.macro_steps} macro steps", flush=True)
    assert abs(rho / want[0] - 1.0) < 1e-6 and abs(p / want[1] - 1.0) < 1e-6


@pytest.mark.parametrize("dim,lmax,steps", [(2, 2, 4), (3, 2, 1)])
def test_periodic_gpu_equals_oracle_and_conserves(torch, dim, lmax, steps):
    print(f"[multilevel periodic] dim={dim} lmax={lmax}", flush=True)
    cap = ml_cases.capacity(dim, 4 if dim == 2 else 2, lmax)
    g, o = _pair(lambda be: ml_cases.blob_solver(be, dim, lmax), dim, cap, gpu.FLUX_HLLC)
    s0, a0 = ml_cases.conserved_sums(g)
    for n in range(steps):
        g.advance()
        o.advance()
        _same(g, o, f"macro step {n + 1}")
    s1, _ = ml_cases.conserved_sums(g)
    drift = np.abs(s1 - s0)[:2] / a0[:2]
    print(f"  drift mass/energy {drift}", flush=True)
    assert np.all(drift <= 1e-13)


def test_restart_written_by_cpu_continues_on_gpu(torch, tmp_path):
    """A multi-level snapshot written by the CPU backend's run, restored by
    the GPU backend: both continue to the same tree and bits."""
    print("[multilevel restart]", flush=True)
    cap = ml_cases.capacity(2, 4, 2)
    o = ml_cases.sod_solver(oracle_ml.OracleBackend(2, cap, flux=gpu.FLUX_ROE), gpu.FLUX_ROE, 2)
    o.create_new_simulation()
    for _ in range(8):
        o.advance()
    path = o.write_restart(str(tmp_path / "ml.h5"))
    g = ml_cases.sod_solver(multilevel.GpuBackend(2, cap, flux=gpu.FLUX_ROE), gpu.FLUX_ROE, 2)
    g.restore(path)
    for n in range(4):
        o.advance()
        g.advance()
        _same(g, o, f"after restore, step {n + 1}")


class _ExchangingBackend(multilevel.GpuBackend):
    """The GPU backend with SwapBuffers as the reference's exchange of the
    Average and RightHandSide buffers (the backend itself copies)."""

    def swap(self, slots):
        lst = self._list(slots)
        gpu.check(self.lib.alpaca_swap_blocks_list(self.dim, self._p(self.avg), self._p(self.rhs),
                                                   self.capacity, self._p(lst), len(slots),
                                                   self.st), "swap_blocks_list")


@pytest.mark.parametrize("dim,lmax,nb0,bc,steps", [
    (3, 3, [2, 2, 2], gpu.BC_ZERO_GRADIENT, 3),
    (2, 4, [4, 4], gpu.BC_ZERO_GRADIENT, 4),
    (2, 3, [4, 2], gpu.BC_SYMMETRY, 6)])
def test_swap_as_copy_equals_exchange(torch, dim, lmax, nb0, bc, steps):
    """The backend's one-way copy for SwapBuffers leaves every Average
    buffer with the bits of the reference's exchange, over trees with level
    jumps at the domain sides, maximum levels 3 and 4."""
    from alpaca_amd import cases
    print(f"[swap as copy] dim={dim} lmax={lmax}", flush=True)
    node = 0.25 if dim == 3 else 0.5
    cap = int(np.prod(nb0)) * sum((1 << dim) ** lv for lv in range(lmax + 1))
    out = []
    for cls in (multilevel.GpuBackend, _ExchangingBackend):
        be = cls(dim, cap, flux=gpu.FLUX_HLLC)
        sol = multilevel.MultiLevelSolver(be, dim, nb0, node, lmax, [bc] * (2 * dim),
                                          cases.blob_node_initial(dim, node), 0.5, 10.0, 1e-3,
                                          lmax)
        sol.create_new_simulation()
        for _ in range(steps):
            sol.advance()
        keys = sorted(sol.slot)
        out.append((keys, be.get_blocks(sol._slots(keys)), sol.current_run_time))
        be.close()
    (ka, a, ta), (kb, b, tb) = out
    assert ka == kb and ta == tb and len(ka) > int(np.prod(nb0))
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
