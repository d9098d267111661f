"""ORACLE — test infrastructure only.  ctypes access to the CPU checkers:

* ``liboracle.so``: the C restatement (oracle/alpaca_oracle.c);
* ``_ref/libalpaca_ref{2,3}d.so``: the harness linking the reference's own
  translation units (built from the reference tree where it is present).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = {2: os.path.join(HERE, "_ref", "libalpaca_ref2d.so"),
           3: os.path.join(HERE, "_ref", "libalpaca_ref3d.so")}

_P = ctypes.POINTER(ctypes.c_double)


class OraGrid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nb", ctypes.c_int * 3), ("bc", ctypes.c_int * 3),
                ("flux", ctypes.c_int), ("gamma", ctypes.c_double), ("b", ctypes.c_double),
                ("cell_size", ctypes.c_double)]


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


_oracle = None
_refs = {}


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _oracle = ctypes.CDLL(ORACLE_LIB)
    return _oracle


def ref_available(dim: int) -> bool:
    return os.path.exists(REF_LIB[dim])


def ref(dim: int):
    if dim not in _refs:
        if not ref_available(dim):
            raise RuntimeError(f"{REF_LIB[dim]} is not built (make -C oracle ref)")
        _refs[dim] = ctypes.CDLL(REF_LIB[dim])
    return _refs[dim]


def ref_level0_ids(nb, dim=3):
    """Level-0 node ids of a block grid nb = (nx, ny, nz) in the order the
    reference's TopologyManager creates them (its own id arithmetic)."""
    nb3 = (ctypes.c_int * 3)(*(list(nb) + [1] * (3 - len(nb))))
    out = np.zeros(nb3[0] * nb3[1] * nb3[2], dtype=np.uint64)
    ref(dim).ref_level0_ids(nb3, out.ctypes.data_as(ctypes.c_void_p))
    return out


def nflux(dim):
    icz = 16 if dim == 3 else 1
    return 3 * (dim + 2) * 17 * 17 * (icz + 1)


def nrhs(dim):
    icz = 16 if dim == 3 else 1
    return (dim + 2) * 16 * 16 * icz


def weno5(v, which="oracle"):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(2)
    if which == "oracle":
        oracle().ora_weno5(_dp(v), _dp(out))
    else:
        ref(2).ref_weno5(_dp(v), _dp(out))
    return out


def eos(dim, gamma, b, cons5, prim5, which="oracle"):
    c = np.ascontiguousarray(cons5, dtype=np.float64)
    p = np.ascontiguousarray(prim5, dtype=np.float64)
    out = np.zeros(5)
    if which == "oracle":
        oracle().ora_eos(dim, ctypes.c_double(gamma), ctypes.c_double(b), _dp(c), _dp(p), _dp(out))
    else:
        ref(dim).ref_eos(ctypes.c_double(gamma), ctypes.c_double(b), _dp(c), _dp(p), _dp(out))
    return out


def hllc(dim, direction, gamma, b, cl, cr, which="oracle"):
    cl = np.ascontiguousarray(cl, dtype=np.float64)
    cr = np.ascontiguousarray(cr, dtype=np.float64)
    out = np.zeros(dim + 2)
    if which == "oracle":
        oracle().ora_hllc(dim, direction, ctypes.c_double(gamma), ctypes.c_double(b), _dp(cl),
                          _dp(cr), _dp(out))
    else:
        ref(dim).ref_hllc(direction, ctypes.c_double(gamma), ctypes.c_double(b), _dp(cl), _dp(cr),
                          _dp(out))
    return out


def roe_eigensystem(dim, direction, gamma, b, ct, cn):
    ne = dim + 2
    ct = np.ascontiguousarray(ct, dtype=np.float64)
    cn = np.ascontiguousarray(cn, dtype=np.float64)
    L = np.zeros((ne, ne))
    R = np.zeros((ne, ne))
    lam = np.zeros(ne)
    ok = oracle().ora_roe_eigensystem(dim, direction, ctypes.c_double(gamma), ctypes.c_double(b),
                                      _dp(ct), _dp(cn), _dp(L), _dp(R), _dp(lam))
    return ok, L, R, lam


def block_fluxes(dim, flux, gamma, b, blk, cell_size, which="oracle"):
    blk = np.ascontiguousarray(blk, dtype=np.float64)
    out = np.zeros(nflux(dim))
    if which == "oracle":
        oracle().ora_block_fluxes(dim, flux, ctypes.c_double(gamma), ctypes.c_double(b), _dp(blk),
                                  ctypes.c_double(cell_size), _dp(out))
    else:
        ref(dim).ref_block_fluxes(flux, ctypes.c_double(gamma), ctypes.c_double(b), _dp(blk),
                                  ctypes.c_double(cell_size), _dp(out))
    icz = 16 if dim == 3 else 1
    return out.reshape(3, dim + 2, 17, 17, icz + 1)


def block_rhs(dim, flux, gamma, b, blk, cell_size, which="oracle"):
    blk = np.ascontiguousarray(blk, dtype=np.float64)
    out = np.zeros(nrhs(dim))
    if which == "oracle":
        oracle().ora_block_rhs(dim, flux, ctypes.c_double(gamma), ctypes.c_double(b), _dp(blk),
                               ctypes.c_double(cell_size), _dp(out))
    else:
        ref(dim).ref_block_rhs(flux, ctypes.c_double(gamma), ctypes.c_double(b), _dp(blk),
                               ctypes.c_double(cell_size), _dp(out))
    icz = 16 if dim == 3 else 1
    return out.reshape(dim + 2, 16, 16, icz)


BC_ZERO_GRADIENT, BC_PERIODIC, BC_SYMMETRY = 0, 1, 2


def run(dim, nb, bc, flux, cfl, t_end, max_steps, state, gamma=1.4, b=0.0,
        cell_size=None, threads=8, which="oracle"):
    """Advance a domain state [NE][NX][NY][NZ]; returns (state, dts, t_final).
    bc: per direction (both sides) 0 zero-gradient, 1 periodic, 2 symmetry."""
    nb3 = list(nb) + [1] * (3 - len(nb))
    per3 = list(bc) + [0] * (3 - len(bc))
    if cell_size is None:
        cell_size = 1.0 / (nb3[0] * 16)
    st = np.ascontiguousarray(state, dtype=np.float64).copy()
    dts = np.zeros(max(max_steps, 1))
    tf = ctypes.c_double()
    if which == "oracle":
        g = OraGrid(dim, (ctypes.c_int * 3)(*nb3), (ctypes.c_int * 3)(*per3), flux, gamma, b,
                    cell_size)
        n = oracle().ora_run(ctypes.byref(g), ctypes.c_double(cfl), ctypes.c_double(t_end),
                             max_steps, threads, _dp(st), _dp(dts), ctypes.byref(tf))
    else:
        n = ref(dim).ref_run(flux, ctypes.c_double(gamma), ctypes.c_double(b),
                             (ctypes.c_int * 3)(*nb3), (ctypes.c_int * 3)(*per3),
                             ctypes.c_double(cell_size), ctypes.c_double(cfl),
                             ctypes.c_double(t_end), max_steps, threads, _dp(st), _dp(dts),
                             ctypes.byref(tf))
    return st, dts[:n], tf.value


# Multiresolution operators (child at sibling position `position` of a parent).
MR_NEUTRAL, MR_REFINE, MR_COARSE = 0, 1, 2


def _block(dim, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.shape == (dim + 2, 24, 24, 24 if dim == 3 else 1), a.shape
    return a


def mr_average(dim, child, parent, position, which="oracle"):
    """Parent block with the child's octant (quadrant) replaced by the
    conservative average of the child's internal cells."""
    child = _block(dim, child)
    out = _block(dim, parent).copy()
    if which == "oracle":
        oracle().ora_mr_average(dim, _dp(child), _dp(out), position)
    else:
        ref(dim).ref_mr_average(_dp(child), _dp(out), position)
    return out


def mr_prediction(dim, parent_field, child_field, position, rng=None, which="oracle"):
    """Child field [24][24][24|1] with the cells of rng = (x_start, x_count,
    y_start, y_count, z_start, z_count) predicted from the parent field."""
    tz = 24 if dim == 3 else 1
    if rng is None:
        rng = (0, 24, 0, 24, 0, tz)
    pf = np.ascontiguousarray(parent_field, dtype=np.float64)
    out = np.ascontiguousarray(child_field, dtype=np.float64).copy()
    assert pf.shape == (24, 24, tz) and out.shape == (24, 24, tz)
    r = (ctypes.c_int * 6)(*rng)
    if which == "oracle":
        oracle().ora_mr_prediction(dim, _dp(pf), _dp(out), position, r)
    else:
        ref(dim).ref_mr_prediction(_dp(pf), _dp(out), position, r)
    return out


MR_LINF, MR_L1, MR_L2 = 0, 1, 2


def mr_detail(dim, parent, child, position, norm=MR_LINF):
    f = oracle().ora_mr_detail
    f.restype = ctypes.c_double
    return f(dim, _dp(_block(dim, parent)), _dp(_block(dim, child)), position, norm)


def mr_detail_linf(dim, parent, child, position):
    return mr_detail(dim, parent, child, position, MR_LINF)


def mr_threshold(dim, maximum_level, reference_level, reference_epsilon, level, which="oracle"):
    if which == "oracle":
        f = oracle().ora_mr_threshold
        f.restype = ctypes.c_double
        return f(dim, maximum_level, reference_level, ctypes.c_double(reference_epsilon), level)
    f = ref(dim).ref_mr_threshold
    f.restype = ctypes.c_double
    return f(maximum_level, reference_level, ctypes.c_double(reference_epsilon), level)


def mr_decision(detail, eps_coarsen):
    return oracle().ora_mr_decision(ctypes.c_double(detail), ctypes.c_double(eps_coarsen))


def ref_mr_child_needs_remeshing(dim, parent, child, position, epsilon, norm=MR_LINF):
    """The reference's ChildNeedsRemeshing<norm> with the coarsening threshold
    on the child's level equal to `epsilon`."""
    return ref(dim).ref_mr_child_needs_remeshing(_dp(_block(dim, parent)), _dp(_block(dim, child)),
                                                 position, ctypes.c_double(epsilon), norm)


def ref_mr_detail(dim, parent, child, position, norm=MR_LINF):
    """The reference's L-infinity detail, read off its remeshing decision: the
    decision is Coarse exactly when detail <= epsilon, so the smallest
    non-negative double epsilon with a Coarse decision is the detail itself
    (bisection over the ordered bit patterns of non-negative doubles)."""
    lo, hi = 0, 0x7FF0000000000000  # 0.0 .. +inf as int64 bit patterns

    def coarse(bits):
        eps = np.array([bits], dtype=np.int64).view(np.float64)[0]
        return ref_mr_child_needs_remeshing(dim, parent, child, position, eps, norm) == MR_COARSE

    if coarse(lo):
        return 0.0
    while hi - lo > 1:          # invariant: not coarse at lo, coarse at hi
        mid = (lo + hi) // 2
        if coarse(mid):
            hi = mid
        else:
            lo = mid
    return float(np.array([hi], dtype=np.int64).view(np.float64)[0])


def mr_average_jump(dim, child, parent, position, which="oracle"):
    """Parent jump surfaces [6][NE][16][16|1] plus the child's averaged ones
    on the faces the child touches."""
    icz = 16 if dim == 3 else 1
    child = np.ascontiguousarray(child, dtype=np.float64)
    out = np.ascontiguousarray(parent, dtype=np.float64).copy()
    assert child.shape == (6, dim + 2, 16, icz) and out.shape == child.shape
    if which == "oracle":
        oracle().ora_mr_average_jump(dim, _dp(child), _dp(out), position)
    else:
        ref(dim).ref_mr_average_jump(_dp(child), _dp(out), position)
    return out


# Level-wise integration (RungeKutta3TVD with micro time steps).
def lts_timestep(micro, n, stage, jump=False):
    m = np.ascontiguousarray(micro, dtype=np.float64)
    f = oracle().ora_lts_timestep
    f.restype = ctypes.c_double
    return f(_dp(m), len(m), n, stage, int(bool(jump)))


def lts_integrate(dim, u_old, u_new, rng, timestep):
    """u_new with the cells of rng (None: internal cells) integrated."""
    out = _block(dim, u_new).copy()
    r = None if rng is None else (ctypes.c_int * 6)(*rng)
    oracle().ora_lts_integrate(dim, _dp(_block(dim, u_old)), _dp(out), r, ctypes.c_double(timestep))
    return out


def lts_integrate_jump(dim, cons, flux, timestep):
    c = np.ascontiguousarray(cons, dtype=np.float64).copy()
    f = np.ascontiguousarray(flux, dtype=np.float64).copy()
    oracle().ora_lts_integrate_jump(dim, _dp(c), _dp(f), ctypes.c_double(timestep))
    return c, f


def lts_prepare_buffer(dim, avg, init, stage):
    out = _block(dim, avg).copy()
    oracle().ora_lts_prepare_buffer(dim, _dp(out), _dp(_block(dim, init)), stage)
    return out


def ref_lts_integrate(dim, op, stage, n, micro, start, size, avg, rhs, init, jcons, jflux):
    """The reference's RungeKutta3TVD on one Node: op 0 PrepareBufferForIntegration
    + IntegrateNode, op 1 IntegrateJumpHalos; returns the five buffers."""
    m = np.ascontiguousarray(micro, dtype=np.float64)
    bufs = [_block(dim, avg).copy(), _block(dim, rhs).copy(), _block(dim, init).copy(),
            np.ascontiguousarray(jcons, dtype=np.float64).copy(),
            np.ascontiguousarray(jflux, dtype=np.float64).copy()]
    st = (ctypes.c_int * 3)(*start)
    sz = (ctypes.c_int * 3)(*size)
    ref(dim).ref_lts_integrate(op, stage, n, _dp(m), len(m), st, sz, *[_dp(b) for b in bufs])
    return bufs


def block_jump_fluxes(dim, flux, gamma, b, blk, jump):
    """Jump-flux surfaces [6][NE][16][16|1] plus the block's boundary face
    fluxes (SpaceSolver::UpdateFluxes' boundary-flux save)."""
    out = np.ascontiguousarray(jump, dtype=np.float64).copy()
    oracle().ora_block_jump_fluxes(dim, flux, ctypes.c_double(gamma), ctypes.c_double(b),
                                   _dp(_block(dim, blk)), _dp(out))
    return out


def lts_jump_flux_adjustment(dim, cells, jump_leaf, jump_nb, one_cell_size):
    """(cells, jump_leaf) after JumpFluxAdjustment's exchange for one leaf;
    jump_nb: six neighbour surfaces or None per face location."""
    c = _block(dim, cells).copy()
    j = np.ascontiguousarray(jump_leaf, dtype=np.float64).copy()
    keep = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in jump_nb]
    ptrs = (_P * 6)(*[(_dp(x) if x is not None else _P()) for x in keep])
    oracle().ora_lts_jump_flux_adjustment(dim, _dp(c), _dp(j), ptrs, ctypes.c_double(one_cell_size))
    return c, j
