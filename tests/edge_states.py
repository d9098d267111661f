"""Haloed blocks that drive the face numerics into their special branches.

Each kind overwrites part of a random block (alpaca_amd.cases.random_block)
with states that reach one branch of the reference's per-face code:

  vacuum    a slab with rho = 1e-17: reconstructed densities <= DBL_EPSILON,
            the face is skipped (finite_volume_scheme.cpp:218-225)
  negative  a slab whose energy is below its kinetic energy (p < 0 = -B): the
            face is skipped (same lines)
  rest      a uniform state at rest: S* = 0 exactly, Signum(0) = 0 and the
            HLLC flux is the average of the two star fluxes
            (hllc_riemann_solver.h:184-195)
  tiny      a slab with rho and p of 1e-240 (square roots below 2^-767 take
            the scaled path of the hardware square root)
  huge      a slab with p = 1e108 (WENO5 smoothness indicators beyond 2^700:
            the divisions take the scaled path)
  zero_rho  a slab with rho = 0: the Roe eigensystem is skipped
            (eigendecomposition.h:250), giving non-finite primes elsewhere

Used by the oracle-against-reference test (CPU) and the GPU-against-oracle
test; NaN positions compare equal.
"""
import numpy as np

from alpaca_amd import cases

KINDS = ("vacuum", "negative", "rest", "tiny", "huge", "zero_rho")


def _slab(dim, lo=10, hi=13):
    idx = [slice(None)] * 4
    idx[1] = slice(lo, hi)
    return tuple(idx)


def edge_block(dim, kind, seed=0, gamma=1.4):
    blk = cases.random_block(dim, seed, gamma=gamma)
    shape = blk.shape[1:]
    if kind == "rest":
        rho = np.full(shape, 1.3)
        p = np.full(shape, 0.7)
        vel = [np.zeros(shape) for _ in range(dim)]
        return np.ascontiguousarray(cases.primes_to_conservatives(dim, gamma, 0.0, rho, vel, p))
    sl = _slab(dim)
    region = blk[sl]
    n = region.shape[1:]
    rng = np.random.default_rng(seed + 100)
    vel = [rng.uniform(-0.5, 0.5, n) for _ in range(dim)]
    if kind == "vacuum":
        rho, p = np.full(n, 1e-17), np.full(n, 1e-5)
    elif kind == "tiny":
        rho, p = np.full(n, 1e-240), np.full(n, 1e-240)
    elif kind == "huge":
        rho, p = np.full(n, 1.0), np.full(n, 1e108)
    elif kind == "zero_rho":
        rho, p = np.zeros(n), np.full(n, 1.0)
        vel = [np.zeros(n) for _ in range(dim)]
    elif kind == "negative":
        rho, p = np.full(n, 1.0), np.full(n, 0.5)
        vel = [np.full(n, 1.0)] + [np.zeros(n) for _ in range(dim - 1)]
    else:
        raise ValueError(kind)
    cons = cases.primes_to_conservatives(dim, gamma, 0.0, rho, vel, p)
    if kind == "zero_rho":
        cons[2:] = 0.0
    if kind == "negative":
        cons[1] = 0.2  # E = 0.2 below the kinetic energy 0.5: p = 0.4 (0.2 - 0.5) < 0
    blk[sl] = cons
    return np.ascontiguousarray(blk)


def canonical(a):
    """Bytes with every NaN replaced by one quiet NaN, for hashing."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    a[np.isnan(a)] = np.nan
    return a.tobytes()
