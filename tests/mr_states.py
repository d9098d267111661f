"""Parent/child block pairs for the multiresolution operator tests.

  smooth     parent and child sample one smooth state at their own cell
             centres (the child covering the parent's octant at `position`):
             small details, so the remeshing decision depends on the threshold
  random     independent random states (large details)
  zeros      smooth, with a slab of exact zeros in Mass and Energy of both
             blocks: |0 - 0| / |0| is NaN (dropped by std::max) and
             |0 - p| / |0| is +inf
  nonfinite  smooth, with +inf, -inf and NaN cells in the parent: the
             prediction carries them into the child's cells

Used by the oracle-against-reference tests (CPU) and the GPU-against-oracle
tests.
"""
import numpy as np

from alpaca_amd import cases

KINDS = ("smooth", "random", "zeros", "nonfinite")
IC, HS, TC = 16, 4, 24


def _centres(n_cells, h, origin):
    return origin + (np.arange(n_cells) - HS + 0.5) * h


def _smooth_state(dim, coords, phase0):
    g = np.meshgrid(*coords, indexing="ij")
    s = sum((k + 1.3) * gi for k, gi in enumerate(g))
    rho = 1.0 + 0.2 * np.sin(2.0 * s + phase0)
    p = 1.0 + 0.2 * np.cos(1.7 * s + phase0)
    vel = [0.3 * np.sin(1.1 * (d + 1) * s + phase0) for d in range(dim)]
    cons = cases.primes_to_conservatives(dim, 1.4, 0.0, rho, vel, p)
    if dim == 2:
        cons = cons[..., None]
    return np.ascontiguousarray(cons)


def pair(dim, kind, position, seed=0):
    """(parent, child) haloed blocks [NE][24][24][24|1]."""
    rng = np.random.default_rng(seed * 131 + position)
    if kind == "random":
        return (cases.random_block(dim, seed * 7 + 1 + position),
                cases.random_block(dim, seed * 7 + 100 + position))
    phase0 = rng.uniform(0.0, 6.0)
    h = 1.0 / IC
    parent = _smooth_state(dim, [_centres(TC, h, 0.0) for _ in range(dim)], phase0)
    origin = [0.5 * ((position >> a) & 1) for a in range(dim)]
    child = _smooth_state(dim, [_centres(TC, h / 2, origin[a]) for a in range(dim)], phase0)
    if kind == "zeros":
        parent[:2, 6:10] = 0.0
        child[:2, 2:9] = 0.0
    elif kind == "nonfinite":
        parent[0, 8, 9] = np.inf
        parent[1, 12, 11] = -np.inf
        parent[2, 14, 7] = np.nan
    elif kind != "smooth":
        raise ValueError(kind)
    return np.ascontiguousarray(parent), np.ascontiguousarray(child)


def ranges(dim):
    """Prediction ranges (x_start, x_count, y_start, y_count, z_start,
    z_count): the whole block, the internal cells (refinement), the -x and +y
    halo slabs (a level jump) and an odd count (floored to even)."""
    tz = TC if dim == 3 else 1
    z = (0, tz)
    zi = (HS, IC) if dim == 3 else (0, 1)
    return [
        (0, TC, 0, TC) + z,
        (HS, IC, HS, IC) + zi,
        (0, HS, 0, TC) + z,
        (0, TC, TC - HS, HS) + z,
        (2, 7, 6, 10) + ((4, 9) if dim == 3 else (0, 1)),
    ]
