"""The oracle on the special branches of the face numerics (CPU).

tests/edge_states.py builds blocks that reach the face-skip conditions
(near-vacuum and negative-pressure reconstructions), S* = 0, the Roe
eigensystem's non-positive-density skip, and square roots and WENO5
divisions in the range where the hardware sequences scale.  The oracle's
per-block fluxes and right-hand sides must hash like the reference's own code
did on the same blocks (golden/edge_blocks.json, from make_golden.py --edge)
and, where oracle/_ref is built, equal it bit for bit with NaNs in the same
places.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import edge_states
import oracle_py

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "edge_blocks.json")
CASES = [(dim, flux, kind) for dim in (2, 3) for flux in (0, 1) for kind in edge_states.KINDS]


def _key(dim, flux, kind):
    return f"{dim}d_{'roe' if flux else 'hllc'}_{kind}"


@pytest.mark.parametrize("dim,flux,kind", CASES)
def test_oracle_matches_reference_on_edge_blocks(dim, flux, kind):
    meta = json.load(open(GOLDEN))[_key(dim, flux, kind)]
    blk = edge_states.edge_block(dim, kind, seed=3)
    with np.errstate(all="ignore"):
        fl = oracle_py.block_fluxes(dim, flux, 1.4, 0.0, blk, 1.0 / 128)
        rhs = oracle_py.block_rhs(dim, flux, 1.4, 0.0, blk, 1.0 / 128)
    assert hashlib.sha256(edge_states.canonical(fl)).hexdigest() == meta["fluxes_sha256"]
    assert hashlib.sha256(edge_states.canonical(rhs)).hexdigest() == meta["rhs_sha256"]
    assert int(np.isnan(fl).sum()) == meta["nan_fluxes"]
    if oracle_py.ref_available(dim):
        with np.errstate(all="ignore"):
            rf = oracle_py.block_fluxes(dim, flux, 1.4, 0.0, blk, 1.0 / 128, which="ref")
        assert np.array_equal(fl, rf, equal_nan=True)


def test_edge_blocks_reach_their_branches():
    """Each kind changes the result the way its branch says."""
    with np.errstate(all="ignore"):
        # vacuum (HLLC): faces next to the slab are skipped -- all equations zero
        fl = oracle_py.block_fluxes(3, 0, 1.4, 0.0, edge_states.edge_block(3, "vacuum", 3), 1 / 128)
        # x-faces between haloed cells a+3 and a+4: the slab holds cells 10..12,
        # so exactly the face planes 7 and 8 (inside it) are skipped entirely
        planes = np.all(fl[0] == 0.0, axis=0).reshape(17, -1).all(axis=1)
        assert planes.tolist() == [a in (7, 8) for a in range(17)]
        # rest: no mass moves through any face, and the momentum flux is the pressure
        fl = oracle_py.block_fluxes(2, 0, 1.4, 0.0, edge_states.edge_block(2, "rest", 3), 1 / 128)
        assert np.all(fl[:2, 0] == 0.0)
        # face arrays [ANOE][IC+1][IC+1][ICZ+1]: transverse entries 1..16 (2D: k = 1)
        assert np.allclose(fl[0, 2, :, 1:, 1], 0.7, rtol=1e-14, atol=0)   # x-faces: momentum-x flux = p
        assert np.allclose(fl[1, 3, 1:, :, 1], 0.7, rtol=1e-14, atol=0)   # y-faces: momentum-y flux = p
        # huge: finite fluxes although the smoothness indicators pass 2^700
        fl = oracle_py.block_fluxes(3, 1, 1.4, 0.0, edge_states.edge_block(3, "huge", 3), 1 / 128)
        assert np.isfinite(fl).all() and np.abs(fl).max() > 2.0 ** 340
