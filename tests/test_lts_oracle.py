"""Level-wise integration (TimeIntegrator<RungeKutta3TVD> with micro time
steps): the oracle's restatement (oracle/alpaca_oracle.c ora_lts_*) pinned
against the reference's own RungeKutta3TVD on a Node (harness, and the
hashes in tests/golden/mr_golden.json), the host time-step arithmetic of the
C ABI (alpaca_lts_timestep, host code) against the oracle, and GetLevels'
schedule against what it must satisfy over a macro step."""
import hashlib
import json
import os

import numpy as np
import pytest

from alpaca_amd import lts
from oracle import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "mr_golden.json")) as fh:
    GOLDEN = json.load(fh)

LTS_HALO = {2: ((4, 20, 0), (16, 4, 1)), 3: ((4, 20, 4), (16, 4, 16))}


def _oracle_ops(dim, op, stage, n, micro, start, size, avg, rhs, init, jc, jf):
    """The reference's op sequence restated with the oracle's functions."""
    avg, rhs, jc, jf = avg.copy(), rhs.copy(), jc.copy(), jf.copy()
    if op == 0:
        avg = O.lts_prepare_buffer(dim, avg, init, stage)
        k = 2 * dim
        c, f = O.lts_integrate_jump(dim, jc[:k], jf[:k], O.lts_timestep(micro, n, stage, True))
        jc[:k], jf[:k] = c, f
        rhs = O.lts_integrate(dim, avg, rhs, None, O.lts_timestep(micro, n, stage))
    else:
        rng = (start[0], size[0], start[1], size[1], start[2], size[2])
        rhs = O.lts_integrate(dim, avg, rhs, rng, O.lts_timestep(micro, n, stage))
    return [avg, rhs, init, jc, jf]


def _inputs(dim, op, stage):
    tz, icz = (24, 16) if dim == 3 else (1, 1)
    rng = np.random.default_rng(1000 * dim + 10 * op + stage)
    micro = rng.uniform(1e-4, 1e-3, 8)
    bufs = [rng.standard_normal((dim + 2, 24, 24, tz)) for _ in range(3)]
    bufs += [rng.standard_normal((6, dim + 2, 16, icz)) for _ in range(2)]
    return micro, bufs


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_oracle_matches_reference_fixture(dim, op, stage):
    micro, bufs = _inputs(dim, op, stage)
    start, size = LTS_HALO[dim]
    out = _oracle_ops(dim, op, stage, 4, micro, start, size, *bufs)
    got = hashlib.sha256(b"".join(o.tobytes() for o in out)).hexdigest()
    assert got == GOLDEN["lts"][f"{dim}d_op{op}_stage{stage}"]


@pytest.mark.skipif(not (O.ref_available(2) and O.ref_available(3)),
                    reason="reference harness not built (oracle/_ref)")
@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_matches_reference_harness(dim):
    rng = np.random.default_rng(77 + dim)
    tz, icz = (24, 16) if dim == 3 else (1, 1)
    for op in (0, 1):
        for stage in range(3):
            for n in (1, 2, 8):
                micro = rng.uniform(1e-5, 1e-2, 8)
                bufs = [rng.standard_normal((dim + 2, 24, 24, tz)) for _ in range(3)]
                bufs += [rng.standard_normal((6, dim + 2, 16, icz)) for _ in range(2)]
                start = (0, 4, 4 if dim == 3 else 0)
                size = (4, 16, 16 if dim == 3 else 1)
                want = O.ref_lts_integrate(dim, op, stage, n, micro, start, size, *bufs)
                got = _oracle_ops(dim, op, stage, n, micro, start, size, *bufs)
                for w, g in zip(want, got):
                    assert np.array_equal(w, g), (op, stage, n)


def test_timestep_sums_newest_first_from_zero():
    micro = [0.1, 1e-17, 0.3, 1e-17, 1e-17]
    # (((0 + 1e-17) + 1e-17) + 0.3) + ...: the tiny newest steps add first
    s = 0.0
    for dt in reversed(micro[-4:]):
        s = s + dt
    assert O.lts_timestep(micro, 4, 0) == s
    assert O.lts_timestep(micro, 4, 1) == s * 0.25
    assert O.lts_timestep(micro, 4, 2) == s * (2.0 / 3.0)
    assert O.lts_timestep(micro, 4, 0, True) == s * (1.0 / 6.0)
    assert O.lts_timestep(micro, 1, 2, True) == 1e-17 * (2.0 / 3.0)


def test_host_timestep_entry_matches_oracle():
    """alpaca_lts_timestep is host arithmetic in the C ABI (no GPU needed)."""
    rng = np.random.default_rng(3)
    ti = lts.LevelTimeIntegrator(3)
    for dt in rng.uniform(1e-4, 1e-2, 8):
        ti.append_micro_timestep(dt)
    for n in (1, 2, 4, 8):
        for stage in range(3):
            for jump in (False, True):
                assert ti.timestep(stage, n, jump) == \
                    O.lts_timestep(ti.micro_timestep_sizes, n, stage, jump)
    total = sum(ti.micro_timestep_sizes)
    ti.finish_macro_timestep()
    assert ti.micro_timestep_sizes == [] and abs(ti.current_run_time - total) < 1e-15


@pytest.mark.parametrize("maximum_level", [0, 1, 2, 3, 4, 6])
def test_level_schedule_over_a_macro_step(maximum_level):
    """Over the 2^L micro steps of a macro step the finest level advances
    every micro step and level l every 2^(L-l)-th one (after micro steps
    2^(L-l) - 1, 2 2^(L-l) - 1, ...), so every level reaches the macro time
    exactly at the last micro step, when all levels advance."""
    L = maximum_level
    counts = [0] * (L + 1)
    for t in range(2 ** L):
        levels = lts.levels_to_update(t, L)
        assert levels == sorted(levels, reverse=True) and levels[0] == L
        for level in levels:
            assert (t + 1) % (2 ** (L - level)) == 0
            counts[level] += 1
    assert counts == [2 ** level for level in range(L + 1)]
    assert lts.levels_to_update(2 ** L - 1, L) == list(range(L, -1, -1))


def _surf(dim, rng):
    icz = 16 if dim == 3 else 1
    return rng.standard_normal((6, dim + 2, 16, icz))


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_exchange_closed_forms(dim):
    """One jump face (West): its boundary layer becomes exactly
    (u - J_leaf*s) + J_nb[East]*s (CMS of one non-zero term is that term), all
    other internal cells u - 0 + 0, the halos untouched, and the leaf's West
    surface takes the neighbour's East one."""
    rng = np.random.default_rng(5 + dim)
    tz = 24 if dim == 3 else 1
    cells = rng.standard_normal((dim + 2, 24, 24, tz))
    cells[0, 10, 10, 0 if dim == 2 else 10] = -0.0
    jl, jn = _surf(dim, rng), _surf(dim, rng)
    s = 64.0
    nb = [None, jn, None, None, None, None]
    out, jout = O.lts_jump_flux_adjustment(dim, cells, jl, nb, s)
    zs = slice(4, 20) if dim == 3 else slice(0, 1)
    layer = cells[:, 4, 4:20, zs]
    want = (layer - jl[1] * s * 1.0) + jn[0] * s * 1.0
    assert np.array_equal(out[:, 4, 4:20, zs], want)
    inner = (slice(None), slice(5, 20), slice(4, 20), zs)
    assert np.array_equal(out[inner], (cells[inner] - 0.0) + 0.0)
    assert not np.signbit(out[0, 10, 10, 0 if dim == 2 else 10])  # -0 became +0
    halo = np.ones(cells.shape, dtype=bool)
    halo[:, 4:20, 4:20, zs] = False
    assert np.array_equal(out[halo], cells[halo])
    assert np.array_equal(jout[1], jn[0])
    assert np.array_equal(np.delete(jout, 1, axis=0), np.delete(jl, 1, axis=0))


@pytest.mark.parametrize("dim", [2, 3])
def test_jump_exchange_conserves(dim):
    """Every face a jump: the leaf's internal total changes by the net
    exchanged flux, sum over faces of (J_nb[opp] - J_leaf) * s * dir (to
    rounding); a sign or face mix-up breaks it."""
    rng = np.random.default_rng(11 + dim)
    tz = 24 if dim == 3 else 1
    cells = rng.uniform(0.5, 1.5, (dim + 2, 24, 24, tz))
    jl = _surf(dim, rng)
    nbs = [_surf(dim, rng) for _ in range(2 * dim)] + [None] * (6 - 2 * dim)
    s = 32.0
    out, _ = O.lts_jump_flux_adjustment(dim, cells, jl, nbs, s)
    zs = slice(4, 20) if dim == 3 else slice(0, 1)
    delta = (out - cells)[:, 4:20, 4:20, zs].reshape(dim + 2, -1).sum(axis=1)
    want = np.zeros(dim + 2)
    for loc in range(2 * dim):
        d = -1.0 if loc % 2 == 0 else 1.0
        opp = loc + 1 if loc % 2 == 0 else loc - 1
        want += ((nbs[loc][opp] - jl[loc]) * s * d).reshape(dim + 2, -1).sum(axis=1)
    assert np.allclose(delta, want, rtol=1e-12, atol=1e-9)
