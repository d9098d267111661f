"""The exact form's FP64 division and square root against x86's correctly
rounded ones on constructed hard cases (results within 2^-100 of a rounding
midpoint) and on random operands.  The reference's CPU build computes a / b
and sqrt(x) correctly rounded, so the exact form can reproduce it bit for
bit only if these agree everywhere the kernels divide."""
import ctypes

import numpy as np
import pytest

from alpaca_amd import gpu
import fp64_hard_cases as hard

pytestmark = pytest.mark.gpu

OP_DIV, OP_WENO_DIV, OP_SQRT = 0, 1, 2
N = 20000


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    gpu.load_library()
    return _t


def _run(torch, op, a, b=None):
    lib = gpu.load_library()
    da = torch.as_tensor(a, device="cuda")
    db = torch.as_tensor(b if b is not None else a, device="cuda")
    out = torch.empty_like(da)
    gpu.check(lib.alpaca_exact_fp64_ops(op, ctypes.c_void_p(da.data_ptr()),
                                        ctypes.c_void_p(db.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr()), len(a),
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
              "exact_fp64_ops")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _mismatch(got, want):
    return int(np.count_nonzero(got.view(np.int64) != want.view(np.int64)))


@pytest.mark.parametrize("op,recip_hard", [(OP_DIV, False), (OP_DIV, True), (OP_WENO_DIV, False)])
def test_division_hard_cases_correctly_rounded(torch, op, recip_hard):
    """Every division of state-derived quantities rounds as x86 does, also
    for divisors just below a power of two.  The WENO weight divisions keep
    the hardware sequence without that check (their divisors are smoothness
    indicators): they are pinned on near-midpoint quotients of other divisors."""
    print(f"[fp64 div] op={op} recip_hard={recip_hard}", flush=True)
    # the WENO weight division sees divisors in [1e-12, 2^700): keep exponents small
    a, b = hard.division_cases(N, seed=11 + op + 2 * recip_hard, recip_hard=recip_hard,
                               exp_range=(-30, 30))
    got = _run(torch, op, a, b)
    assert _mismatch(got, a / b) == 0


def test_division_random_and_decimal_operands(torch):
    """Random doubles, and quotients of short decimal numbers (the shape of
    input-file data), over the division of physical quantities."""
    print("[fp64 div] random", flush=True)
    rng = np.random.default_rng(5)
    a = rng.standard_normal(N) * np.exp(rng.uniform(-20, 20, N))
    b = rng.standard_normal(N) * np.exp(rng.uniform(-20, 20, N))
    da = np.round(rng.uniform(0.0, 10.0, N), 3)
    db = np.round(rng.uniform(0.001, 10.0, N), 3)
    # the Sod enthalpy quotient that once rounded the other way
    sod_a = np.array([float.fromhex("0x1.6666666666663p-2")])
    sod_b = np.array([float.fromhex("0x1.ffffffffffffbp-4")])
    A = np.concatenate([a, da, sod_a])
    B = np.concatenate([b, db, sod_b])
    got = _run(torch, OP_DIV, A, B)
    assert _mismatch(got, A / B) == 0
    assert got[-1] == 2.8000000000000003


def test_sqrt_hard_and_random_cases(torch):
    print("[fp64 sqrt]", flush=True)
    rng = np.random.default_rng(9)
    x = np.concatenate([hard.sqrt_cases(N, seed=13),
                        np.exp(rng.uniform(-600, 600, N)),
                        np.round(rng.uniform(0.0, 10.0, N), 4)])
    got = _run(torch, OP_SQRT, x)
    assert _mismatch(got, np.sqrt(x)) == 0
