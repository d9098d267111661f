"""Constructed hard cases of FP64 division and square root.

A quotient or square root that lies extremely close to a rounding midpoint
(half-way between two doubles) is where a Newton-Raphson sequence without a
final exact decision can round the wrong way.  These generators build such
inputs with integer arithmetic:

* division: a = A, b = B (53-bit integers, then scaled by powers of two) with
  A * 2^54 = M * B - t, M odd in [2^53, 2^54) and |t| small, so that a / b is
  within |t| / (B * 2^54) of the midpoint M * 2^-54;
  ``recip_hard`` draws B = 2^53 - k (k odd and small), whose reciprocal is
  itself within k^2 * 2^-107 of a midpoint: the pairs that defeat the
  hardware sequence's reciprocal, and the shape of the densities just below a
  power of two that a Sod-type run produces (0.125 after round-off);
* square root: x = K * 2^55 with K * 2^55 = M^2 - t (M odd, 54 bits, t = 1
  mod 8 and |t| < 2^13), so that sqrt(x) is within about |t| / (2M) of the
  midpoint M.

numpy's float64 division and square root on x86 are IEEE correctly rounded:
the reference's build computes exactly those.
"""
import numpy as np

_B53 = 1 << 53
_B54 = 1 << 54
_B55 = 1 << 55


def _divisor(rng, recip_hard):
    if recip_hard:
        return _B53 - int(2 * rng.integers(0, 64) + 1)
    return int(rng.integers(1 << 52, _B53)) | 1


def division_cases(n, seed=0, recip_hard=False, exp_range=(-40, 40)):
    """n pairs (a, b) whose quotient is within 2^-100 (relative) of a midpoint."""
    rng = np.random.default_rng(seed)
    a_out, b_out = [], []
    while len(a_out) < n:
        B = _divisor(rng, recip_hard)
        inv = pow(B, -1, _B54)
        t = int(rng.integers(-16, 16)) * 2 + 1
        M = (t * inv) % _B54
        if M < _B53:
            continue
        num = M * B - t
        assert num % _B54 == 0
        A = num >> 54
        if not ((1 << 52) <= A < _B53):
            continue
        ea = int(rng.integers(*exp_range))
        eb = int(rng.integers(*exp_range))
        a_out.append(np.ldexp(float(A), ea - 52))
        b_out.append(np.ldexp(float(B), eb - 52))
    return np.array(a_out), np.array(b_out)


def _sqrt_mod(t, bits):
    """A root m of m^2 = t (mod 2^bits), t = 1 (mod 8)."""
    m = 1
    for k in range(3, bits):
        mod = 1 << (k + 1)
        if (m * m - t) % mod:
            m += 1 << (k - 1)
    return m


def sqrt_cases(n, seed=0, exp_range=(-60, 60)):
    """n arguments x whose square root is within 2^-100 (relative) of a midpoint."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        # the roots of m^2 = t (mod 2^55) are +-m and +-m + 2^54: four per t
        t = 1 + 8 * int(rng.integers(-1024, 1024))
        m = _sqrt_mod(t % _B55, 55)
        if rng.integers(0, 2):
            m = (-m) % _B55
        M = m % _B54
        if M < _B53 or (M * M - t) % _B55:
            continue
        K = (M * M - t) >> 55
        if not ((1 << 52) <= K < _B53):
            continue
        # x = (M^2 - t) 2^(2j - 108): sqrt(x) = sqrt(M^2 - t) 2^(j - 54)
        e = 2 * int(rng.integers(*exp_range))
        out.append(np.ldexp(float(K), 55 - 108 + e))
    return np.array(out)


def midpoint_distance_ulps(exact_num, exact_den, q):
    """Distance of the exact quotient from the nearest midpoint, in ulps of q
    (pure-Python rational arithmetic; used to check the generators)."""
    from fractions import Fraction
    x = Fraction(exact_num) / Fraction(exact_den)
    m, e = np.frexp(q)
    ulp = Fraction(2) ** (int(e) - 53)
    lo = Fraction(float(q)) - ulp / 2
    hi = Fraction(float(q)) + ulp / 2
    return float(min(abs(x - lo), abs(x - hi)) / ulp)
