"""Seeded synthetic initial states for the uniform level-0 cases.

ALPACA evaluates prime-state initial conditions at cell centres
``x = origin + (i + 1/2) dx`` (prime_state_initializer.cpp:79) and converts
them to conservatives with the stiffened-gas energy
``E = (p + gamma*B)/(gamma - 1) + 1/2 |u|^2 rho`` (generic_stiffened_gas.h:61-71,
prime_state_handler.h:197-276).  The states themselves are this build's
seeded choices: a four-quadrant 2D Riemann problem and a 3D field of Gaussian
density/pressure blobs with a planar Mach-2 shock.

All arrays are domain-global internal cells ``[NE][NX][NY][NZ]`` (k fastest,
NZ = 1 in 2D) in the equation order Mass, Energy, MomX, MomY, (MomZ).
"""
from __future__ import annotations

import math

import numpy as np

IC = 16
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014) for reproducible draws."""

    def __init__(self, *key: int):
        s = 0x9E3779B97F4A7C15
        for k in key:
            s = (s ^ (int(k) & MASK64)) * 0xBF58476D1CE4E5B9 & MASK64
            s = (s ^ (s >> 31)) & MASK64
        self.state = s

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def _dams(dim, a, b, c):
    """DimensionAwareConsistencyManagedSum (mathematical_functions.h:79-108)."""
    if dim == 2:
        return a + b
    s1 = (a + b) + c
    s2 = (a + c) + b
    s3 = (b + c) + a
    return 0.5 * (np.minimum(np.minimum(s1, s2), s3) + np.maximum(np.maximum(s1, s2), s3))


def conservatives_to_primes(dim, gamma, b, cons):
    """(rho, [u, v, (w)], p) of conservatives cons[NE][...] with the
    ConservativesToPrimeStates arithmetic (prime_state_handler.h:279-370):
    u = m * (1 / rho), stiffened-gas pressure."""
    rho = cons[0]
    one = 1.0 / rho
    vel = [cons[2 + d] * one for d in range(dim)]
    m2 = [cons[2 + d] * cons[2 + d] for d in range(dim)] + [0.0] * (3 - dim)
    q = 0.5 * _dams(dim, m2[0], m2[1], m2[2])
    p = -gamma * b + (gamma - 1.0) * (cons[1] - q / rho)
    return rho, vel, p


def primes_to_conservatives(dim, gamma, b, rho, vel, p):
    """PrimeStatesToConservatives with the stiffened-gas energy."""
    u = vel[0]
    v = vel[1]
    w = vel[2] if dim == 3 else np.zeros_like(u)
    ne = dim + 2
    out = np.empty((ne,) + rho.shape, dtype=np.float64)
    out[0] = rho
    out[1] = (p + gamma * b) / (gamma - 1.0) + (0.5 * _dams(dim, u * u, v * v, w * w) * rho)
    out[2] = u * rho
    out[3] = v * rho
    if dim == 3:
        out[4] = w * rho
    return out


def quadrant_states(seed: int, case: int = 0):
    """Per-quadrant (rho, p, u, v) drawn in order SW, SE, NW, NE."""
    rng = SplitMix64(0xA1CA, seed, case)
    q = []
    for _ in range(4):
        rho = rng.uniform(0.1, 1.5)
        p = rng.uniform(0.05, 1.5)
        u = rng.uniform(-0.8, 0.8)
        v = rng.uniform(-0.8, 0.8)
        q.append((rho, p, u, v))
    return q


def four_quadrant_2d(nbx: int, nby: int, seed: int, gamma: float = 1.4, case: int = 0,
                     length: float = 1.0):
    """2D four-quadrant problem on [0, length]^2 with nbx x nby blocks of 16^2."""
    nx, ny = nbx * IC, nby * IC
    dx = length / nx
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * (length / ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    q = quadrant_states(seed, case)
    east = X >= 0.5 * length
    north = Y >= 0.5 * length
    idx = east.astype(np.int64) + 2 * north.astype(np.int64)
    tab = np.array(q, dtype=np.float64)
    rho = tab[idx, 0]
    p = tab[idx, 1]
    u = tab[idx, 2]
    v = tab[idx, 3]
    cons = primes_to_conservatives(2, gamma, 0.0, rho, (u, v), p)
    return cons[..., None]


_LN2_HI = 6.93147180369123816490e-01  # ln 2 to 32 bits (low bits zero: k*_LN2_HI is exact)
_LN2_LO = 1.90821492927058770002e-10  # ln 2 - _LN2_HI
_INV_LN2 = 1.44269504088896338700e+00
_EXP_TAYLOR = [1.0 / math.factorial(n) for n in range(14)]


def det_exp(x):
    """exp(x) for x <= 0 from IEEE +, -, * and exact scaling only.

    Cody-Waite reduction x = k ln2 + r (|r| <= ln2/2) and the degree-13 Taylor
    polynomial of exp(r) in Horner form (truncation < 1e-17 relative); every
    IEEE-754 machine returns the same bits, unlike library exp routines, so a
    state built from it can be regenerated anywhere and compared by hash.
    """
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x * _INV_LN2)
    r = (x - k * _LN2_HI) - k * _LN2_LO
    p = np.full_like(r, _EXP_TAYLOR[13])
    for c in reversed(_EXP_TAYLOR[:13]):
        p = p * r + c
    kk = np.clip(k, -1100, 0).astype(np.int64)
    return np.where(x < -740.0, 0.0, np.ldexp(p, kk))


def _blob_params(seed: int, nblobs: int):
    rng = SplitMix64(0xB10B, seed)
    blobs = []
    for _ in range(nblobs):
        cx, cy, cz = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1)
        r = rng.uniform(0.02, 0.08)
        a = rng.uniform(0.2, 2.0)
        blobs.append((cx, cy, cz, r, a))
    return blobs


def _blob_profiles(n: int, blobs):
    """Per blob the three 1D factors g(x) = exp(-d(x)^2 / (2 r^2)) with d the
    periodic distance of the cell centre to the blob centre."""
    c1 = (np.arange(n) + 0.5) * (1.0 / n)
    out = []
    for cx, cy, cz, r, a in blobs:
        g = []
        for c in (cx, cy, cz):
            d = np.abs(c1 - c)
            d = np.minimum(d, 1.0 - d)
            g.append(det_exp(-(d * d) / (2.0 * r * r)))
        out.append((a, g[0], g[1], g[2]))
    return c1, out


def _shock_state(gamma: float):
    m = 2.0
    rho_post = (gamma + 1.0) * m * m / ((gamma - 1.0) * m * m + 2.0)
    p_post = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m * m - 1.0)
    u_post = m * math.sqrt(gamma) * (1.0 - 1.0 / rho_post)
    return rho_post, p_post, u_post


def blobs_shock_3d(nb: int, seed: int, gamma: float = 1.4, nblobs: int = 32,
                   shock_x: float = 0.1, chunk: int = 16):
    """3D uniform background + seeded Gaussian blobs + planar Mach-2 shock.

    Background rho = p = 1 at rest.  Each blob adds
    A gx(x) gy(y) gz(z) = A exp(-|x-c|^2 / (2 r^2)) (periodic distance per
    axis) to rho and p with c ~ U[0,1]^3, r ~ U[0.02, 0.08], A ~ U[0.2, 2].
    For x < shock_x the Rankine-Hugoniot post-shock state of a Mach-2 shock
    into the background: rho = 8/3, p = 4.5, u = 2 sqrt(gamma) (1 - 3/8).
    Only IEEE +, -, *, / and det_exp are used, so the state is the same bits
    on every machine (and in blobs_shock_3d_torch).
    """
    n = nb * IC
    c1, prof = _blob_profiles(n, _blob_params(seed, nblobs))
    rho_post, p_post, u_post = _shock_state(gamma)
    cons = np.empty((5, n, n, n), dtype=np.float64)
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        add = np.zeros((i1 - i0, n, n))
        for a, gx, gy, gz in prof:
            add += a * ((gx[i0:i1, None, None] * gy[None, :, None]) * gz[None, None, :])
        post = np.broadcast_to((c1[i0:i1] < shock_x)[:, None, None], add.shape)
        rho = np.where(post, rho_post + add, 1.0 + add)
        p = np.where(post, p_post + add, 1.0 + add)
        u = np.where(post, u_post, 0.0)
        z = np.zeros_like(add)
        cons[:, i0:i1] = primes_to_conservatives(3, gamma, 0.0, rho, (u, z, z), p)
    return cons


def random_block(dim: int, seed: int, gamma: float = 1.4, smooth: bool = False):
    """One haloed block [NE][24][24][24|1] of random positive states."""
    rng = np.random.default_rng(seed)
    shape = (24, 24, 24) if dim == 3 else (24, 24)
    if smooth:
        g = np.meshgrid(*[np.linspace(0, 1, 24)] * dim, indexing="ij")
        phase = sum((k + 1) * gi for k, gi in enumerate(g))
        rho = 1.0 + 0.3 * np.sin(2 * np.pi * phase + rng.uniform(0, 6))
        p = 1.0 + 0.3 * np.cos(2 * np.pi * phase + rng.uniform(0, 6))
        vel = [0.4 * np.sin(2 * np.pi * phase * (d + 1) + rng.uniform(0, 6)) for d in range(dim)]
    else:
        rho = rng.uniform(0.1, 2.0, shape)
        p = rng.uniform(0.05, 2.0, shape)
        vel = [rng.uniform(-1.0, 1.0, shape) for _ in range(dim)]
    cons = primes_to_conservatives(dim, gamma, 0.0, rho, vel, p)
    if dim == 2:
        cons = cons[..., None]
    return np.ascontiguousarray(cons)


def blobs_shock_3d_torch(nb: int, seed: int, gamma: float = 1.4, nblobs: int = 32,
                         shock_x: float = 0.1, device="cuda", chunk: int = 32):
    """blobs_shock_3d evaluated on a device (the 1D factors on the host, the 3D
    products, sums and conversion as tensor-tensor operations with the same
    order); used where the domain is too large for a host-side evaluation."""
    import torch

    n = nb * IC
    c1, prof = _blob_profiles(n, _blob_params(seed, nblobs))
    rho_post, p_post, u_post = _shock_state(gamma)
    dev = [(a, torch.from_numpy(gx).to(device), torch.from_numpy(gy).to(device),
            torch.from_numpy(gz).to(device)) for a, gx, gy, gz in prof]
    xc = torch.from_numpy(c1).to(device)
    cons = torch.empty((5, n, n, n), dtype=torch.float64, device=device)
    gm1 = torch.tensor(gamma - 1.0, dtype=torch.float64, device=device)
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        add = torch.zeros((i1 - i0, n, n), dtype=torch.float64, device=device)
        for a, gx, gy, gz in dev:
            add += a * ((gx[i0:i1, None, None] * gy[None, :, None]) * gz[None, None, :])
        post = (xc[i0:i1] < shock_x)[:, None, None].expand_as(add)
        rho = torch.where(post, rho_post + add, 1.0 + add)
        p = torch.where(post, p_post + add, 1.0 + add)
        u = torch.where(post, torch.full_like(add, u_post), torch.zeros_like(add))
        z = torch.zeros_like(add)
        uu = u * u
        zz = z * z
        s1 = (uu + zz) + zz
        s2 = (uu + zz) + zz
        s3 = (zz + zz) + uu
        q = 0.5 * (torch.minimum(torch.minimum(s1, s2), s3) +
                   torch.maximum(torch.maximum(s1, s2), s3))
        cons[0, i0:i1] = rho
        cons[1, i0:i1] = p / gm1 + (0.5 * q * rho)
        cons[2, i0:i1] = u * rho
        cons[3, i0:i1] = z
        cons[4, i0:i1] = z
    return cons


def batched_cases(ncases: int = 512, seed: int = 0):
    """The batched-ensemble workload: ncases independent 2D four-quadrant
    problems, grid sizes 32^2/64^2/128^2/256^2 cells (2/4/8/16 blocks a side)
    with weights 0.4/0.3/0.2/0.1, gamma ~ U[1.2, 1.67], states as
    four_quadrant_2d (case index i).  Returns [(nb, gamma, state0)]."""
    rng = SplitMix64(0xBA7C, seed)
    specs = []
    for i in range(ncases):
        u = rng.uniform(0.0, 1.0)
        nb = 2 if u < 0.4 else (4 if u < 0.7 else (8 if u < 0.9 else 16))
        gamma = rng.uniform(1.2, 1.67)
        specs.append((nb, gamma, four_quadrant_2d(nb, nb, seed=seed, gamma=gamma, case=i)))
    return specs


def blob_node_initial(dim: int, node_size: float, gamma: float = 1.4):
    """Initial condition of a multi-level node for the periodic blob case: a
    dense, high-pressure blob (radius^2 < 0.02 around (0.4, 0.55, 0.25)) in a
    uniform flow (u, v) = (0.5, 0.3).  Returns initial(level, key) giving the
    conservatives [NE][16][16][16|1] of node key = (level, i, j, k) at its
    cell centres, the form the multi-level driver takes."""
    ic = 16

    def initial(level, key):
        dx = node_size / (1 << level) / ic
        ax = [(key[a + 1] * ic + np.arange(ic) + 0.5) * dx for a in range(dim)]
        g = np.meshgrid(*ax, indexing="ij")
        centre = (0.4, 0.55, 0.25)
        r2 = sum((g[a] - centre[a]) ** 2 for a in range(dim))
        inside = r2 < 0.02
        rho = 1.0 + 2.0 * inside
        p = 1.0 + 4.0 * inside
        vel = [0.5 * np.ones_like(rho), 0.3 * np.ones_like(rho)] + (
            [np.zeros_like(rho)] if dim == 3 else [])
        c = primes_to_conservatives(dim, gamma, 0.0, rho, vel, p)
        return c[..., None] if dim == 2 else c
    return initial

