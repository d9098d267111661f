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


def blobs_shock_3d(nb: int, seed: int, gamma: float = 1.4, nblobs: int = 32,
                   shock_x: float = 0.1, chunk: int = 16):
    """3D uniform background + seeded Gaussian blobs + planar Mach-2 shock.

    Background rho = p = 1 at rest.  Each blob adds A exp(-|x-c|^2 / (2 r^2))
    (periodic distance) to rho and p with c ~ U[0,1]^3, r ~ U[0.02, 0.08],
    A ~ U[0.2, 2].  For x < shock_x the Rankine-Hugoniot post-shock state of a
    Mach-2 shock into the background: rho = 8/3, p = 4.5,
    u = 2 sqrt(gamma) (1 - 3/8).
    """
    n = nb * IC
    dx = 1.0 / n
    c1 = (np.arange(n) + 0.5) * dx
    rng = SplitMix64(0xB10B, seed)
    blobs = []
    for _ in range(nblobs):
        cx, cy, cz = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1)
        r = rng.uniform(0.02, 0.08)
        a = rng.uniform(0.2, 2.0)
        blobs.append((cx, cy, cz, r, a))
    m = 2.0
    rho_post = (gamma + 1.0) * m * m / ((gamma - 1.0) * m * m + 2.0)
    p_post = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m * m - 1.0)
    u_post = m * math.sqrt(gamma) * (1.0 - 1.0 / rho_post)
    cons = np.empty((5, n, n, n), dtype=np.float64)
    for i0 in range(0, n, chunk):
        xs = c1[i0:i0 + chunk]
        X, Y, Z = np.meshgrid(xs, c1, c1, indexing="ij")
        add = np.zeros_like(X)
        for cx, cy, cz, r, a in blobs:
            dxp = np.abs(X - cx)
            dxp = np.minimum(dxp, 1.0 - dxp)
            dyp = np.abs(Y - cy)
            dyp = np.minimum(dyp, 1.0 - dyp)
            dzp = np.abs(Z - cz)
            dzp = np.minimum(dzp, 1.0 - dzp)
            add += a * np.exp(-(dxp * dxp + dyp * dyp + dzp * dzp) / (2.0 * r * r))
        rho = 1.0 + add
        p = 1.0 + add
        u = np.zeros_like(X)
        post = X < shock_x
        rho = np.where(post, rho_post + add, rho)
        p = np.where(post, p_post + add, p)
        u = np.where(post, u_post, u)
        z = np.zeros_like(X)
        cons[:, i0:i0 + chunk] = primes_to_conservatives(3, gamma, 0.0, rho, (u, z, z), p)
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
    """blobs_shock_3d evaluated with torch on a device (same definition; used
    where the domain is too large for a host-side numpy evaluation)."""
    import torch

    n = nb * IC
    dx = 1.0 / n
    c1 = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * dx
    rng = SplitMix64(0xB10B, seed)
    blobs = []
    for _ in range(nblobs):
        cx, cy, cz = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1)
        r = rng.uniform(0.02, 0.08)
        a = rng.uniform(0.2, 2.0)
        blobs.append((cx, cy, cz, r, a))
    m = 2.0
    rho_post = (gamma + 1.0) * m * m / ((gamma - 1.0) * m * m + 2.0)
    p_post = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m * m - 1.0)
    u_post = m * math.sqrt(gamma) * (1.0 - 1.0 / rho_post)
    cons = torch.empty((5, n, n, n), dtype=torch.float64, device=device)
    for i0 in range(0, n, chunk):
        xs = c1[i0:i0 + chunk]
        X, Y, Z = torch.meshgrid(xs, c1, c1, indexing="ij")
        add = torch.zeros_like(X)
        for cx, cy, cz, r, a in blobs:
            dxp = (X - cx).abs()
            dxp = torch.minimum(dxp, 1.0 - dxp)
            dyp = (Y - cy).abs()
            dyp = torch.minimum(dyp, 1.0 - dyp)
            dzp = (Z - cz).abs()
            dzp = torch.minimum(dzp, 1.0 - dzp)
            add += a * torch.exp(-(dxp * dxp + dyp * dyp + dzp * dzp) / (2.0 * r * r))
        post = X < shock_x
        rho = torch.where(post, rho_post + add, 1.0 + add)
        p = torch.where(post, p_post + add, 1.0 + add)
        u = torch.where(post, torch.full_like(X, u_post), torch.zeros_like(X))
        z = torch.zeros_like(X)
        w = z
        s1 = (u * u + z * z) + w * w
        s2 = (u * u + w * w) + z * z
        s3 = (z * z + w * w) + u * u
        q = 0.5 * (torch.minimum(torch.minimum(s1, s2), s3) + torch.maximum(torch.maximum(s1, s2), s3))
        cons[0, i0:i0 + chunk] = rho
        cons[1, i0:i0 + chunk] = p / (gamma - 1.0) + (0.5 * q * rho)
        cons[2, i0:i0 + chunk] = u * rho
        cons[3, i0:i0 + chunk] = z
        cons[4, i0:i0 + chunk] = z
    return cons
