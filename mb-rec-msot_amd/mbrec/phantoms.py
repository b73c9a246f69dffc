"""Seeded synthetic phantoms and noise for the benchmark configurations.

All randomness comes from numpy.random.Generator(PCG64(seed)).  Images are
returned as (Nx, Nz) arrays (MATLAB [Nz x Nx] memory order).
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _grid(disc):
    return disc.region["xPositionsFov"].T, disc.region["zPositionsFov"].T   # (Nx, Nz)


def _box(X, Z, xlo, xhi, zlo, zhi):
    """Index slices (along x = axis 0, z = axis 1) of the grid points inside a box: shapes
    are evaluated there only (the same values as on the whole grid, zero outside)."""
    xv, zv = X[:, 0], Z[0, :]
    i0, i1 = np.searchsorted(xv, xlo, "left"), np.searchsorted(xv, xhi, "right")
    j0, j1 = np.searchsorted(zv, zlo, "left"), np.searchsorted(zv, zhi, "right")
    return slice(max(i0 - 1, 0), i1 + 1), slice(max(j0 - 1, 0), j1 + 1)


def disks(disc, n: int = 20, rmin: float = 0.2e-3, rmax: float = 3e-3, amin: float = 0.1,
          amax: float = 1.0, seed: int = 0) -> np.ndarray:
    """Sum of n disks, centres uniform in the field of view, radii U(rmin, rmax),
    amplitudes U(amin, amax)."""
    g = rng(seed)
    X, Z = _grid(disc)
    x0, x1, z0, z1 = X.min(), X.max(), Z.min(), Z.max()
    img = np.zeros(X.shape)
    for _ in range(n):
        cx, cz = g.uniform(x0, x1), g.uniform(z0, z1)
        r = g.uniform(rmin, rmax)
        a = g.uniform(amin, amax)
        bi, bj = _box(X, Z, cx - r, cx + r, cz - r, cz + r)
        img[bi, bj] += a * (((X[bi, bj] - cx) ** 2 + (Z[bi, bj] - cz) ** 2) <= r * r)
    return img


def ellipses(disc, n: int = 10, seed: int = 0) -> np.ndarray:
    g = rng(seed)
    X, Z = _grid(disc)
    x0, x1, z0, z1 = X.min(), X.max(), Z.min(), Z.max()
    img = np.zeros(X.shape)
    for _ in range(n):
        cx, cz = g.uniform(x0, x1), g.uniform(z0, z1)
        a, b = g.uniform(0.3e-3, 4e-3, size=2)
        th = g.uniform(0, np.pi)
        amp = g.uniform(0.1, 1.0)
        m = max(a, b)
        bi, bj = _box(X, Z, cx - m, cx + m, cz - m, cz + m)
        Xb, Zb = X[bi, bj], Z[bi, bj]
        xr = (Xb - cx) * np.cos(th) + (Zb - cz) * np.sin(th)
        zr = -(Xb - cx) * np.sin(th) + (Zb - cz) * np.cos(th)
        img[bi, bj] += amp * ((xr / a) ** 2 + (zr / b) ** 2 <= 1.0)
    return img


def vessel_tree(disc, depth: int = 5, rmin: float = 0.1e-3, rmax: float = 1e-3, seed: int = 0) -> np.ndarray:
    """Seeded branching tubes: each segment splits into two with shrinking radius."""
    g = rng(seed)
    X, Z = _grid(disc)
    x0, x1, z0, z1 = X.min(), X.max(), Z.min(), Z.max()
    span = min(x1 - x0, z1 - z0)
    img = np.zeros(X.shape)

    def seg(px, pz, qx, qz, r, amp):
        vx, vz = qx - px, qz - pz
        L2 = vx * vx + vz * vz + 1e-30
        bi, bj = _box(X, Z, min(px, qx) - r, max(px, qx) + r, min(pz, qz) - r, max(pz, qz) + r)
        Xb, Zb = X[bi, bj], Z[bi, bj]
        t = np.clip(((Xb - px) * vx + (Zb - pz) * vz) / L2, 0.0, 1.0)
        d2 = (Xb - px - t * vx) ** 2 + (Zb - pz - t * vz) ** 2
        np.maximum(img[bi, bj], amp * (d2 <= r * r), out=img[bi, bj])

    stack = [(g.uniform(x0, x1), z0 + 0.05 * span, g.uniform(-0.3, 0.3) + np.pi / 2, rmax, 0, 1.0)]
    while stack:
        px, pz, ang, r, lvl, amp = stack.pop()
        L = span * g.uniform(0.15, 0.3)
        qx, qz = px + L * np.cos(ang), pz + L * np.sin(ang)
        seg(px, pz, qx, qz, r, amp)
        if lvl + 1 < depth and r > rmin:
            for s in (-1, 1):
                stack.append((qx, qz, ang + s * g.uniform(0.3, 0.8), max(rmin, r * g.uniform(0.6, 0.85)),
                              lvl + 1, amp * g.uniform(0.7, 1.0)))
    return img


def add_noise(sino: np.ndarray, level: float = 0.01, seed: int = 0) -> np.ndarray:
    """Gaussian noise with standard deviation level * max|sino|."""
    g = rng(seed)
    return sino + level * np.abs(sino).max() * g.standard_normal(sino.shape)


def smooth_spectrum(n: int, seed: int = 0) -> np.ndarray:
    """Seeded smooth positive amplitude spectrum over n wavelengths."""
    g = rng(seed)
    k = np.arange(n) / max(1, n - 1)
    c = g.uniform(0.2, 1.0, size=4)
    s = c[0] + c[1] * np.cos(np.pi * k) + c[2] * np.sin(2 * np.pi * k * c[3])
    s = s - s.min() + 0.2
    return s / s.max()
