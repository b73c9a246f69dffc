"""Shared builders for seeded test problems (no GPU needed)."""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    assert nb > 0 and np.isfinite(nb), "reference is zero or not finite: the comparison would be vacuous"
    return np.linalg.norm(a - b) / nb


def random_tt(P, D, nT, R, seed=0, margin=None, spill=0.0):
    """Travel times [P x D] (MATLAB shape): uniform in [margin, nT-1-margin]; with
    spill > 0 a fraction of entries is pushed beyond both ends of the sinogram."""
    g = rng(seed)
    m = R + 1 if margin is None else margin
    tt = g.uniform(m, nT - 1 - m, size=(P, D))
    if spill > 0:
        mask = g.uniform(size=tt.shape) < spill
        tt[mask] = g.uniform(-3 * R - 2, nT + 3 * R + 2, size=mask.sum())
    return tt


def random_sir(P, D, Ls=9, Nsir=7, seed=0, empty=True):
    """SIR table [Ls x Nsir], windows [2 x Nsir] (1-based), indices [P x D] int32."""
    g = rng(seed)
    sir = g.standard_normal((Ls, Nsir))
    se = np.zeros((2, Nsir), dtype=np.int32)
    for k in range(Nsir):
        a = g.integers(1, Ls + 1)
        b = g.integers(a, Ls + 1)
        se[:, k] = (a, b)
    if empty and Nsir > 1:
        se[:, Nsir - 1] = (3, 2)          # empty window (no taps)
    sidx = g.integers(1, Nsir + 1, size=(P, D)).astype(np.int32)
    return sidx, sir, se


def geometry_tt(Nz, Nx, h, D, radius, coverage_deg, c, fs, delay=0.0):
    """Travel times (D, Nx, Nz) of a centred grid and an arc probe (single SoS)."""
    xv = (np.arange(Nx) - (Nx - 1) / 2) * h
    zv = (np.arange(Nz) - (Nz - 1) / 2) * h
    X, Z = np.meshgrid(xv, zv)          # (Nz, Nx)
    off = (180 - coverage_deg) / 2
    ang = (off + np.arange(D) * coverage_deg / (D - 1)) * np.pi / 180
    xd, zd = radius * np.cos(ang), -radius * np.sin(ang)
    dist = np.sqrt((xd[:, None, None] - X.T[None]) ** 2 + (zd[:, None, None] - Z.T[None]) ** 2)
    return dist / c * fs - delay
