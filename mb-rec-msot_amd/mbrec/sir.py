"""Spatial impulse response (SIR / "SPMR") tables in the reference's polar layout.

The reference computes its polar SPMR table with Field II
(is_tir_derivation/calculate_polarSPMR.m), a closed binary that is not used
here.  This module generates the table from the element geometry instead:
the element's in-plane width is split into ``numSubElements`` point
sub-elements, each contributes a unit hat at its arrival-time offset relative
to the element centre, and the result is convolved with the pixel-model
N-shape phi(c t) * (-t) (the pixel response that the no-SIR model folds into
its IR, MSOTmodel.m:338-340).

Table layout and indexing follow MSOTmodel.m:463-616:
* distances / azimuth angles of every (pixel, detector) pair (:573-597),
* table grid from calculate_required_SPMR_distances_and_angels (:599-616),
* SPMR [Ls x N_angle*N_dist] (returned C-order as (Nsir, Ls)),
* SPMRTimestepIndexStartAndEnd [2 x Nsir] 1-based inclusive windows,
* SPMRIndices int32 [P x D] 1-based: (round(1+(dist-dmin)/dd)-1)*N_angle + 1+round(angle_deg/da).
"""
from __future__ import annotations

import numpy as np

from .discretization import matlab_round

SPMR_LENGTH = 91   # 2*45+1 taps before trimming (calculate_polarSPMR.m:44)


def polar_coordinates(xg, zg, xd, zd):
    """Distances and azimuth angles from every detector to every pixel (single SoS branch,
    MSOTmodel.m:579-596).  xg, zg: pixel coordinates flattened in pixel order (P,);
    xd, zd: detector coordinates (D,).  Returns (D, P) arrays."""
    dx = xd[:, None] - xg[None, :]
    dz = zd[:, None] - zg[None, :]
    dist = np.sqrt(dx * dx + dz * dz)
    rq = np.sqrt(xd * xd + zd * zd)[:, None]
    cosang = (dx * xd[:, None] + dz * zd[:, None]) / (rq * dist)
    ang = np.arccos(np.clip(cosang, -1.0, 1.0))
    return dist, ang


def required_grid(dist, ang, resolution_fov, c):
    """MSOTmodel.m:599-616."""
    dd = resolution_fov * (1450.0 / c)
    dmin = dist.min() - dd
    dmax = dist.max() + dd
    da = np.arctan2(dd, dmax) / np.pi * 180.0
    amax = np.ceil((np.abs(ang).max() / np.pi * 180.0) * da) / da
    return dmin, dmax, dd, amax, da


def spmr_table(dists, angles_deg, width, n_sub, c, fs, pixel_model, Ls=SPMR_LENGTH,
               trim=1e-7):
    """SPMR for every (distance, angle) grid point: returns (N_dist*N_angle, Ls) with the
    angle index fastest, and the 1-based inclusive windows (Nsir, 2)."""
    nd, na = len(dists), len(angles_deg)
    rho = np.repeat(dists, na)                     # distance slow, angle fast
    alpha = np.tile(angles_deg, nd) * np.pi / 180.0
    e = (np.arange(n_sub) + 0.5) * width / n_sub - width / 2.0
    lat = np.zeros((nd * na, Ls))
    centre = Ls // 2
    sa = np.sin(alpha)
    for ei in e:
        di = np.sqrt(rho * rho + ei * ei - 2.0 * rho * ei * sa)
        dt = (di - rho) / c * fs + centre
        lo = np.floor(dt).astype(np.int64)
        fr = dt - lo
        ok = (lo >= 0) & (lo + 1 < Ls)
        rows = np.nonzero(ok)[0]
        np.add.at(lat, (rows, lo[ok]), (1.0 - fr[ok]) / n_sub)
        np.add.at(lat, (rows, lo[ok] + 1), fr[ok] / n_sub)
    # pixel-model N-shape, support where it is non-negligible
    kn = np.arange(-centre, centre + 1)
    t = kn / fs
    nsh = pixel_model(c * t) * (-t)
    keep = np.abs(nsh) > 1e-12 * np.abs(nsh).max()
    idx = np.nonzero(keep)[0]
    nsh = nsh[idx[0]:idx[-1] + 1]
    off = kn[idx[0]]
    out = np.zeros_like(lat)
    for j, h in enumerate(nsh):
        sh = j + off
        if sh >= 0:
            out[:, sh:] += h * lat[:, :Ls - sh]
        else:
            out[:, :Ls + sh] += h * lat[:, -sh:]
    mx = np.abs(out).max()
    mask = np.abs(out) > trim * mx
    any_ = mask.any(axis=1)
    first = np.where(any_, mask.argmax(axis=1), 0)
    last = np.where(any_, Ls - 1 - mask[:, ::-1].argmax(axis=1), -1)
    se = np.stack([first + 1, last + 1], axis=1).astype(np.int32)
    return out, se


def spmr_indices(dist, ang, dmin, dd, da, n_angle):
    """SPMR index of every pair (MSOTmodel.m:472-478), 1-based: distance slot
    round(1 + (dist - dmin) / dd), angle slot 1 + round(angle_deg / da) (MATLAB round, half
    away from zero), linearised angle-fastest over n_angle angles."""
    i_d = matlab_round(1 + (np.asarray(dist) - dmin) / dd)
    i_a = 1 + matlab_round(np.asarray(ang) * 180.0 / np.pi / da)
    return ((i_d - 1) * n_angle + i_a).astype(np.int32)


def sir_tables_for_model(xg, zg, xd, zd, resolution_fov, c, fs, width, n_sub, pixel_model,
                         Ls=SPMR_LENGTH):
    """get_SPMR_and_SPMR_indices (MSOTmodel.m:463-480) with a generated table.
    Returns SPMR (Nsir, Ls), windows (Nsir, 2) int32, indices (D, P) int32 (1-based)."""
    dist, ang = polar_coordinates(xg, zg, xd, zd)
    dmin, dmax, dd, amax, da = required_grid(dist, ang, resolution_fov, c)
    dists = np.arange(dmin, dmax + 0.5 * dd, dd)
    angles = np.arange(0.0, amax + 0.5 * da, da)
    table, se = spmr_table(dists, angles, width, n_sub, c, fs, pixel_model, Ls)
    k = spmr_indices(dist, ang, dmin, dd, da, len(angles))
    return table, se, k, dict(dmin=dmin, dd=dd, da=da, n_angle=len(angles), n_dist=len(dists))
