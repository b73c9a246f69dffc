"""Spatial impulse response (SIR / "SPMR") tables in the reference's polar layout, on the
device, with the polarSPMR file.

The reference computes its polar SPMR table with Field II
(is_tir_derivation/calculate_polarSPMR.m), a closed binary that is not used here.  The
generator (mbrec_spmr_table_dev) splits the element's in-plane width into
``numSubElements`` point sub-elements, deposits each one's arrival-time offset relative to
the element centre as a linear hat, convolves with the pixel-model N-shape
phi(c t) * (-t) (the pixel response the no-SIR model folds into its IR,
MSOTmodel.m:338-340) and trims every entry like calculate_polarSPMR.m:152-161 (samples
below 1e-4 of the entry's largest set to zero, window = first / last non-zero sample).

Table layout, grid and indexing follow MSOTmodel.m:463-616 (mbrec_spmr_grid_dev):
* distances / azimuth angles of every (pixel, detector) pair (:573-597),
* table grid from calculate_required_SPMR_distances_and_angels (:599-616), axes as
  calculate_polarSPMR.m:22-25 lays them out,
* SPMR [Ls x N_angle*N_dist] (returned C-order as (Nsir, Ls)),
* SPMRTimestepIndexStartAndEnd [2 x Nsir] 1-based inclusive windows,
* SPMRIndices int32 [P x D] 1-based: (round(1+(dist-dmin)/dd)-1)*N_angle + 1+round(angle_deg/da).

polarSPMR file (MSOTmodel.m:482-571, calculate_polarSPMR.m:188-193): with ``spmr_dir`` the
table is read from polarSPMR_<probe>_c_<c>_pixelModel_<hash>.mat when that file covers the
required grid (same version and pixel model, min distance <=, max distance >=, steps <=,
max angle >=, tolerance 1e-10), its own grid then giving the indices; otherwise it is
computed over the union of the file's and the required grid and written there (MAT v5:
SPMR_polar [1 x N_angle x N_dist x Ls], SPMR_timestep_index_start_and_end
[1 x N_angle x N_dist x 2] int32, distance_mat_SPMR, angle_azimuth_mat_SPMR, ...).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from .discretization import matlab_round

SPMR_LENGTH = 91            # 2*45+1 taps before trimming (calculate_polarSPMR.m:44)
SPMR_TRIM = 1e-4            # calculate_polarSPMR.m:153
VALID_SPMR_VERSION = "2021-03-05"     # MSOTmodel.m:13
GRID_TOLERANCE = 1e-10                # MSOTmodel.m:14


def polar_coordinates(xg, zg, xd, zd):
    """Distances and azimuth angles from every detector to every pixel (MSOTmodel.m:573-597),
    host form.  xg, zg: pixel coordinates flattened in pixel order (P,), or per detector
    (D, P); xd, zd: detector coordinates (D,).  Returns (D, P) arrays."""
    xg, zg = np.asarray(xg), np.asarray(zg)
    dx = xd[:, None] - (xg if xg.ndim == 2 else xg[None, :])
    dz = zd[:, None] - (zg if zg.ndim == 2 else zg[None, :])
    dist = np.sqrt(dx * dx + dz * dz)
    rq = np.sqrt(xd * xd + zd * zd)[:, None]
    cosang = (dx * xd[:, None] + dz * zd[:, None]) / (rq * dist)
    ang = np.arccos(np.clip(cosang, -1.0, 1.0))
    return dist, ang


def spmr_indices(dist, ang, dmin, dd, da, n_angle):
    """SPMR index of every pair (MSOTmodel.m:472-478), 1-based: distance slot
    round(1 + (dist - dmin) / dd), angle slot 1 + round(angle_deg / da) (MATLAB round, half
    away from zero), linearised angle-fastest over n_angle angles."""
    i_d = matlab_round(1 + (np.asarray(dist) - dmin) / dd)
    i_a = 1 + matlab_round(np.asarray(ang) * 180.0 / np.pi / da)
    return ((i_d - 1) * n_angle + i_a).astype(np.int32)


def nshape_taps(pixel_model, c, fs, Ls=SPMR_LENGTH):
    """The pixel-model N-shape phi(c t) * (-t) on the Ls-sample grid centred at Ls // 2,
    trimmed to where it is above 1e-12 of its peak: (taps, sample offset of the first)."""
    centre = Ls // 2
    kn = np.arange(-centre, centre + 1)
    t = kn / fs
    nsh = pixel_model(c * t) * (-t)
    keep = np.nonzero(np.abs(nsh) > 1e-12 * np.abs(nsh).max())[0]
    return np.ascontiguousarray(nsh[keep[0]:keep[-1] + 1], dtype=np.float64), int(kn[keep[0]])


def grid_axes(dmin, dmax, dd, amax, da):
    """calculate_polarSPMR.m:22-25: distances dmin + (0..ceil((dmax-dmin)/dd)) dd, angles
    (0..ceil(amax/da)) da."""
    dists = dmin + np.arange(int(np.ceil((dmax - dmin) / dd)) + 1) * dd
    angles = np.arange(int(np.ceil(amax / da)) + 1) * da
    return dists, angles


def spmr_table(dists, angles_deg, width, n_sub, c, fs, pixel_model, Ls=SPMR_LENGTH, trim=SPMR_TRIM):
    """The table on the device for a distance x angle grid: (Nsir, Ls) FP64 with the angle
    index fastest, and the 1-based inclusive windows (Nsir, 2) int32."""
    import torch
    nsh, off = nshape_taps(pixel_model, c, fs, Ls)
    d = np.ascontiguousarray(dists, dtype=np.float64)
    a = np.ascontiguousarray(angles_deg, dtype=np.float64)
    ne = d.size * a.size
    tab = torch.empty((ne, Ls), dtype=torch.float64, device="cuda")
    win = torch.empty((ne, 2), dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().mbrec_spmr_table_dev(d.ctypes.data, d.size, a.ctypes.data, a.size, float(width), int(n_sub),
                                               float(c), float(fs), nsh.ctypes.data, nsh.size, off, int(Ls),
                                               float(trim), tab.data_ptr(), win.data_ptr(), _lib.stream_handle()))
    return tab.cpu().numpy(), win.cpu().numpy()


def _device_xy(xg, zg):
    import torch
    xg, zg = np.asarray(xg, dtype=np.float64), np.asarray(zg, dtype=np.float64)
    return (torch.as_tensor(np.ascontiguousarray(xg)).cuda(), torch.as_tensor(np.ascontiguousarray(zg)).cuda(),
            1 if xg.ndim == 2 else 0, xg.shape[-1])


def required_grid(xg, zg, xd, zd, resolution_fov, c):
    """MSOTmodel.m:599-616 on the device: (dmin, dmax, dd, amax_deg, da_deg)."""
    xv, zv, per, P = _device_xy(xg, zg)
    xd = np.ascontiguousarray(xd, dtype=np.float64)
    zd = np.ascontiguousarray(zd, dtype=np.float64)
    g = np.zeros(7)
    _lib.check(_lib.lib().mbrec_spmr_grid_dev(xv.data_ptr(), zv.data_ptr(), per, P, xd.ctypes.data, zd.ctypes.data,
                                              xd.size, float(resolution_fov), float(c), g.ctypes.data, None,
                                              _lib.stream_handle()))
    return tuple(float(v) for v in g[:5])


def device_indices(xg, zg, xd, zd, dmin, dd, da, n_angle):
    """MSOTmodel.m:470-478 on the device for a given grid: (D, P) int32, 1-based."""
    import torch
    xv, zv, per, P = _device_xy(xg, zg)
    xd = np.ascontiguousarray(xd, dtype=np.float64)
    zd = np.ascontiguousarray(zd, dtype=np.float64)
    out = torch.empty((xd.size, P), dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().mbrec_spmr_index_dev(xv.data_ptr(), zv.data_ptr(), per, P, xd.ctypes.data, zd.ctypes.data,
                                               xd.size, float(dmin), float(dd), float(da), int(n_angle),
                                               out.data_ptr(), _lib.stream_handle()))
    return out.cpu().numpy()


def pixel_model_id(pixel_model):
    """(hash, string) of the pixel model (MSOTmodel.m:310-318 evaluates it at -10:10 and
    hashes that with DataHash; here an MD5 of the FP64 samples, and the samples' text as
    the descriptive string)."""
    import hashlib
    v = np.asarray(pixel_model(np.arange(-10, 11, dtype=np.float64)), dtype=np.float64)
    return hashlib.md5(v.tobytes()).hexdigest(), "pixelModel(-10:10) = [" + " ".join(f"{x:.17g}" for x in v) + "]"


def spmr_file_name(probe_id, c, pixel_model_hash):
    """polarSPMR_<probe>_c_<c>_pixelModel_<hash>.mat (MSOTmodel.m:505, calculate_polarSPMR.m:40)."""
    return f"polarSPMR_{probe_id}_c_{c:g}_pixelModel_{pixel_model_hash}.mat"


def save_polar_spmr(path, table, windows, dists, angles, c, pixel_model_string, fs):
    """calculate_polarSPMR.m:188-193 (2-D, one elevation slice), MAT v5."""
    from scipy.io import savemat
    nd, na = len(dists), len(angles)
    Ls = table.shape[1]
    polar = np.transpose(table.reshape(nd, na, Ls), (1, 0, 2))[None]              # [1 x na x nd x Ls]
    se = np.transpose(np.asarray(windows, dtype=np.int32).reshape(nd, na, 2), (1, 0, 2))[None]
    savemat(path, {"SPMR_polar": polar, "SPMR_timestep_index_start_and_end": se,
                   "speed_of_sound_coupling": float(c), "distance_mat_SPMR": np.asarray(dists)[None],
                   "angle_azimuth_mat_SPMR": np.asarray(angles)[None], "elevation_mat_SPMR": np.zeros((1, 1)),
                   "filter_type": "none", "filter_cutoff": float(fs / 2), "filter_order": 0,
                   "spmr_version": VALID_SPMR_VERSION, "pixel_model_string_SPMR": pixel_model_string,
                   "field_sampling_frequency": float(fs)}, do_compression=False)


def load_polar_spmr(path):
    """(table (Nsir, Ls), windows (Nsir, 2), dists, angles, version, pixel model string)
    from a polarSPMR file (MSOTmodel.m:559-571: squeeze, permute [3 1 2], windows reshaped
    to [2 x Nsir])."""
    from scipy.io import loadmat
    z = loadmat(path)
    polar = np.asarray(z["SPMR_polar"], dtype=np.float64)
    polar = polar.reshape(polar.shape[-3:]) if polar.ndim == 4 else polar               # [na x nd x Ls]
    na, nd, Ls = polar.shape
    table = np.transpose(polar, (1, 0, 2)).reshape(nd * na, Ls)
    se = np.asarray(z["SPMR_timestep_index_start_and_end"], dtype=np.int32)
    se = se.reshape(se.shape[-3:]) if se.ndim == 4 else se
    windows = np.transpose(se, (1, 0, 2)).reshape(nd * na, 2)
    version = str(np.asarray(z["spmr_version"]).ravel()[0]) if "spmr_version" in z else "0"
    pstr = str(np.asarray(z["pixel_model_string_SPMR"]).ravel()[0]) if "pixel_model_string_SPMR" in z else ""
    return table, windows, np.asarray(z["distance_mat_SPMR"]).ravel(), np.asarray(z["angle_azimuth_mat_SPMR"]).ravel(), \
        version, pstr


def _le(a, b):
    return a < b or abs(a - b) < GRID_TOLERANCE


def _ge(a, b):
    return a > b or abs(a - b) < GRID_TOLERANCE


def sir_tables_for_model(xg, zg, xd, zd, resolution_fov, c, fs, width, n_sub, pixel_model,
                         Ls=SPMR_LENGTH, spmr_dir=None, probe_id="probe"):
    """get_SPMR_and_SPMR_indices (MSOTmodel.m:463-480) on the device, with the polarSPMR
    file when ``spmr_dir`` is given.  xg, zg: pixel coordinates (P,) or virtual positions
    (D, P).  Returns SPMR (Nsir, Ls), windows (Nsir, 2) int32, indices (D, P) int32
    (1-based) and the grid."""
    dmin, dmax, dd, amax, da = required_grid(xg, zg, xd, zd, resolution_fov, c)
    path = None
    if spmr_dir is not None:
        h, pstr = pixel_model_id(pixel_model)
        path = os.path.join(spmr_dir, spmr_file_name(probe_id, c, h))
        if os.path.exists(path):
            table, se, dists, angles, version, fstr = load_polar_spmr(path)
            f_dd = dists[1] - dists[0] if len(dists) > 1 else dd
            f_da = angles[1] - angles[0] if len(angles) > 1 else da
            ok = (version == VALID_SPMR_VERSION and fstr == pstr and _le(dists[0], dmin) and _ge(dists[-1], dmax) and
                  _le(f_dd, dd) and _ge(angles[-1], amax) and _le(f_da, da))
            if ok:
                sidx = device_indices(xg, zg, xd, zd, dists[0], f_dd, f_da, len(angles))
                return table, se, sidx, dict(dmin=float(dists[0]), dd=float(f_dd), da=float(f_da),
                                             n_angle=len(angles), n_dist=len(dists), file=path, loaded=True)
            if version == VALID_SPMR_VERSION:      # the more general of both grids (MSOTmodel.m:527-535)
                dmin, dmax = min(dists[0], dmin), max(dists[-1], dmax)
                dd, amax, da = min(f_dd, dd), max(angles[-1], amax), min(f_da, da)
    dists, angles = grid_axes(dmin, dmax, dd, amax, da)
    table, se = spmr_table(dists, angles, width, n_sub, c, fs, pixel_model, Ls)
    sidx = device_indices(xg, zg, xd, zd, dmin, dd, da, len(angles))
    if path is not None:
        os.makedirs(spmr_dir, exist_ok=True)
        save_polar_spmr(path, table, se, dists, angles, c, pstr, fs)
    return table, se, sidx, dict(dmin=float(dmin), dd=float(dd), da=float(da), n_angle=len(angles),
                                 n_dist=len(dists), file=path, loaded=False)
