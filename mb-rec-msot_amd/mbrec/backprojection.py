"""Backprojection-type reconstructions (reconstruction_backprojection/*.m) on the device.

Arrays follow MATLAB memory order with reversed axes: sinograms (W, D, nT) or (D, nT)
(MATLAB [nT x D x W]), pixel grids (Nx, Nz) (MATLAB [Nz x Nx]).  Results are FP64.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).cuda()


def dmas_cf(sigMat, fov_x, fov_z, speed_of_sound, sampling_frequency,
            cropped_or_unrecorded_signals_at_sinogram_start, angular_coverage, detector_radius):
    """dmas_cf.m:19-107.  sigMat (W, D, nT) or (D, nT) float32/float64 (computed from its
    FP32 values); fov_x, fov_z (Nx, Nz) pixel coordinates.  Returns
    (p0_rec_dmas_cf, p0_rec_dmas, cf, p0_rec_das), each (W, Nx, Nz) or (Nx, Nz)."""
    import torch
    sig = np.asarray(sigMat)
    squeeze = sig.ndim == 2
    sig = sig.reshape((-1,) + sig.shape[-2:])
    W, D, nT = sig.shape
    fx = np.asarray(fov_x, dtype=np.float64)
    fz = np.asarray(fov_z, dtype=np.float64)
    shape = fx.shape
    P = fx.size
    ds = _dev(sig, np.float32)
    dx, dz = _dev(fx.ravel(), np.float64), _dev(fz.ravel(), np.float64)
    out = torch.empty((4, W, P), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().mbrec_dmas_cf_dev(dx.data_ptr(), dz.data_ptr(), P, ds.data_ptr(), W, D, nT,
                                            float(speed_of_sound), float(sampling_frequency),
                                            int(cropped_or_unrecorded_signals_at_sinogram_start),
                                            float(angular_coverage), float(detector_radius), out.data_ptr(),
                                            _lib.stream_handle()))
    o = out.cpu().numpy().reshape((4, W) + shape)
    if squeeze:
        o = o[:, 0]
    return o[0], o[1], o[2], o[3]


def reconstruct_dmas_cf(model, sinogram):
    """reconstruct_dmas_cf.m:1-14."""
    X = model.Discretization.region["xPositionsFov"].T
    Z = model.Discretization.region["zPositionsFov"].T
    return dmas_cf(sinogram, X, Z, model.Medium.speedOfSound, model.Probe.DAC.frequency,
                   model.DataPreprocessing.numCroppedSamplesAtSinogramStart + model.Probe.DAC.delayBeforeRecording,
                   model.Probe.detector.angularCoverage, model.Probe.detector.radius)


def bp_kernel(t, device_out=False):
    """bp_kernel.m:1-12: K = -imag(trapz(lambda, H0 .* conj(H0) .* lambda)), H0 = besselh(0, t'*lambda).
    Returns the (Nt, Nt) FP64 matrix (numpy, or the device tensor in column-major
    memory when device_out=True)."""
    import torch
    t = np.ascontiguousarray(np.asarray(t, dtype=np.float64).ravel())
    Nt = t.size
    K = torch.empty((Nt, Nt), dtype=torch.float64, device="cuda")   # memory = column-major [Nt x Nt]
    _lib.check(_lib.lib().mbrec_bp_kernel_dev(t.ctypes.data, Nt, K.data_ptr(), _lib.stream_handle()))
    if device_out:
        return K
    return K.cpu().numpy().T.copy()


def backproject_waveeq(sigMat, fov_x, fov_y, speed_of_sound, sampling_frequency,
                       cropped_or_unrecorded_signals_at_sinogram_start, angular_coverage, detector_radius,
                       psf, K):
    """backproject_waveeq.m:1-50.  sigMat (W, D, nT) or (D, nT); fov_x, fov_y meshgrid
    coordinates given as (Nx, Nz) arrays (MATLAB [Nz x Nx]); psf a (pm, pn) = (time,
    detector) array (MATLAB shape); K the (Nt, Nt) bp_kernel matrix (numpy) or the device
    tensor from bp_kernel(..., device_out=True).  Returns (W, Nx, Nz) or (Nx, Nz)."""
    import torch
    sig = np.asarray(sigMat, dtype=np.float64)
    squeeze = sig.ndim == 2
    sig = sig.reshape((-1,) + sig.shape[-2:])
    W, D, nT = sig.shape
    fx = np.asarray(fov_x, dtype=np.float64)
    fy = np.asarray(fov_y, dtype=np.float64)
    Nx, Nz = fx.shape
    xv = np.ascontiguousarray(fx[:, 0])
    zv = np.ascontiguousarray(fy[0, :])
    ps = np.atleast_2d(np.asarray(psf, dtype=np.float64))
    if np.ndim(psf) == 1:
        ps = ps.T
    pm, pn = ps.shape
    psf_f = np.asfortranarray(ps).ravel(order="F").copy()
    if isinstance(K, torch.Tensor):
        Kd = K
        Nt = K.shape[0]
    else:
        Kn = np.asarray(K, dtype=np.float64)
        Nt = Kn.shape[0]
        Kd = _dev(np.asfortranarray(Kn).ravel(order="F"), np.float64)
    ds = _dev(sig, np.float64)
    out = torch.empty((W, Nx, Nz), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().mbrec_backproject_waveeq_dev(
        xv.ctypes.data, zv.ctypes.data, Nz, Nx, ds.data_ptr(), W, D, nT, float(speed_of_sound),
        float(sampling_frequency), int(cropped_or_unrecorded_signals_at_sinogram_start), float(angular_coverage),
        float(detector_radius), psf_f.ctypes.data, pm, pn, Kd.data_ptr(), Nt, out.data_ptr(),
        _lib.stream_handle()))
    o = out.cpu().numpy()
    return o[0] if squeeze else o
