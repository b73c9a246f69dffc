"""Sinogram preprocessing on the device (load_and_preprocess/).

``filter_butter_zero_phase``, ``apply_butterworth_window_to_sinogram`` and
``interpolate_signals_of_broken_transducers`` run in libmbrec (mbrec_preprocess.hip).
``upsample_detectors_in_sinogram`` restates upsample_detectors_in_sinogram.m:1-13:
interp2(..., 'spline') on a grid whose time axis is queried at its own knots, i.e.
a not-a-knot cubic spline across detectors (knots 1:f:D*f, queries 1..D*f, the
queries past the last knot extrapolated with the last cubic piece).  The spline is
linear in the data, so it is one [D*f x D] matrix U, applied on the device as an
FP64 GEMM per sinogram.
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import CubicSpline

from . import _lib


def spline_upsampling_matrix(D: int, factor: int) -> np.ndarray:
    """U [D*factor x D]: column j = not-a-knot spline through e_j at knots 1:f:D*f,
    evaluated at 1..D*f (with extrapolation)."""
    knots = np.arange(1, D * factor + 1, factor, dtype=np.float64)
    q = np.arange(1, D * factor + 1, dtype=np.float64)
    cs = CubicSpline(knots, np.eye(D), axis=0, bc_type="not-a-knot", extrapolate=True)
    return cs(q)


def upsample_detectors_in_sinogram(sinograms, factor: int):
    """sinograms (F, W, D, nT), (W, D, nT) or (D, nT) (MATLAB [nT x D x W x F]) ->
    the same layout with D*factor detectors, FP64."""
    import torch
    s = np.asarray(sinograms, dtype=np.float64)
    lead = s.shape[:-2]
    D, nT = s.shape[-2:]
    nb = int(np.prod(lead)) if lead else 1
    U = spline_upsampling_matrix(D, int(factor))
    Dup = U.shape[0]
    Ud = torch.as_tensor(np.asfortranarray(U).ravel(order="F").copy()).cuda()
    sd = torch.as_tensor(np.ascontiguousarray(s.reshape(nb, D, nT))).cuda()
    out = torch.empty((nb, Dup, nT), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().mbrec_upsample_detectors_dev(Ud.data_ptr(), Dup, D, sd.data_ptr(), nT, nb,
                                                       out.data_ptr(), _lib.stream_handle()))
    return out.cpu().numpy().reshape(lead + (Dup, nT))


def _columns(x):
    """(array, lead shape, D, nT, host?) -> device FP64 [cols][nT] tensor."""
    import torch
    host = not hasattr(x, "data_ptr")
    t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).cuda() if host else x.double().contiguous()
    return t, host


def filter_butter_zero_phase(signal_matrix, sampling_frequency, filter_frequencies, use_fft=True):
    """filter_butter_zero_phase.m:1-53 on the device, FFT branch: demean along time (when
    the high-pass edge is > 0), then the zero-phase response |H|^2 of the 8th-order
    Butterworth low-pass (filter_frequencies[1]) and the 2nd-order high-pass
    (filter_frequencies[0]) applied circularly along time.  signal_matrix: time along the
    last axis ((W, D, nT) = MATLAB [nT x D x W]), numpy or device tensor; FP64 result of the
    same kind and shape."""
    import torch
    if not use_fft:
        raise NotImplementedError("only the FFT branch (use_fft = true) is provided")
    t, host = _columns(signal_matrix)
    nT = t.shape[-1]
    out = torch.empty_like(t)
    _lib.check(_lib.lib().mbrec_filter_butter_zero_phase_dev(
        t.data_ptr(), int(nT), int(t.numel() // nT), float(sampling_frequency), float(filter_frequencies[0]),
        float(filter_frequencies[1]), out.data_ptr(), _lib.stream_handle()))
    return out.cpu().numpy() if host else out


def apply_butterworth_window_to_sinogram(sinograms, n_butter, low_cutoff, high_cutoff):
    """apply_butterworth_window_to_sinogram.m:1-10 on the device: each time series becomes
    mean + (s - mean) |h(pi t / N)| with h the order-n_butter digital Butterworth band-pass
    [low_cutoff, high_cutoff] / N (N = number of samples).  Time along the last axis; FP64
    result of the same kind and shape."""
    import torch
    t, host = _columns(sinograms)
    nT = t.shape[-1]
    out = t.clone()
    _lib.check(_lib.lib().mbrec_apply_butterworth_window_dev(
        out.data_ptr(), int(nT), int(out.numel() // nT), int(n_butter), float(low_cutoff), float(high_cutoff),
        _lib.stream_handle()))
    return out.cpu().numpy() if host else out


def interpolate_signals_of_broken_transducers(sigMat, broken_transducers):
    """interpolate_signals_of_broken_transducers.m:1-20 on the device: every detector's
    signal replaced by the not-a-knot cubic spline across the working detectors (1-based
    indices not in broken_transducers), evaluated at its index (extrapolated past the end
    ones), per time sample.  sigMat (..., D, nT); FP64 result of the same kind and shape."""
    import torch
    broken = np.unique(np.asarray(broken_transducers, dtype=np.int64).ravel())
    if broken.size == 0:
        return sigMat
    t, host = _columns(sigMat)
    D, nT = t.shape[-2:]
    nb = t.numel() // (D * nT)
    out = torch.empty_like(t)
    br = np.ascontiguousarray(broken, dtype=np.int32)
    _lib.check(_lib.lib().mbrec_interpolate_broken_transducers_dev(
        t.data_ptr(), int(D), int(nT), int(nb), br.ctypes.data, int(br.size), out.data_ptr(), _lib.stream_handle()))
    return out.cpu().numpy() if host else out
