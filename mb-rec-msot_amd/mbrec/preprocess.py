"""Sinogram preprocessing on the device (load_and_preprocess/).

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
