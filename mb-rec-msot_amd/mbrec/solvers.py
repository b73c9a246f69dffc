"""Non-negative reconstruction methods (reconstruction_model_based/reconstruction/).

The solver loops run on the device (libmbrec); these functions mirror the
reference's MATLAB entry points: same argument order, same constants, same
outputs.  ``data_filt`` is (W, D, nT) or (D, nT) (MATLAB [nT x D x W] memory
order); ``rec_img`` is (W, Nx, Nz) or (Nx, Nz).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib


def _is_identity(f):
    return f is None or f == "identity" or getattr(f, "is_identity", False)


def identity(x):
    return x


identity.is_identity = True


def _to_device(x):
    import torch
    if isinstance(x, np.ndarray):
        return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32)).cuda(), True
    if x.dtype != torch.float32:
        raise TypeError("device tensors must be float32")
    return x.contiguous(), False


def rec_nn_with_L2_reg(model, data_filt, num_iter_nn_outer, lambdaL2=0.0, RegL2=None, RegL2T=None,
                       lambdaL2_2=None, RegL2_2=None, RegL2T_2=None, *, coupled=True, init=None):
    """rec_nn_with_L2_reg.m:1-66 (RegL2 = identity).

    Returns (rec_img, lCurveErrImg, lCurveErrReg, lCurveErrReg_2, err, t) like the
    reference; ``err`` stays zero (calculate_objective_per_iter = false, :35) and
    ``t`` holds the solve's wall time split evenly over the outer iterations.
    ``coupled=True`` is the reference's semantics for a multi-wavelength call
    (shared LSQR scalars, nnls_L2_less.m:56 on stacked vectors); ``coupled=False``
    solves every sinogram independently.  ``init`` (optional image) seeds the
    outer iterate (warm start); the reference's own start value is ignored by
    nnls_L2_less.m:28-29 and so is not computed here.
    """
    import torch
    if lambdaL2 and lambdaL2 > 0 and not (_is_identity(RegL2) and _is_identity(RegL2T)):
        raise NotImplementedError("only RegL2 = identity is provided on the device")
    if lambdaL2_2 and lambdaL2_2 > 0:
        raise NotImplementedError("the second (Laplacian) L2 term is not provided on the device")
    y, host = _to_device(data_filt)
    D, nT = model.shape_sinogram
    B = y.numel() // (D * nT)
    y = y.view(B, D, nT)
    rec = torch.empty((B, *model.shape_image), dtype=torch.float32, device=y.device)
    initp = None
    if init is not None:
        it, _ = _to_device(init)
        if it.numel() != rec.numel():
            raise ValueError("init has the wrong number of pixels")
        initp = it.contiguous()
    G = 1 if coupled else B
    lc = np.zeros(2 * G, dtype=np.float64)
    L = _lib.lib()
    t0 = time.perf_counter()
    _lib.check(L.mbrec_rec_nn_with_L2_reg_dev(
        model.plan.h, y.data_ptr(), int(B), int(bool(coupled)), int(num_iter_nn_outer),
        float(lambdaL2 or 0.0), initp.data_ptr() if initp is not None else None, rec.data_ptr(),
        lc.ctypes.data, _lib.stream_handle()))
    el = time.perf_counter() - t0
    errimg = lc[0::2] if not coupled else lc[0]
    errreg = lc[1::2] if not coupled else lc[1]
    err = np.zeros(int(num_iter_nn_outer))
    t = np.full(int(num_iter_nn_outer), el / max(1, int(num_iter_nn_outer)))
    if host:
        out = rec.cpu().numpy().astype(np.float64)
        if np.ndim(data_filt) == 2:
            out = out[0]
        rec = out
    return rec, errimg, errreg, 0.0, err, t


def rec_nn_with_L2_reg_batch(model, data_filt, num_iter_nn_outer, lambdaL2, init=None, init_dual=None):
    """B independent rec_nn_with_L2_reg.m solves (RegL2 = identity), each with its own
    lambdaL2 (array [B]), in one batched device call.  data_filt (B, D, nT) device
    tensor; init / init_dual (optional (B, Nx, Nz)) seed the outer iterate w and the
    scaled dual v of nnls_L2_less.m:28-29.  Returns the (B, Nx, Nz) device tensor."""
    import torch
    y, _ = _to_device(data_filt)
    D, nT = model.shape_sinogram
    B = y.numel() // (D * nT)
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lambdaL2, dtype=np.float64), (B,)))
    rec = torch.empty((B, *model.shape_image), dtype=torch.float32, device=y.device)
    ptr = []
    for t in (init, init_dual):
        if t is None:
            ptr.append(None)
            continue
        td, _ = _to_device(t)
        if td.numel() != rec.numel():
            raise ValueError("init / init_dual have the wrong number of pixels")
        ptr.append(td)
    _lib.check(_lib.lib().mbrec_rec_nn_with_L2_reg_batch_dev(
        model.plan.h, y.data_ptr(), int(B), int(num_iter_nn_outer), lam.ctypes.data,
        ptr[0].data_ptr() if ptr[0] is not None else None, ptr[1].data_ptr() if ptr[1] is not None else None,
        rec.data_ptr(), None, _lib.stream_handle()))
    return rec


def rec_nn_with_L1_eye_reg(model, data_filt, num_iter, regParam, *, coupled=True):
    """rec_nn_with_L1_eye_reg.m:1-50: non-negative L1 (identity) reconstruction by split
    Bregman (nnls_L1_eye_matrix_less.m) with MATLAB-cgs inner solves, started from A'y.

    Returns (rec_img, lCurveErrImg, lCurveErrReg, err, t); lCurveErrImg = ||A rec - y||^2
    (:38) and lCurveErrReg = ||rec||_1 (:40), per sinogram when coupled=False."""
    import torch
    y, host = _to_device(data_filt)
    D, nT = model.shape_sinogram
    B = y.numel() // (D * nT)
    y = y.view(B, D, nT)
    rec = torch.empty((B, *model.shape_image), dtype=torch.float32, device=y.device)
    t0 = time.perf_counter()
    _lib.check(_lib.lib().mbrec_rec_nn_with_L1_eye_reg_dev(
        model.plan.h, y.data_ptr(), int(B), int(bool(coupled)), int(num_iter), float(regParam),
        rec.data_ptr(), _lib.stream_handle()))
    el = time.perf_counter() - t0
    f = model.plan.forward(rec.view(B, -1))
    res = ((f.double() - y.double()) ** 2).sum(dim=(1, 2))
    l1 = rec.double().abs().sum(dim=(1, 2))
    if coupled:
        e_img, e_reg = float(res.sum()), float(l1.sum())
    else:
        e_img, e_reg = res.cpu().numpy(), l1.cpu().numpy()
    err = np.zeros(int(num_iter))
    t = np.full(int(num_iter), el / max(1, int(num_iter)))
    if host:
        out = rec.cpu().numpy().astype(np.float64)
        rec = out[0] if np.ndim(data_filt) == 2 else out
    return rec, e_img, e_reg, err, t


SPARSA_MINITER_DEFAULT = 5   # SpaRSA 2.0's default MiniterA; the reference does not pass one


def rec_nn_with_TV_reg(model, data_filt, num_iter, regParam, *, coupled=True, inner_iterations=2,
                       inner_min_iterations=SPARSA_MINITER_DEFAULT):
    """rec_nn_with_TV_reg.m:1-55: non-negative TV reconstruction (square grids only).

    Returns (rec_img, lCurveErrImg, lCurveErrReg, err, t) with lCurveErrImg = ||A rec - y||^2
    (:51) and lCurveErrReg = Phi(rec) = isotropic TV (:52), per sinogram when coupled=False.
    The inner SpaRSA takes ``inner_iterations`` = MaxiterA (:27, passed at
    nnls_sparsa_based_reg_matrix_less.m:64) and ``inner_min_iterations`` = MiniterA (not
    passed by the reference, so SpaRSA's default applies): no fewer than MiniterA steps, and
    past MiniterA no more than MaxiterA, stopping early on a relative objective change <= 1e-8."""
    import torch
    y, host = _to_device(data_filt)
    D, nT = model.shape_sinogram
    B = y.numel() // (D * nT)
    y = y.view(B, D, nT)
    rec = torch.empty((B, *model.shape_image), dtype=torch.float32, device=y.device)
    G = 1 if coupled else B
    lc = np.zeros(2 * G)
    t0 = time.perf_counter()
    _lib.check(_lib.lib().mbrec_rec_nn_with_TV_reg_dev(
        model.plan.h, y.data_ptr(), int(B), int(bool(coupled)), int(num_iter), float(regParam),
        int(inner_iterations), int(inner_min_iterations), rec.data_ptr(), lc.ctypes.data, _lib.stream_handle()))
    el = time.perf_counter() - t0
    e_img = lc[0] if coupled else lc[0::2]
    e_reg = lc[1] if coupled else lc[1::2]
    err = np.zeros(int(num_iter))
    t = np.full(int(num_iter), el / max(1, int(num_iter)))
    if host:
        out = rec.cpu().numpy().astype(np.float64)
        rec = out[0] if np.ndim(data_filt) == 2 else out
    return rec, e_img, e_reg, err, t
