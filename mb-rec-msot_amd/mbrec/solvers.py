"""Non-negative reconstruction methods (reconstruction_model_based/reconstruction/).

The solver loops run on the device (libmbrec); these functions mirror the
reference's MATLAB entry points: same argument order, same constants, same
outputs.  ``data_filt`` is (W, D, nT) or (D, nT) (MATLAB [nT x D x W] memory
order); ``rec_img`` is (W, Nx, Nz) or (Nx, Nz).
"""
from __future__ import annotations

import ctypes
import time
from ctypes import c_int as C_int

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


def laplacian_per_wavelength(x, shape=None):
    """laplacian_per_wavelength.m:1-10: MATLAB del2(img, 1) of every image of x (a device
    tensor (W, Nx, Nz), or (W, P) with ``shape`` = (Nx, Nz)) on the device.  Passed as
    RegL2 / RegL2T (or RegL2_2 / RegL2T_2) of rec_nn_with_L2_reg it selects the device
    solver's built-in del2 term, used as its own transpose as the reference does
    (reconstruction_overview_script.m:50-51)."""
    import torch
    t = x if x.dtype == torch.float32 else x.float()
    t = t.contiguous()
    Nx, Nz = shape if shape is not None else t.shape[-2:]
    B = t.numel() // (Nx * Nz)
    out = torch.empty_like(t)
    _lib.check(_lib.lib().mbrec_del2_dev(t.data_ptr(), out.data_ptr(), int(Nz), int(Nx), int(B),
                                         _lib.stream_handle()))
    return out


laplacian_per_wavelength.is_del2 = True


class _DeviceImages:
    """A device buffer seen through __cuda_array_interface__ (no copy)."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f4", "data": (int(ptr), False),
                                         "version": 3, "strides": None}


def _term(lam, R, RT, shape, keep):
    """mbrec_l2_term for one L2 term (lambda, R, R'): identity, del2 or a Python callback
    (R, RT map device tensors (B, Nx, Nz) -> (B, Nx, Nz))."""
    import torch
    t = _lib.L2Term()
    t.lambda_ = float(lam or 0.0)
    if _is_identity(R) and _is_identity(RT):
        t.kind = 0
    elif getattr(R, "is_del2", False) and getattr(RT, "is_del2", False):
        t.kind = 1
    else:
        Nx, Nz = shape

        def apply(ctx, pin, pout, B, P, transp, stream):
            try:
                src = torch.as_tensor(_DeviceImages(pin, (B, Nx, Nz)), device="cuda")
                dst = torch.as_tensor(_DeviceImages(pout, (B, Nx, Nz)), device="cuda")
                dst.copy_((RT if transp else R)(src))
                return 0
            except Exception:       # reported by the library as mbrec:regOperator
                return 1
        cb = _lib.REG_APPLY(apply)
        keep.append(cb)
        t.kind = 2
        t.apply = cb
    return t


def rec_nn_with_L2_reg(model, data_filt, num_iter_nn_outer, lambdaL2=0.0, RegL2=None, RegL2T=None,
                       lambdaL2_2=None, RegL2_2=None, RegL2T_2=None, *, coupled=True, init=None):
    """rec_nn_with_L2_reg.m:1-66.

    RegL2 / RegL2T and RegL2_2 / RegL2T_2: None or ``identity`` (the identity),
    ``laplacian_per_wavelength`` (MATLAB del2 per image, built into the device solver), or
    any pair of callables mapping device image tensors (B, Nx, Nz) to (B, Nx, Nz) (run on
    the current stream inside the solve).  As in incorporate_l2_reg_into_model_functions.m
    the second term enters only with lambdaL2 > 0.
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
    plain = RegL2 is None and RegL2T is None and RegL2_2 is None and not (lambdaL2_2 and lambdaL2_2 > 0)
    if not plain:
        return _rec_nn_with_L2_reg_terms(model, data_filt, num_iter_nn_outer, lambdaL2, RegL2, RegL2T, lambdaL2_2,
                                         RegL2_2, RegL2T_2, coupled, init)
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


def _rec_nn_with_L2_reg_terms(model, data_filt, num_iter, lambdaL2, RegL2, RegL2T, lambdaL2_2, RegL2_2, RegL2T_2,
                              coupled, init):
    import torch
    y, host = _to_device(data_filt)
    D, nT = model.shape_sinogram
    B = y.numel() // (D * nT)
    y = y.view(B, D, nT)
    rec = torch.empty((B, *model.shape_image), dtype=torch.float32, device=y.device)
    keep = []
    terms = [_term(lambdaL2, RegL2, RegL2T, model.shape_image, keep)]
    if RegL2_2 is not None or (lambdaL2_2 and lambdaL2_2 > 0):
        terms.append(_term(lambdaL2_2, RegL2_2, RegL2T_2, model.shape_image, keep))
    arr = (_lib.L2Term * len(terms))(*terms)
    initp = None
    if init is not None:
        it, _ = _to_device(init)
        if it.numel() != rec.numel():
            raise ValueError("init has the wrong number of pixels")
        initp = it.contiguous()
    G = 1 if coupled else B
    lc = np.zeros(3 * G, dtype=np.float64)
    t0 = time.perf_counter()
    _lib.check(_lib.lib().mbrec_rec_nn_with_L2_reg_terms_dev(
        model.plan.h, y.data_ptr(), int(B), int(bool(coupled)), int(num_iter), arr, len(terms),
        initp.data_ptr() if initp is not None else None, rec.data_ptr(), lc.ctypes.data, _lib.stream_handle()))
    el = time.perf_counter() - t0
    pick = (lambda j: lc[j]) if coupled else (lambda j: lc[j::3])
    err = np.zeros(int(num_iter))
    t = np.full(int(num_iter), el / max(1, int(num_iter)))
    if host:
        out = rec.cpu().numpy().astype(np.float64)
        rec = out[0] if np.ndim(data_filt) == 2 else out
    return rec, pick(0), pick(1), pick(2), err, t


def rec_nn_with_L2_reg_batch(model, data_filt, num_iter_nn_outer, lambdaL2, init=None, init_dual=None,
                             return_dual=False):
    """B independent rec_nn_with_L2_reg.m solves (RegL2 = identity), each with its own
    lambdaL2 (array [B]), in one batched device call.  data_filt (B, D, nT) device
    tensor; init / init_dual (optional (B, Nx, Nz)) seed the outer iterate w and the
    scaled dual v of nnls_L2_less.m:28-29.  Returns the (B, Nx, Nz) device tensor, or
    (rec, dual) with return_dual=True: passing them back as init / init_dual continues
    the same iteration sequence."""
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
    dual = torch.empty_like(rec) if return_dual else None
    _lib.check(_lib.lib().mbrec_rec_nn_with_L2_reg_batch_dev(
        model.plan.h, y.data_ptr(), int(B), int(num_iter_nn_outer), lam.ctypes.data,
        ptr[0].data_ptr() if ptr[0] is not None else None, ptr[1].data_ptr() if ptr[1] is not None else None,
        rec.data_ptr(), dual.data_ptr() if dual is not None else None, None, _lib.stream_handle()))
    return (rec, dual) if return_dual else rec


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


def SB_ITV2(g, mu):
    """SB_ITV2.m:1-40: arg min_u 1/2||u - g||^2 + mu ITV(u) by split Bregman on the device.

    g: one square image (N, N) or a stack (B, N, N) in this package's layout ((Nx, Nz),
    z fastest), numpy or a device tensor; mu a scalar or one value per image.  The inner
    cgs(I + B'B, ., 1e-5, 100) of SB_ITV2.m:31 is the exact DCT Neumann solve.  Returns
    (u, iterations) with u of the same kind and shape (FP32 on the device)."""
    import torch
    host = not hasattr(g, "data_ptr")
    t = torch.as_tensor(np.asarray(g, dtype=np.float32)).cuda() if host else g.float()
    t = t.contiguous()
    N = t.shape[-1]
    if t.shape[-2] != N:
        raise ValueError("SB_ITV2: square images only")
    B = t.numel() // (N * N)
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (B,)))
    u = torch.empty_like(t)
    it = C_int()
    _lib.check(_lib.lib().mbrec_sb_itv2_dev(t.data_ptr(), int(N), int(B), m.ctypes.data, u.data_ptr(),
                                            ctypes.addressof(it), _lib.stream_handle()))
    return (u.cpu().numpy().astype(np.float64) if host else u), it.value
