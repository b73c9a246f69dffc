"""The four MEX entry points with the reference's call signatures and errors.

Arrays are given in MATLAB shape (e.g. image [W x P], sirIndices [P x D]) and
are read in column-major order, exactly as a MATLAB mxArray is laid out; a
1-D array is a column.  Type checks raise ``MexError`` with the reference's
identifiers (forwardMex.cpp:30-50 and the corresponding lines of the other
three files).  Computation runs in libmbrec (FP32 on the device, FP64 out).
"""
from __future__ import annotations

import numpy as np

from . import _lib


class MexError(_lib.MbrecError):
    pass


def _mat(a):
    a = np.asarray(a)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    return a


def _check_double(name, i, a):
    if a.dtype != np.float64:
        _raise_type(name, i, False)


def _check_int32(name, i, a):
    if a.dtype != np.int32:
        _raise_type(name, i, True)


def _raise_type(name, i, want_int32):
    L = _lib.lib()
    rc = L.mbrec_mex_type_error(name.encode(), int(i), int(want_int32))
    raise MexError(L.mbrec_last_error_id().decode(), L.mbrec_last_error().decode(), rc)


def _f(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def forwardMex(image, sirIndices, travelTimes, sir, numTimesteps, sirTimestepsStartAndEnd, R):
    """sig = forwardMex(image, sirIndices, travelTimes, sir, numTimesteps, sirSE, R) -> [W x nT x D]."""
    args = [_mat(a) for a in (image, sirIndices, travelTimes, sir, numTimesteps, sirTimestepsStartAndEnd, R)]
    kinds = [0, 1, 0, 0, 1, 1, 0]
    for i, (a, k) in enumerate(zip(args, kinds)):
        (_check_int32 if k else _check_double)("forwardMex", i, a)
    img, sidx, tt, s, nt, se, r = args
    W, P = img.shape[0], img.shape[1]
    D = sidx.shape[1]
    nT = int(nt.ravel()[0])
    out = np.zeros((W, nT, D), dtype=np.float64, order="F")
    img, sidx, tt, s, se = _f(img, np.float64), _f(sidx, np.int32), _f(tt, np.float64), _f(s, np.float64), _f(se, np.int32)
    _lib.check(_lib.lib().mbrec_forwardMex(img.ctypes.data, W, P, sidx.ctypes.data, tt.ctypes.data, D,
                                           s.ctypes.data, s.shape[0], s.shape[1], nT, se.ctypes.data,
                                           float(r.ravel()[0]), out.ctypes.data))
    return out


def transposeMex(sigMat, sirIndices, travelTimes, sir, sirTimestepsStartAndEnd, R):
    """img = transposeMex(sigMat, sirIndices, travelTimes, sir, sirSE, R) -> [W x P]."""
    sig = np.asarray(sigMat)
    if sig.ndim < 3:
        sig = sig.reshape(sig.shape + (1,) * (3 - sig.ndim))
    args = [sig] + [_mat(a) for a in (sirIndices, travelTimes, sir, sirTimestepsStartAndEnd, R)]
    kinds = [0, 1, 0, 0, 1, 0]
    for i, (a, k) in enumerate(zip(args, kinds)):
        (_check_int32 if k else _check_double)("transposeMex", i, a)
    sig, sidx, tt, s, se, r = args
    W, nT, D = sig.shape
    P = sidx.shape[0]
    out = np.zeros((W, P), dtype=np.float64, order="F")
    sig, sidx, tt, s, se = _f(sig, np.float64), _f(sidx, np.int32), _f(tt, np.float64), _f(s, np.float64), _f(se, np.int32)
    _lib.check(_lib.lib().mbrec_transposeMex(sig.ctypes.data, W, nT, D, sidx.ctypes.data, tt.ctypes.data, P,
                                             s.ctypes.data, s.shape[0], s.shape[1], se.ctypes.data,
                                             float(r.ravel()[0]), out.ctypes.data))
    return out


def forwardDelayAndSumMex(image, travelTimes, numTimesteps, R):
    """sig = forwardDelayAndSumMex(image, travelTimes, numTimesteps, R) -> [W x nT x D]."""
    args = [_mat(a) for a in (image, travelTimes, numTimesteps, R)]
    kinds = [0, 0, 1, 0]
    for i, (a, k) in enumerate(zip(args, kinds)):
        (_check_int32 if k else _check_double)("forwardDelayAndSumMex", i, a)
    img, tt, nt, r = args
    W, P = img.shape
    D = tt.shape[1]
    nT = int(nt.ravel()[0])
    out = np.zeros((W, nT, D), dtype=np.float64, order="F")
    img, tt = _f(img, np.float64), _f(tt, np.float64)
    _lib.check(_lib.lib().mbrec_forwardDelayAndSumMex(img.ctypes.data, W, P, tt.ctypes.data, D, nT,
                                                      float(r.ravel()[0]), out.ctypes.data))
    return out


def transposeDelayAndSumMex(sigMat, travelTimes, R):
    """img = transposeDelayAndSumMex(sigMat, travelTimes, R) -> [W x P]."""
    sig = np.asarray(sigMat)
    if sig.ndim < 3:
        sig = sig.reshape(sig.shape + (1,) * (3 - sig.ndim))
    args = [sig] + [_mat(a) for a in (travelTimes, R)]
    for i, a in enumerate(args):
        _check_double("transposeDelayAndSumMex", i, a)
    sig, tt, r = args
    W, nT, D = sig.shape
    P = tt.shape[0]
    out = np.zeros((W, P), dtype=np.float64, order="F")
    sig, tt = _f(sig, np.float64), _f(tt, np.float64)
    _lib.check(_lib.lib().mbrec_transposeDelayAndSumMex(sig.ctypes.data, W, nT, D, tt.ctypes.data, P,
                                                        float(r.ravel()[0]), out.ctypes.data))
    return out
