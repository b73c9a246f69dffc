"""Oracle for the MSOT operator / NNLS hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker (or the timed CPU baseline), never as the
product path.  It wraps the FP64 C restatement in mbrec_oracle.c and adds
independent numpy formulations that pin that restatement:

* ``dense_das`` / ``dense_sir``: the operator assembled from the hat-function
  definition max(0, 1 - |r - tau|/R) summed over all integers r (no ceil/floor
  tap bounds), i.e. a different formulation of forwardMex.cpp:118-159 /
  forwardDelayAndSumMex.cpp:73-101.
* ``prefilter_fft``: the reference's FFT form (Discretization.m:54-72 with
  shift_inv_proj.m:66-71) against the oracle's direct-DFT separable kernel.
* ``butter_zpk_textbook``: Butterworth design from the analog prototype and the
  bilinear transform with pre-warping (textbook), independent of scipy.signal.
* ``lsqr_krylov``: the LSQR iterate x_k is the least-squares minimiser over
  x0 + K_k(A'A, A'r0); solved densely on tiny problems.

Parity status of each oracle function is stated in its docstring.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmbrec_oracle.so")
_lib = None

_d = C.c_double
_i = C.c_int
_p = C.c_void_p


class OrcModel(C.Structure):
    _fields_ = [("Nz", _i), ("Nx", _i), ("D", _i), ("nT", _i), ("tt", _p), ("R", _d),
                ("use_sir", _i), ("sirIdx", _p), ("sir", _p), ("Ls", _i), ("se", _p),
                ("kz", _p), ("kx", _p), ("ir", _p), ("Lir", _i), ("ir_per_det", _i),
                ("sigma", _d)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_forward_das.argtypes = [_p, _i, _i, _p, _i, _i, _d, _p]
        L.orc_transpose_das.argtypes = [_p, _i, _i, _i, _p, _i, _d, _p]
        L.orc_forward_sir.argtypes = [_p, _i, _i, _p, _p, _i, _p, _i, _i, _p, _d, _p]
        L.orc_transpose_sir.argtypes = [_p, _i, _i, _i, _p, _p, _i, _p, _i, _p, _d, _p]
        L.orc_eir_conv.argtypes = [_p, _i, _i, _i, _p, _i, _i, _p]
        L.orc_eir_corr.argtypes = [_p, _i, _i, _i, _p, _i, _i, _p]
        L.orc_prefilter_kernel_1d.argtypes = [_p, _i, _p]
        L.orc_prefilter_apply.argtypes = [_p, _i, _i, _i, _p, _p, _i, _p]
        L.orc_model_forward.argtypes = [C.POINTER(OrcModel), _p, _i, _p]
        L.orc_model_transpose.argtypes = [C.POINTER(OrcModel), _p, _i, _p]
        L.orc_rec_nn_with_L2_reg.argtypes = [C.POINTER(OrcModel), _p, _i, _i, _d, _p, _p, _p]
        L.orc_rec_nn_with_L1_eye_reg.argtypes = [C.POINTER(OrcModel), _p, _i, _i, _d, _p, _p]
        L.orc_num_threads.restype = _i
        _lib = L
    return _lib


def _F(a, dt):
    return np.asfortranarray(a, dtype=dt)


# ---------------------------------------------------------------------------
# MEX loops (MATLAB-shaped arrays, column-major)
# ---------------------------------------------------------------------------
def forward_das(image, tt, nT, R):
    """forwardDelayAndSumMex.cpp:73-101 (out-of-range samples dropped).
    image [W x P], tt [P x D] -> [W x nT x D].  Pinned by dense_das and the
    closed-form single-pixel tests."""
    image, tt = _F(np.atleast_2d(image), np.float64), _F(tt, np.float64)
    W, P = image.shape
    D = tt.shape[1]
    out = np.zeros((W, nT, D), order="F")
    lib().orc_forward_das(image.ctypes.data, W, P, tt.ctypes.data, D, int(nT), float(R), out.ctypes.data)
    return out


def transpose_das(sig, tt, R):
    """transposeDelayAndSumMex.cpp:171-200.  sig [W x nT x D] -> [W x P]."""
    sig, tt = _F(sig, np.float64), _F(tt, np.float64)
    W, nT, D = sig.shape
    P = tt.shape[0]
    out = np.zeros((W, P), order="F")
    lib().orc_transpose_das(sig.ctypes.data, W, nT, D, tt.ctypes.data, P, float(R), out.ctypes.data)
    return out


def forward_sir(image, sidx, tt, sir, nT, se, R):
    """forwardMex.cpp:118-161.  Pinned by dense_sir and the SIR=delta reduction to DAS."""
    image, tt, sir = _F(np.atleast_2d(image), np.float64), _F(tt, np.float64), _F(sir, np.float64)
    sidx, se = _F(sidx, np.int32), _F(se, np.int32)
    W, P = image.shape
    D = tt.shape[1]
    out = np.zeros((W, nT, D), order="F")
    lib().orc_forward_sir(image.ctypes.data, W, P, sidx.ctypes.data, tt.ctypes.data, D, sir.ctypes.data,
                          sir.shape[0], int(nT), se.ctypes.data, float(R), out.ctypes.data)
    return out


def transpose_sir(sig, sidx, tt, sir, se, R):
    """transposeMex.cpp:277-320."""
    sig, tt, sir = _F(sig, np.float64), _F(tt, np.float64), _F(sir, np.float64)
    sidx, se = _F(sidx, np.int32), _F(se, np.int32)
    W, nT, D = sig.shape
    P = tt.shape[0]
    out = np.zeros((W, P), order="F")
    lib().orc_transpose_sir(sig.ctypes.data, W, nT, D, sidx.ctypes.data, tt.ctypes.data, P, sir.ctypes.data,
                            sir.shape[0], se.ctypes.data, float(R), out.ctypes.data)
    return out


# ---------------------------------------------------------------------------
# independent dense formulations (pins for the loops above)
# ---------------------------------------------------------------------------
def hat(r, tau, R):
    return np.maximum(0.0, 1.0 - np.abs(r - tau) / R)


def dense_das(tt, nT, R):
    """Dense [nT*D x P] matrix: A[d*nT + t, x] = max(0, 1 - |t - tau|/R), t in [0, nT)."""
    P, D = tt.shape
    t = np.arange(nT)[:, None]
    A = np.zeros((nT * D, P))
    for d in range(D):
        A[d * nT:(d + 1) * nT] = hat(t, tt[None, :, d], R)
    return A


def dense_sir(sidx, tt, sir, se, nT, R):
    """Dense [nT*D x P]: A[d*nT + t, x] = sum_s sir[s, k] * max(0, 1 - |t - s + Ls//2 - tau|/R)
    over the 0-based window s in [se[0,k]-1, se[1,k]-1] (flat column-major read of sir)."""
    P, D = tt.shape
    Ls = sir.shape[0]
    flat = np.asfortranarray(sir).ravel(order="F")
    A = np.zeros((nT * D, P))
    t = np.arange(nT)
    for d in range(D):
        for x in range(P):
            k = sidx[x, d] - 1
            for s in range(se[0, k] - 1, se[1, k]):
                A[d * nT + t, x] += flat[k * Ls + s] * hat(t - s + Ls // 2, tt[x, d], R)
    return A


# ---------------------------------------------------------------------------
# EIR, prefilter
# ---------------------------------------------------------------------------
def eir_conv(sino, ir, per_det=False):
    """MSOTmodel.m:259-263 as direct convolution; sino (W, D, nT) C-order."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    ir = np.ascontiguousarray(ir, dtype=np.float64)
    W, D, nT = sino.shape
    out = np.empty_like(sino)
    lib().orc_eir_conv(sino.ctypes.data, W, D, nT, ir.ctypes.data, ir.shape[-1], int(per_det), out.ctypes.data)
    return out


def eir_corr(sino, ir, per_det=False):
    """MSOTmodel.m:265-269 (flip(IR) convolution)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    ir = np.ascontiguousarray(ir, dtype=np.float64)
    W, D, nT = sino.shape
    out = np.empty_like(sino)
    lib().orc_eir_corr(sino.ctypes.data, W, D, nT, ir.ctypes.data, ir.shape[-1], int(per_det), out.ctypes.data)
    return out


def eir_conv_fft(sino, ir):
    """The reference's own FFT form: ifft(fft(s, L) .* fft(IR, L)), L = nT + Lir - 1, crop."""
    sino = np.asarray(sino, dtype=np.float64)
    nT, Lir = sino.shape[-1], ir.shape[-1]
    L = nT + Lir - 1
    f = np.real(np.fft.ifft(np.fft.fft(sino, L, axis=-1) * np.fft.fft(ir, L, axis=-1), axis=-1))
    h = Lir // 2
    return f[..., h:L - h]


def eir_corr_fft(sino, ir):
    return eir_conv_fft(sino, np.flip(ir, axis=-1))


def prefilter_kernel_1d(g_shifted):
    """Direct FP64 DFT: k = real(ifft(1 ./ fft(g))) (oracle restatement, mbrec_oracle.c)."""
    g = np.ascontiguousarray(g_shifted, dtype=np.float64)
    k = np.empty_like(g)
    lib().orc_prefilter_kernel_1d(g.ctypes.data, len(g), k.ctypes.data)
    return k


def prefilter_apply(img, kz, kx, transp=False):
    """Separable circular convolution; img (W, Nx, Nz) C-order."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    W, Nx, Nz = img.shape
    out = np.empty_like(img)
    kz = np.ascontiguousarray(kz, dtype=np.float64)
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    lib().orc_prefilter_apply(img.ctypes.data, W, Nz, Nx, kz.ctypes.data, kx.ctypes.data, int(transp), out.ctypes.data)
    return out


def gaussian_profile_shifted(N, h, sigma):
    """Centred grid h*(floor(-N/2) : floor(N/2-1)) (Discretization.m:80-81), Gaussian, fftshift."""
    pos = h * np.arange(np.floor(-N / 2), np.floor(N / 2 - 1) + 1)
    return np.fft.fftshift(np.exp(-pos ** 2 / (2 * sigma ** 2)))


def prefilter_fft(img, Nz, Nx, hz, hx, sigma, transp=False):
    """Reference FFT form of the prefilter with the 2-D Gaussian pixel and lattice [1 1]:
    prefilt = conj(phi_hat ./ fftn(autocorr)) (shift_inv_proj.m:66-71)."""
    lz = hz * np.arange(np.floor(-Nz / 2), np.floor(Nz / 2 - 1) + 1)
    lx = hx * np.arange(np.floor(-Nx / 2), np.floor(Nx / 2 - 1) + 1)
    xg, zg = np.meshgrid(lx, lz)
    phi = np.fft.fftshift(np.exp(-(xg ** 2 + zg ** 2) / (2 * sigma ** 2)))
    ph = np.fft.fft2(phi)
    ac = np.real(np.fft.ifft2(np.abs(ph) ** 2))
    pre = np.conj(ph / np.fft.fft2(ac))
    if transp:
        pre = np.conj(pre)
    a = np.swapaxes(np.asarray(img, dtype=np.float64).reshape(-1, Nx, Nz), 1, 2)
    out = np.real(np.fft.ifft2(np.fft.fft2(a, axes=(1, 2)) * pre[None], axes=(1, 2)))
    return np.swapaxes(out, 1, 2)


def butter_zpk_textbook(order, wn, btype):
    """Digital Butterworth zeros/poles/gain from the analog prototype
    p_k = exp(i pi (2k + n - 1) / (2n)), pre-warping w = 2 fs tan(pi wn / 2) with fs = 2,
    low->low / low->high frequency transform and the bilinear transform z = (2fs+s)/(2fs-s)."""
    n = order
    pa = np.exp(1j * np.pi * (2 * np.arange(1, n + 1) + n - 1) / (2 * n))
    fs2 = 4.0
    warped = fs2 * np.tan(np.pi * wn / 2.0)
    if btype == "low":
        p = pa * warped
        z = np.zeros(0)
        k = warped ** n
    else:
        p = warped / pa
        z = np.zeros(n)
        k = 1.0
    zd = (fs2 + z) / (fs2 - z)
    pd = (fs2 + p) / (fs2 - p)
    kd = k * np.real(np.prod(fs2 - z) / np.prod(fs2 - p))
    if btype == "low":
        zd = -np.ones(n)
    return zd, pd, kd


def zero_phase_response_textbook(n, fs, fmin, fmax):
    w = np.exp(2j * np.pi * np.arange(n) / n)
    H = np.ones(n, dtype=complex)
    for order, f, bt, on in ((8, fmax, "low", fmax > 0), (2, fmin, "high", fmin > 0)):
        if not on:
            continue
        z, p, k = butter_zpk_textbook(order, f * 2.0 / fs, bt)
        H *= k * np.prod(w[:, None] - z[None], axis=1) / np.prod(w[:, None] - p[None], axis=1)
    return np.abs(H) ** 2


# ---------------------------------------------------------------------------
# composed model and solvers (FP64 C)
# ---------------------------------------------------------------------------
class Model:
    """Holds the FP64 tables of one model for the C oracle.  tt (D, P) C-order
    == MATLAB [P x D]; sidx (D, P) int32; sir (Nsir, Ls) C-order == [Ls x Nsir];
    se (Nsir, 2) int32; kz/kx full circular kernels (length Nz, Nx) or None;
    ir (Lir,) or (D, Lir) or None."""

    def __init__(self, Nz, Nx, tt, nT, R, sigma=1.0, sidx=None, sir=None, se=None, kz=None, kx=None, ir=None):
        self._keep = []
        self.Nz, self.Nx, self.nT, self.R = Nz, Nx, nT, R
        self.tt = self._k(tt, np.float64)
        self.D = self.tt.shape[0]
        m = OrcModel()
        m.Nz, m.Nx, m.D, m.nT, m.R, m.sigma = Nz, Nx, self.D, nT, R, sigma
        m.tt = self.tt.ctypes.data
        if sir is not None:
            m.use_sir = 1
            m.sirIdx = self._k(sidx, np.int32).ctypes.data
            s = self._k(sir, np.float64)
            m.sir = s.ctypes.data
            m.Ls = s.shape[1]
            m.se = self._k(se, np.int32).ctypes.data
        self.kz = self.kx = self.ir = None
        if kz is not None:
            self.kz, self.kx = self._k(kz, np.float64), self._k(kx, np.float64)
            m.kz, m.kx = self.kz.ctypes.data, self.kx.ctypes.data
        if ir is not None:
            i = self.ir = self._k(ir, np.float64)
            m.ir = i.ctypes.data
            m.Lir = i.shape[-1]
            m.ir_per_det = 1 if i.ndim == 2 else 0
        self.m = m

    def _k(self, a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        self._keep.append(a)
        return a

    def forward(self, img):
        img = np.ascontiguousarray(img, dtype=np.float64)
        W = img.size // (self.Nz * self.Nx)
        out = np.empty((W, self.D, self.nT))
        lib().orc_model_forward(C.byref(self.m), img.ctypes.data, W, out.ctypes.data)
        return out

    def transpose(self, sino):
        sino = np.ascontiguousarray(sino, dtype=np.float64)
        W = sino.size // (self.D * self.nT)
        out = np.empty((W, self.Nx, self.Nz))
        lib().orc_model_transpose(C.byref(self.m), sino.ctypes.data, W, out.ctypes.data)
        return out

    def rec_nn_with_L2_reg(self, data, num_iter, lambdaL2=0.0):
        """rec_nn_with_L2_reg.m / nnls_L2_less.m with MATLAB-lsqr semantics (stacked W).
        Returns (rec (W, Nx, Nz), lcurve[2], counts[2])."""
        data = np.ascontiguousarray(data, dtype=np.float64)
        W = data.size // (self.D * self.nT)
        rec = np.empty((W, self.Nx, self.Nz))
        lc = np.zeros(2)
        cnt = np.zeros(2, dtype=np.int32)
        lib().orc_rec_nn_with_L2_reg(C.byref(self.m), data.ctypes.data, W, int(num_iter), float(lambdaL2),
                                     rec.ctypes.data, lc.ctypes.data, cnt.ctypes.data)
        return rec, lc, cnt

    def rec_nn_with_L1_eye_reg(self, data, num_iter, reg):
        data = np.ascontiguousarray(data, dtype=np.float64)
        W = data.size // (self.D * self.nT)
        rec = np.empty((W, self.Nx, self.Nz))
        cnt = np.zeros(2, dtype=np.int32)
        lib().orc_rec_nn_with_L1_eye_reg(C.byref(self.m), data.ctypes.data, W, int(num_iter), float(reg),
                                         rec.ctypes.data, cnt.ctypes.data)
        return rec, cnt


def full_kernels(Nz, Nx, hz, hx, sigma):
    """Full-length circular prefilter kernels from the oracle's direct DFT."""
    return (prefilter_kernel_1d(gaussian_profile_shifted(Nz, hz, sigma)),
            prefilter_kernel_1d(gaussian_profile_shifted(Nx, hx, sigma)))


def num_threads():
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------------------
# LSQR (numpy restatement of the same recurrences, for dense tiny problems)
# ---------------------------------------------------------------------------
def lsqr_matlab(A, b, tol, maxit, x0):
    """Same recurrences as orc_lsqr, on a dense matrix (used to cross-check the C code)."""
    n2b = np.linalg.norm(b)
    x = x0.astype(np.float64).copy()
    if n2b == 0:
        return np.zeros_like(x)
    tolb = tol * n2b
    u = b - A @ x
    beta = np.linalg.norm(u)
    normr = beta
    if beta != 0:
        u = u / beta
    c, s, phibar = 1.0, 0.0, beta
    v = A.T @ u
    alpha = np.linalg.norm(v)
    if alpha != 0:
        v = v / alpha
    d = np.zeros_like(x)
    normar = alpha * beta
    if normar == 0:
        return x
    norma = 0.0
    for _ in range(maxit):
        z = v
        u = A @ z - alpha * u
        beta = np.linalg.norm(u)
        u = u / beta
        norma = np.sqrt(norma ** 2 + alpha ** 2 + beta ** 2)
        thet = -s * alpha
        rhot = c * alpha
        rho = np.sqrt(rhot ** 2 + beta ** 2)
        c = rhot / rho
        s = -beta / rho
        phi = c * phibar
        phibar = s * phibar
        d = (z - thet * d) / rho
        x = x + phi * d
        normr = abs(s) * normr
        v = A.T @ u - beta * v
        alpha = np.linalg.norm(v)
        v = v / alpha
        normar = alpha * abs(s * phi)
        if normr <= tolb or normar / (norma * normr) <= tol:
            break
    return x


def lsqr_krylov(A, b, maxit, x0):
    """Exact LSQR iterate: argmin ||b - A x|| over x in x0 + span{(A'A)^j A' r0, j < maxit}."""
    r0 = b - A @ x0
    K = []
    q = A.T @ r0
    for _ in range(maxit):
        K.append(q)
        q = A.T @ (A @ q)
    Kb = np.stack(K, axis=1)
    coef, *_ = np.linalg.lstsq(A @ Kb, r0, rcond=None)
    return x0 + Kb @ coef
