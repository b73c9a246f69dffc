"""Electrical / total impulse responses (setup-time tables).

* ``filter_butter_zero_phase`` restates load_and_preprocess/filter_butter_zero_phase.m:1-53
  (FFT branch): demean, 8th-order Butterworth low-pass and 2nd-order high-pass
  designed by bilinear transform, |H|^2 sampled at N circular bins, circular
  filtering along axis 0.
* ``impulse_response_without_sir`` / ``impulse_response_with_sir`` restate
  MSOTmodel.m:320-348.
* ``gauspuls_eir`` is the synthetic 201-tap EIR of the benchmark configurations
  (Gaussian-modulated sinusoid, MATLAB gauspuls definition).
"""
from __future__ import annotations

import numpy as np
from scipy import signal

EIR_LENGTH = 201  # Probe.m:13 (eirLength)


def gauspuls_eir(fs: float, fc: float = 4e6, bw: float = 1.0, bwr: float = -6.0,
                 length: int = EIR_LENGTH) -> np.ndarray:
    """Gaussian-modulated sinusoid centred at tap length//2 (gauspuls(t, fc, bw, bwr))."""
    t = (np.arange(length) - length // 2) / fs
    ref = 10.0 ** (bwr / 20.0)
    a = -((np.pi * fc * bw) ** 2) / (4.0 * np.log(ref))
    return np.exp(-a * t * t) * np.cos(2.0 * np.pi * fc * t)


def butter_zero_phase_response(n: int, fs: float, fmin: float, fmax: float) -> np.ndarray:
    """|H(e^{i 2 pi k/n})|^2 of the cascaded Butterworth filters, k = 0..n-1
    (filter_butter_zero_phase.m:23-45: butter / zp2sos, then freqz(..., n, 'whole', fs))."""
    zs, ps, k = [], [], 1.0
    if fmax > 0:
        z, p, kk = signal.butter(8, fmax * 2.0 / fs, btype="low", output="zpk")
        zs.append(z); ps.append(p); k *= kk
    if fmin > 0:
        z, p, kk = signal.butter(2, fmin * 2.0 / fs, btype="high", output="zpk")
        zs.append(z); ps.append(p); k *= kk
    z = np.concatenate(zs) if zs else np.zeros(0)
    p = np.concatenate(ps) if ps else np.zeros(0)
    _, h = signal.freqz_zpk(z, p, k, worN=n, whole=True)
    return np.abs(h) ** 2


def filter_butter_zero_phase(x: np.ndarray, fs: float, band, use_fft: bool = True) -> np.ndarray:
    """Zero-phase Butterworth band-pass along axis 0 (filter_butter_zero_phase.m:16-48)."""
    if not use_fft:
        raise NotImplementedError("only the FFT branch (use_fft=true) is provided")
    fmin, fmax = float(band[0]), float(band[1])
    x = np.asarray(x, dtype=np.float64)
    y = x - x.mean(axis=0, keepdims=True) if fmin > 0 else x.copy()
    n = x.shape[0]
    H2 = butter_zero_phase_response(n, fs, fmin, fmax)
    shape = (n,) + (1,) * (x.ndim - 1)
    return np.real(np.fft.ifft(H2.reshape(shape) * np.fft.fft(y, axis=0), axis=0))


def impulse_response_with_sir(eir: np.ndarray, fs: float, fmin: float, fmax: float) -> np.ndarray:
    """get_impulse_response_for_model_with_spmr (MSOTmodel.m:320-327): filtered EIR."""
    return filter_butter_zero_phase(eir, fs, [fmin, fmax], True)


def impulse_response_without_sir(eir: np.ndarray, fs: float, c: float, pixel_model,
                                 fmin: float, fmax: float) -> np.ndarray:
    """get_impulse_response_for_model_without_spmr (MSOTmodel.m:329-348):
    EIR convolved ('same') with the pixel-model N-shape phi(c t) * (-t), then filtered."""
    L = eir.shape[0]
    half = (L - 1) // 2
    t = np.arange(-half, half + 1) / fs
    pmr = pixel_model(t * c) * (-t)
    eir2 = eir.reshape(L, -1)
    full = np.stack([np.convolve(eir2[:, j], pmr, mode="full") for j in range(eir2.shape[1])], axis=1)
    start = len(pmr) // 2        # 'same': central part, full(floor(nv/2)+1 : floor(nv/2)+nu)
    ir = full[start:start + L, :]
    ir = ir.reshape(eir.shape)
    return filter_butter_zero_phase(ir, fs, [fmin, fmax], True)
