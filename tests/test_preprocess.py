"""Sinogram preprocessing (load_and_preprocess/): zero-phase Butterworth filter,
Butterworth window, broken-transducer spline fill.

CPU tests pin the oracle (FFT filtering = circular convolution with ifft(|H|^2) and the
textbook design pinned against scipy in test_oracle.py; the window's documented 1/sqrt(2)
at its cut-offs; cubic reproduction of the spline fill); GPU tests compare the device
entry points with it (FP64 both sides)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, rng

FS = 4e7


def test_zero_phase_filter_is_circular_convolution_with_ifft_of_H2():
    x = rng(1).standard_normal((3, 64))
    y = O.filter_butter_zero_phase(x, FS, 1e5, 12e6)
    H2 = O.zero_phase_response_textbook(64, FS, 1e5, 12e6)
    k = np.real(np.fft.ifft(H2))
    xd = x - x.mean(axis=-1, keepdims=True)
    direct = np.array([[sum(k[(t - m) % 64] * xd[c, m] for m in range(64)) for t in range(64)] for c in range(3)])
    np.testing.assert_allclose(y, direct, atol=1e-12)
    # no high-pass edge: no demeaning, the mean passes with |H(0)|^2 = 1
    y0 = O.filter_butter_zero_phase(x + 5.0, FS, 0.0, 12e6)
    np.testing.assert_allclose(y0.mean(axis=-1), (x + 5.0).mean(axis=-1), rtol=1e-9)


def test_window_is_one_over_sqrt2_at_its_cutoffs():
    """The window is the band-pass magnitude: 1/sqrt(2) at the cut-offs low / N and high / N
    of the normalised frequency (apply_butterworth_window_to_sinogram.m:2-4), < 1 elsewhere."""
    from scipy import signal
    N, n, lo, hi = 2030, 4, 30.0, 1700.0
    w = O.butterworth_window(N, n, lo, hi)
    z, p, k = signal.butter(n, [lo / N, hi / N], btype="bandpass", output="zpk")
    for f in (lo, hi):
        _, h = signal.freqz_zpk(z, p, k, worN=[np.pi * f / N])
        assert abs(abs(h[0]) - 1 / np.sqrt(2)) < 1e-9
    assert w.max() <= 1 + 1e-12 and w[0] < 1e-6
    assert abs(w[int((lo + hi) / 2)] - 1) < 1e-3


def test_broken_fill_reproduces_cubics_and_keeps_working_detectors():
    D, nT = 24, 5
    d = np.arange(1, D + 1, dtype=float)[:, None]
    t = np.arange(nT, dtype=float)[None, :]
    x = 0.01 * d ** 3 - 0.3 * d ** 2 * (1 + t) + 2 * d - t
    broken = [1, 7, 8, 24]
    xb = x.copy()
    xb[[0, 6, 7, 23]] = 999.0
    y = O.interpolate_broken_transducers(xb[None], broken)[0]
    np.testing.assert_allclose(y, x, rtol=1e-10, atol=1e-9)


@pytest.mark.gpu
def test_device_zero_phase_filter_matches_oracle(require_gpu):
    from mbrec.preprocess import filter_butter_zero_phase
    print("zero-phase filter", flush=True)
    x = rng(2).standard_normal((2, 256, 2030)) + 0.3
    for band in ((1e5, 12e6), (0.0, 8e6), (5e5, 0.0)):
        got = filter_butter_zero_phase(x, FS, band)
        want = O.filter_butter_zero_phase(x, FS, *band)
        assert rel_l2(got, want) < 1e-10, band


@pytest.mark.gpu
def test_device_butterworth_window_matches_oracle(require_gpu):
    from mbrec.preprocess import apply_butterworth_window_to_sinogram
    print("butterworth window", flush=True)
    x = rng(3).standard_normal((2, 64, 2030)) + 1.0
    for n, lo, hi in ((4, 30.0, 1700.0), (2, 10.0, 1000.0), (6, 100.0, 1900.0)):
        got = apply_butterworth_window_to_sinogram(x, n, lo, hi)
        want = O.apply_butterworth_window(x, n, lo, hi)
        assert rel_l2(got, want) < 1e-10, (n, lo, hi)


@pytest.mark.gpu
def test_device_broken_fill_matches_oracle(require_gpu):
    from mbrec.preprocess import interpolate_signals_of_broken_transducers
    print("broken transducers", flush=True)
    x = rng(4).standard_normal((3, 256, 300))
    for broken in ([5], [1, 2, 100, 256], list(range(40, 60))):
        got = interpolate_signals_of_broken_transducers(x, broken)
        want = O.interpolate_broken_transducers(x, broken)
        assert rel_l2(got, want) < 1e-10, broken
