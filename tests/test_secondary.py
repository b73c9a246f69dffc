"""Secondary paths: dmas_cf, backproject_waveeq, bp_kernel, upsample_detectors_in_sinogram.

CPU tests pin the numpy oracle (identities, library routines); GPU tests compare the
device implementation with it (FP64 both sides: 1e-10 relative, 1e-8 for the GEMM paths).
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, rng

C, FS, COV, RAD = 1500.0, 4e7, 174.0, 0.040


def _grid(N, h):
    v = (np.arange(N) - (N - 1) / 2) * h
    X, Z = np.meshgrid(v, v)           # MATLAB [Nz x Nx]
    return X.T.copy(), Z.T.copy()      # (Nx, Nz)


def _cropped(nT):
    return int(round(RAD / C * FS)) - nT // 2


# --- CPU pins ------------------------------------------------------------------------
def test_dmas_identity_and_cf_bound():
    D, nT = 9, 300
    X, Z = _grid(6, 1e-4)
    sig = rng(1).standard_normal((1, D, nT))
    out = O.dmas_cf(sig, X, Z, C, FS, _cropped(nT), COV, RAD)
    # brute force sum_{i<j} s^_i s^_j for one pixel
    s = sig / np.abs(sig).max()
    sq = np.sign(s) * np.sqrt(np.abs(s))
    ang = -((180 - COV) / 2 + np.arange(D) * COV / (D - 1)) * np.pi / 180
    x, z = X[2, 3], Z[2, 3]
    vals = []
    for p in range(D):
        T = int(O.matlab_round((np.hypot(x - np.cos(ang[p]) * RAD, z - np.sin(ang[p]) * RAD) / C
                                 - (1 + _cropped(nT)) / FS) * FS + 1))
        vals.append(sq[0, p, T - 1])
    brute = sum(vals[i] * vals[j] for i in range(D) for j in range(i + 1, D))
    assert abs(out[1, 0, 2, 3] - brute) < 1e-12 * max(1, abs(brute))
    assert np.all(out[2] >= 0) and np.all(out[2] <= 1)
    np.testing.assert_allclose(out[0], out[1] * out[2], rtol=1e-14)


def test_bp_kernel_antisymmetric_and_gemm_form():
    t = (np.arange(1, 41) + 3) / FS
    K = O.bp_kernel(t)
    np.testing.assert_allclose(K, -K.T, atol=1e-12 * np.abs(K).max())
    from scipy.special import jv, yv
    dl = np.pi / (t[-1] - t[0])
    lam = dl * np.arange(1, t.size + 1)
    w = np.full(t.size, dl); w[0] = w[-1] = dl / 2
    A = jv(0, np.outer(t, lam)) * (w * lam)
    B = yv(0, np.outer(t, lam))
    M = A @ B.T
    np.testing.assert_allclose(K, M - M.T, rtol=1e-9, atol=1e-9 * np.abs(K).max())


def test_divergence_of_linear_field_is_exact():
    X, Z = _grid(7, 2e-4)
    U, V = 3.0 * X + 1.0, -0.5 * Z + 2.0
    np.testing.assert_allclose(O.divergence(X[:, 0], Z[0, :], U, V), 2.5, atol=1e-9)


def test_spline_matrix_reproduces_cubics():
    """A not-a-knot spline reproduces any cubic exactly (including extrapolation)."""
    D, f = 16, 2
    U = O.not_a_knot_spline_matrix(D, f)
    knots = np.arange(1, D * f + 1, f, dtype=float)
    q = np.arange(1, D * f + 1, dtype=float)
    cub = lambda x: 0.3 * x ** 3 - 2 * x ** 2 + x - 5
    np.testing.assert_allclose(U @ cub(knots), cub(q), rtol=1e-10, atol=1e-8)


def test_bp_radial_integrals_match_cumulative_trapezoid():
    """The time integral of backproject_waveeq.m:24-25 is scipy's cumulative trapezoid
    (initial value 0) divided by the acquisition time of each sample."""
    from scipy.integrate import cumulative_trapezoid
    sig = rng(3).standard_normal((2, 5, 60))
    tsec = (np.arange(1, 61) + 7) / FS
    want = cumulative_trapezoid(sig / C, tsec, axis=-1, initial=0) / tsec
    np.testing.assert_allclose(O.bp_radial_integrals(sig, tsec, C), want, rtol=1e-13, atol=0)


def test_bp_psf_same_matches_convolve2d_for_odd_kernels():
    from scipy.signal import convolve2d
    S = rng(31).standard_normal((2, 7, 40))            # (W, D, nT)
    psf = rng(32).uniform(0.1, 1, (3, 5))              # (time, detector)
    got = O.bp_psf_same(S, psf)
    for w in range(2):
        want = convolve2d(S[w].T, psf, mode="same") / psf.sum()
        np.testing.assert_allclose(got[w].T, want, rtol=1e-12, atol=1e-14)


def test_bp_psf_same_matlab_centring_for_even_kernels():
    """MATLAB's documented conv example: conv([-1 2 3 -2 0 1 2], [2 4 -1 1], 'same') =
    [15 5 -9 7 6 7 -1] (the full result from index floor(4/2) + 1); scipy's 'same' starts
    one sample earlier for even kernels.  Along time and along detectors."""
    u = np.array([-1.0, 2, 3, -2, 0, 1, 2])
    v = np.array([2.0, 4, -1, 1])
    want = np.array([15.0, 5, -9, 7, 6, 7, -1]) / v.sum()
    np.testing.assert_allclose(O.bp_psf_same(u[None, None, :], v[:, None])[0, 0], want, atol=1e-14)
    np.testing.assert_allclose(O.bp_psf_same(u[None, :, None], v[None, :])[0, :, 0], want, atol=1e-14)


def _brute_backproject(sig, fx, fy, cropped, K):
    """backproject_waveeq.m with psf = 1, written as scalar loops: trapezoid sums, the
    K(t, t) S dt product element by element, the nearest-sample gather per pixel and
    detector, MATLAB gradient (central inside, one-sided at the edges), fliplr."""
    import math
    W, D, nT = sig.shape
    dt = 1.0 / FS
    off = (180.0 - COV) / 2
    th = [-(off + p * COV / (D - 1)) * math.pi / 180 for p in range(D)]
    t = [(n + 1 + cropped) * dt for n in range(nT)]
    Nx, Nz = fx.shape
    xs, zs = fx[:, 0], fy[0, :]
    out = np.zeros((W, Nx, Nz))
    for w in range(W):
        S = np.zeros((nT, D))
        for p in range(D):
            acc = 0.0
            for n in range(nT):
                if n:
                    acc += (t[n] - t[n - 1]) * (sig[w, p, n - 1] + sig[w, p, n]) / (2 * C)
                S[n, p] = acc / t[n]
        Swl = np.zeros((nT, D))
        for i in range(nT):
            for p in range(D):
                Swl[i, p] = sum(K[i + cropped, j + cropped] * S[j, p] for j in range(nT)) * dt
        px = np.zeros((Nx, Nz)); py = np.zeros((Nx, Nz))
        for ix in range(Nx):
            for iz in range(Nz):
                for p in range(D):
                    d = math.hypot(fx[ix, iz] - math.cos(th[p]) * RAD, fy[ix, iz] - math.sin(th[p]) * RAD)
                    v = (d / C - t[0]) * FS + 1
                    T = int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))
                    px[ix, iz] += Swl[T - 1, p] * math.cos(th[p]) * dt * C
                    py[ix, iz] += Swl[T - 1, p] * math.sin(th[p]) * dt * C

        def grad(f, coord, i, n, get):
            if i == 0:
                return (get(1) - get(0)) / (coord[1] - coord[0])
            if i == n - 1:
                return (get(n - 1) - get(n - 2)) / (coord[n - 1] - coord[n - 2])
            return (get(i + 1) - get(i - 1)) / (coord[i + 1] - coord[i - 1])
        for ix in range(Nx):
            for iz in range(Nz):
                dU = grad(px, xs, ix, Nx, lambda k: px[k, iz])
                dV = grad(py, zs, iz, Nz, lambda k: py[ix, k])
                out[w, Nx - 1 - ix, iz] = dU + dV
    return out


def test_backproject_waveeq_matches_brute_force():
    """Oracle composition (identity PSF, a random K of Nt > nT + cropped) against the scalar
    loop form: pins the K(t, t) block, the gather, the divergence and the fliplr."""
    D, nT, N = 7, 40, 5
    cropped = int(round(RAD / C * FS)) - nT // 2
    X, Z = _grid(N, 2e-4)
    sig = rng(33).standard_normal((2, D, nT))
    Nt = nT + cropped + 3
    K = rng(34).standard_normal((Nt, Nt))
    got = O.backproject_waveeq(sig, X, Z, C, FS, cropped, COV, RAD, np.ones((1, 1)), K)
    want = _brute_backproject(sig, X, Z, cropped, K)
    assert np.abs(want).max() > 0
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


# --- GPU parity ----------------------------------------------------------------------
@pytest.mark.gpu
def test_dmas_cf_matches_oracle(require_gpu):
    from mbrec.backprojection import dmas_cf
    print("dmas_cf parity", flush=True)
    D, nT = 64, 400
    X, Z = _grid(48, 1e-4)
    sig = rng(4).standard_normal((2, D, nT)).astype(np.float32)
    got = dmas_cf(sig, X, Z, C, FS, _cropped(nT), COV, RAD)
    want = O.dmas_cf(sig.astype(np.float64), X, Z, C, FS, _cropped(nT), COV, RAD)
    for k in range(4):
        assert rel_l2(got[k], want[k]) < 1e-10, k


@pytest.mark.gpu
def test_dmas_cf_index_error(require_gpu):
    from mbrec import MbrecError
    from mbrec.backprojection import dmas_cf
    X, Z = _grid(8, 1e-4)
    with pytest.raises(MbrecError) as ei:
        dmas_cf(np.ones((4, 50), np.float32), X, Z, C, FS, 0, COV, RAD)
    assert ei.value.identifier == "dmas_cf:indexOutOfRange"


@pytest.mark.gpu
def test_bp_kernel_matches_oracle(require_gpu):
    from mbrec.backprojection import bp_kernel
    print("bp_kernel parity", flush=True)
    t = (np.arange(1, 97) + 5) / FS
    assert rel_l2(bp_kernel(t), O.bp_kernel(t)) < 1e-8


@pytest.mark.gpu
def test_backproject_waveeq_matches_oracle(require_gpu):
    from mbrec.backprojection import backproject_waveeq, bp_kernel
    print("backproject parity", flush=True)
    D, nT = 32, 160
    cropped = _cropped(nT)
    X, Z = _grid(24, 1e-4)
    sig = rng(5).standard_normal((2, D, nT))
    Nt = nT + cropped
    K = rng(6).standard_normal((Nt, Nt))
    psf = np.outer(np.hanning(5), np.hanning(3))
    got = backproject_waveeq(sig, X, Z, C, FS, cropped, COV, RAD, psf, K)
    want = O.backproject_waveeq(sig, X, Z, C, FS, cropped, COV, RAD, psf, K)
    assert rel_l2(got, want) < 1e-8


@pytest.mark.gpu
def test_upsample_detectors_matches_spline(require_gpu):
    from mbrec.preprocess import upsample_detectors_in_sinogram
    print("upsample parity", flush=True)
    s = rng(7).standard_normal((2, 3, 256, 120))
    got = upsample_detectors_in_sinogram(s, 2)
    U = O.not_a_knot_spline_matrix(256, 2)
    want = np.einsum("ud,fwdt->fwut", U, s)
    assert got.shape == (2, 3, 512, 120)
    assert rel_l2(got, want) < 1e-12
    # detectors at the knots are reproduced exactly
    np.testing.assert_allclose(got[..., ::2, :], s, atol=1e-12)
