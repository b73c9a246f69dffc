"""CPU tests that pin the FP64 oracle (no GPU).

Each oracle function is checked against something it does not share code with:
hand-derived known answers, a different formulation (dense hat-function
matrices, FFT forms, exact Krylov minimisers, textbook filter design) or a
mathematical invariant.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, random_sir, random_tt, rng

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


# --- known answers -----------------------------------------------------------
def test_das_single_pixel_known_answer():
    c = GOLD["das_single_pixel"]
    y = O.forward_das(np.array([[c["value"]]]), np.array([[c["tau"]]]), c["nT"], c["R"])
    want = np.zeros(c["nT"])
    for k, v in c["expect"].items():
        want[int(k)] = v
    np.testing.assert_allclose(y[0, :, 0], want, atol=1e-14)


def test_das_edge_samples_dropped():
    c = GOLD["das_edge_drop"]
    y = O.forward_das(np.array([[c["value"]]]), np.array([[c["tau"]]]), c["nT"], c["R"])
    want = np.array([c["expect"][str(i)] for i in range(c["nT"])])
    np.testing.assert_allclose(y[0, :, 0], want, atol=1e-14)


def test_sir_single_pixel_known_answer():
    c = GOLD["sir_single_pixel"]
    sir = np.array(c["sir"]).reshape(-1, 1)
    se = np.array(c["window"], dtype=np.int32).reshape(2, 1)
    y = O.forward_sir(np.array([[c["value"]]]), np.array([[1]], dtype=np.int32), np.array([[c["tau"]]]),
                      sir, c["nT"], se, c["R"])
    want = np.zeros(c["nT"])
    for k, v in c["expect"].items():
        want[int(k)] = v
    np.testing.assert_allclose(y[0, :, 0], want, atol=1e-13)


def test_das_transpose_known_answer():
    c = GOLD["das_transpose_single_detector"]
    sig = np.array(c["signal"], dtype=np.float64).reshape(1, -1, 1)
    x = O.transpose_das(sig, np.array([[c["tau"]]]), c["R"])
    assert abs(x[0, 0] - c["expect"]) < 1e-13


# --- different formulation: dense hat-function matrices ------------------------
@pytest.mark.parametrize("R", [2.6666666666666665, 3.3333333333333335, 1.6, 2.0, 0.75])
def test_das_matches_dense_hat_matrix(R):
    P, D, nT = 37, 5, 64
    tt = random_tt(P, D, nT, R, seed=1, spill=0.2)
    A = O.dense_das(tt, nT, R)
    x = rng(2).standard_normal(P)
    y = O.forward_das(x.reshape(1, P), tt, nT, R)
    np.testing.assert_allclose(y[0].T.ravel(), A @ x, rtol=1e-12, atol=1e-12)
    s = rng(3).standard_normal((1, nT, D))
    xt = O.transpose_das(s, tt, R)
    np.testing.assert_allclose(xt[0], A.T @ s[0].T.ravel(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("R", [2.6666666666666665, 1.6])
def test_sir_matches_dense_hat_matrix(R):
    P, D, nT = 23, 4, 48
    tt = random_tt(P, D, nT, R, seed=4, spill=0.2)
    sidx, sir, se = random_sir(P, D, Ls=9, Nsir=6, seed=5)
    A = O.dense_sir(sidx, tt, sir, se, nT, R)
    x = rng(6).standard_normal(P)
    y = O.forward_sir(x.reshape(1, P), sidx, tt, sir, nT, se, R)
    np.testing.assert_allclose(y[0].T.ravel(), A @ x, rtol=1e-12, atol=1e-12)
    s = rng(7).standard_normal((1, nT, D))
    xt = O.transpose_sir(s, sidx, tt, sir, se, R)
    np.testing.assert_allclose(xt[0], A.T @ s[0].T.ravel(), rtol=1e-12, atol=1e-12)


def test_sir_delta_reduces_to_das():
    """A unit tap at floor(Ls/2) with that single-tap window makes forwardMex == DAS."""
    P, D, nT, R, Ls = 50, 6, 80, 2.6666666666666665, 7
    tt = random_tt(P, D, nT, R, seed=8, spill=0.1)
    sir = np.zeros((Ls, 1)); sir[Ls // 2, 0] = 1.0
    se = np.array([[Ls // 2 + 1], [Ls // 2 + 1]], dtype=np.int32)
    sidx = np.ones((P, D), dtype=np.int32)
    x = rng(9).standard_normal((2, P))
    np.testing.assert_allclose(O.forward_sir(x, sidx, tt, sir, nT, se, R), O.forward_das(x, tt, nT, R), atol=1e-13)
    s = rng(10).standard_normal((2, nT, D))
    np.testing.assert_allclose(O.transpose_sir(s, sidx, tt, sir, se, R), O.transpose_das(s, tt, R), atol=1e-13)


def test_wavelength_batch_equals_single_calls():
    P, D, nT, R = 40, 5, 60, 2.5
    tt = random_tt(P, D, nT, R, seed=11)
    sidx, sir, se = random_sir(P, D, seed=12)
    x = rng(13).standard_normal((5, P))
    yb = O.forward_sir(x, sidx, tt, sir, nT, se, R)
    for w in range(5):
        np.testing.assert_allclose(yb[w], O.forward_sir(x[w:w + 1], sidx, tt, sir, nT, se, R)[0], atol=1e-13)


def test_adjoint_identity_fp64():
    P, D, nT, R = 60, 7, 90, 2.6666666666666665
    tt = random_tt(P, D, nT, R, seed=14, spill=0.1)
    sidx, sir, se = random_sir(P, D, seed=15)
    x = rng(16).standard_normal((3, P))
    s = rng(17).standard_normal((3, nT, D))
    for fwd, adj in ((lambda v: O.forward_das(v, tt, nT, R), lambda v: O.transpose_das(v, tt, R)),
                     (lambda v: O.forward_sir(v, sidx, tt, sir, nT, se, R),
                      lambda v: O.transpose_sir(v, sidx, tt, sir, se, R))):
        lhs = np.vdot(fwd(x), s)
        rhs = np.vdot(x, adj(s))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


# --- EIR -----------------------------------------------------------------------
def test_eir_conv_matches_numpy_and_fft_forms():
    g = rng(18)
    s = g.standard_normal((2, 3, 150))
    ir = g.standard_normal(201)
    y = O.eir_conv(s, ir)
    ref = np.stack([[np.convolve(s[w, d], ir)[100:250] for d in range(3)] for w in range(2)])
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(y, O.eir_conv_fft(s, ir), rtol=1e-10, atol=1e-10)
    z = O.eir_corr(s, ir)
    np.testing.assert_allclose(z, O.eir_corr_fft(s, ir), rtol=1e-10, atol=1e-10)
    t = g.standard_normal(s.shape)
    assert abs(np.vdot(O.eir_conv(t, ir), s) - np.vdot(t, z)) < 1e-10 * abs(np.vdot(t, z))


def test_eir_per_detector():
    g = rng(19)
    s = g.standard_normal((1, 4, 120))
    ir = g.standard_normal((4, 201))
    y = O.eir_conv(s, ir, per_det=True)
    for d in range(4):
        np.testing.assert_allclose(y[0, d], np.convolve(s[0, d], ir[d])[100:220], atol=1e-12)


# --- prefilter -----------------------------------------------------------------
@pytest.mark.parametrize("N,h,sig", [(40, 100e-6, 50e-6), (33, 125e-6, 62.5e-6), (16, 60e-6, 30e-6)])
def test_prefilter_kernel_dft_vs_fft(N, h, sig):
    g = O.gaussian_profile_shifted(N, h, sig)
    k = O.prefilter_kernel_1d(g)
    np.testing.assert_allclose(k, np.real(np.fft.ifft(1.0 / np.fft.fft(g))), rtol=1e-10, atol=1e-12)
    # deconvolution property: g (*) k = delta (circular)
    conv = np.real(np.fft.ifft(np.fft.fft(g) * np.fft.fft(k)))
    e = np.zeros(N); e[0] = 1
    np.testing.assert_allclose(conv, e, atol=1e-10)


@pytest.mark.parametrize("Nz,Nx", [(24, 20), (17, 22)])
def test_prefilter_separable_matches_reference_fft(Nz, Nx):
    h, sig = 100e-6, 50e-6
    kz, kx = O.full_kernels(Nz, Nx, h, h, sig)
    img = rng(20).standard_normal((2, Nx, Nz))
    for tr in (False, True):
        a = O.prefilter_apply(img, kz, kx, transp=tr)
        b = O.prefilter_fft(img, Nz, Nx, h, h, sig, transp=tr)
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("Nz,Nx", [(12, 10), (13, 11), (24, 24), (6, 7)])
def test_product_prefilter_taps_rebuild_the_circular_kernel(Nz, Nx):
    """The device prefilter's tap lists (offsets -K..K, the N/2 tap of an even axis kept
    once), placed circularly, give the oracle's full 2-D kernel (taps below 1e-10 of the
    largest dropped)."""
    from mbrec.discretization import Discretization
    h = 1.25e-4
    d = Discretization([-(Nx - 1) / 2 * h, (Nx - 1) / 2 * h, -(Nz - 1) / 2 * h, (Nz - 1) / 2 * h], [Nx, Nz])
    assert 2 * d.Kz <= Nz and 2 * d.Kx <= Nx

    def circ(t, K, N):
        c = np.zeros(N)
        for a in range(-K, K + 1):
            c[a % N] += t[a + K]
        return c
    kz, kx = O.full_kernels(Nz, Nx, d.dz, d.dx, d.pixelSigma)
    full = np.outer(kz, kx)
    np.testing.assert_allclose(np.outer(circ(d.kz, d.Kz, Nz), circ(d.kx, d.Kx, Nx)), full, rtol=0,
                               atol=1e-9 * np.abs(full).max())


# --- IR design -----------------------------------------------------------------
def test_butterworth_design_matches_textbook():
    from mbrec.irf import butter_zero_phase_response
    fs = 4e7
    for n in (201, 256):
        a = butter_zero_phase_response(n, fs, 1e5, 12e6)
        b = O.zero_phase_response_textbook(n, fs, 1e5, 12e6)
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)
    # |H|^2 of the 8th-order low-pass alone is 1/2 at its cut-off (fmax on a bin: n = 40)
    lp = O.zero_phase_response_textbook(40, fs, 0.0, 12e6)
    assert abs(lp[12] - 0.5) < 1e-9


def test_gauspuls_definition():
    from mbrec.irf import gauspuls_eir
    fs, fc = 4e7, 4e6
    e = gauspuls_eir(fs, fc, 1.0, -6.0)
    assert e.shape == (201,) and e[100] == 1.0
    # the -6 dB point of the envelope lies at f = fc*(1 +/- bw/2) in the spectrum
    spec = np.abs(np.fft.rfft(e, 1 << 16))
    f = np.fft.rfftfreq(1 << 16, 1 / fs)
    pk = spec.max()
    for fe in (fc * 0.5, fc * 1.5):
        v = np.interp(fe, f, spec) / pk
        assert abs(20 * np.log10(v) + 6.0) < 0.2


# --- geometry tables ---------------------------------------------------------
def test_travel_times_hand_values():
    """Three detectors on a 50 mm half circle (coverage 180 deg: at 0, 90, 180 deg, i.e.
    (0.05, 0), (0, -0.05), (-0.05, 0)), c = 1500 m/s, fs = 30 MHz (1 m = 2e4 samples),
    delay 10, 5 cropped samples.  Pixel (0, 0): 0.05 m -> 1000 - 15 = 985 for all three.
    Pixel (x, z) = (0.03, -0.01): to (0, -0.05) a 3-4-5 triangle, 0.05 m -> 985; to
    (0.05, 0) sqrt(0.0005) m -> 447.2135955 - 15; to (-0.05, 0) sqrt(0.0065) m -> 1612.4515497 - 15."""
    xd, zd = O.arc_detectors(3, 0.05, 180.0)
    np.testing.assert_allclose(xd, [0.05, 0.0, -0.05], atol=1e-15)
    np.testing.assert_allclose(zd, [0.0, -0.05, 0.0], atol=1e-15)
    tt = O.travel_times(np.array([[0.0], [0.03]]), np.array([[0.0], [-0.01]]), xd, zd, 1500.0, 3e7, 10, 5)
    np.testing.assert_allclose(tt[:, 0], [985.0, 985.0, 985.0], atol=1e-9)
    np.testing.assert_allclose(tt[:, 1], [447.2135955 - 15, 985.0, 1612.4515497 - 15], atol=1e-6)


def test_travel_times_match_independent_geometry():
    """The oracle's table on a 9 x 7 grid and 16 detectors equals the independent helper's
    (the product's device table is compared with it in test_geometry.py)."""
    from helpers import geometry_tt
    from mbrec.discretization import Discretization
    Nz, Nx, h, D = 9, 7, 1e-4, 16
    disc = Discretization([-(Nx - 1) / 2 * h, (Nx - 1) / 2 * h, -(Nz - 1) / 2 * h, (Nz - 1) / 2 * h], [Nx, Nz])
    X, Z = disc.region["xPositionsFov"].T, disc.region["zPositionsFov"].T
    xd, zd = O.arc_detectors(D, 0.040, 174.0)
    tt = O.travel_times(X, Z, xd, zd, 1500.0, 4e7, 7, 3)
    want = geometry_tt(Nz, Nx, h, D, 0.040, 174.0, 1500.0, 4e7, delay=10).reshape(D, -1)
    np.testing.assert_allclose(tt, want, rtol=0, atol=1e-9)


def test_spmr_polar_hand_values():
    """Detector at (0, -0.05): pixel (0.03, -0.01) is at 0.05 m and acos(0.8) (the 3-4-5
    angle) off the detector's inward direction; pixel (0, 0) at 0.05 m and angle 0."""
    dist, ang = O.spmr_polar(np.array([[0.03], [0.0]]), np.array([[-0.01], [0.0]]), np.array([0.0]),
                             np.array([-0.05]))
    np.testing.assert_allclose(dist[0], [0.05, 0.05], atol=1e-15)
    np.testing.assert_allclose(ang[0], [np.arccos(0.8), 0.0], atol=1e-7)
    from mbrec.sir import polar_coordinates
    d2, a2 = polar_coordinates(np.array([0.03, 0.0]), np.array([-0.01, 0.0]), np.array([0.0]), np.array([-0.05]))
    np.testing.assert_allclose(d2, dist, atol=1e-15)
    np.testing.assert_allclose(a2, ang, atol=1e-7)


def test_spmr_indices_hand_values():
    """dmin = 1.0, dd = 0.25, da = 0.5 deg, 40 angles (MSOTmodel.m:472-478):
    (dist 1.35, 3.6 deg):  round(1 + 1.4) = 2, 1 + round(7.2) = 8      -> (2-1)*40 + 8 = 48;
    (dist 1.625, 0 deg):   round(1 + 2.5) = 4 (half away from zero), 1 -> 3*40 + 1 = 121;
    (dist 0.975, 10.2 deg): round(1 - 0.1) = 1, 1 + round(20.4) = 21   -> 21."""
    from mbrec.sir import spmr_indices
    dist = np.array([1.35, 1.625, 0.975])
    ang = np.array([3.6, 0.0, 10.2]) * np.pi / 180
    want = [48, 121, 21]
    np.testing.assert_array_equal(O.spmr_indices(dist, ang, 1.0, 0.25, 0.5, 40), want)
    np.testing.assert_array_equal(spmr_indices(dist, ang, 1.0, 0.25, 0.5, 40), want)


# --- solvers ---------------------------------------------------------------------
def _tiny_model(seed=21, sir=False, Nz=8, Nx=7, D=5, nT=70):
    R = 2.6666666666666665
    P = Nz * Nx
    tt = random_tt(P, D, nT, R, seed=seed)
    kz, kx = O.full_kernels(Nz, Nx, 1e-4, 1e-4, 5e-5)
    ir = rng(seed + 1).standard_normal(9)
    kw = dict(kz=kz, kx=kx, ir=ir)
    if sir:
        sidx, s, se = random_sir(P, D, seed=seed + 2)
        kw.update(sidx=sidx.T, sir=s.T, se=se.T)
    return O.Model(Nz, Nx, tt.T, nT, R, sigma=3.0, **kw)


def _dense_of(m):
    P = m.Nz * m.Nx
    A = np.zeros((m.D * m.nT, P))
    e = np.zeros((1, m.Nx, m.Nz))
    for j in range(P):
        e.ravel()[:] = 0
        e.ravel()[j] = 1
        A[:, j] = m.forward(e).ravel()
    return A


def test_model_transpose_is_adjoint():
    for sir in (False, True):
        m = _tiny_model(sir=sir)
        x = rng(30).standard_normal((2, m.Nx, m.Nz))
        y = rng(31).standard_normal((2, m.D, m.nT))
        lhs, rhs = np.vdot(m.forward(x), y), np.vdot(x, m.transpose(y))
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_model_composition_matches_parts():
    """A = crop(IR * MEX(F p)) / sigma composed from the separately pinned parts."""
    m = _tiny_model(seed=40)
    x = rng(41).standard_normal((1, m.Nx, m.Nz))
    c = O.prefilter_apply(x, m.kz, m.kx)
    tt = m.tt.T
    s = O.forward_das(c.reshape(1, -1), tt, m.nT, m.R)          # [1 x nT x D]
    s = np.transpose(s, (0, 2, 1))                               # (1, D, nT)
    want = O.eir_conv(s, m.ir) / 3.0
    np.testing.assert_allclose(m.forward(x), want, rtol=1e-12, atol=1e-12)


def _nnls_L2_reference(A, y, nIter, lamL2):
    """nnls_L2_less with each LSQR call replaced by its exact Krylov minimiser."""
    m, n = A.shape
    Areg = np.vstack([A, np.sqrt(lamL2) * np.eye(n)]) if lamL2 > 0 else A
    b = np.concatenate([y, np.zeros(n)]) if lamL2 > 0 else y
    sb = np.sqrt(1e-3)
    Aaug = np.vstack([Areg, sb * np.eye(n)])
    w = np.zeros(n); v = np.zeros(n)
    for _ in range(nIter):
        baug = np.concatenate([b, sb * (w + v)])
        u = O.lsqr_krylov(Aaug, baug, 2, w)
        w = np.maximum(u - v, 0)
        v = v + w - u
    return w


@pytest.mark.parametrize("lamL2", [0.0, 0.05])
def test_nnls_L2_matches_exact_krylov_restatement(lamL2):
    m = _tiny_model(seed=50)
    A = _dense_of(m)
    xt = np.maximum(rng(51).standard_normal(A.shape[1]), 0)
    y = A @ xt + 0.01 * rng(52).standard_normal(A.shape[0])
    rec, lc, cnt = m.rec_nn_with_L2_reg(y.reshape(1, m.D, m.nT), 12, lamL2)
    want = _nnls_L2_reference(A, y, 12, lamL2)
    assert rel_l2(rec.ravel(), want) < 1e-8
    # 1 A'y start value + 12 x (3 A + 3 A') + 1 A for the L-curve term
    assert tuple(cnt) == (12 * 3, 12 * 3 + 1)
    assert abs(lc[0] - np.sum((A @ rec.ravel() - y) ** 2)) < 1e-9 * lc[0]


def test_lsqr_restatement_matches_krylov_and_scipy():
    from scipy.sparse.linalg import lsqr as sp_lsqr
    g = rng(60)
    A = g.standard_normal((30, 12)); b = g.standard_normal(30); x0 = g.standard_normal(12)
    for k in (1, 2, 3):
        a = O.lsqr_matlab(A, b, 1e-12, k, x0)
        np.testing.assert_allclose(a, O.lsqr_krylov(A, b, k, x0), rtol=1e-9, atol=1e-11)
        s = sp_lsqr(A, b, atol=0, btol=0, iter_lim=k, x0=x0)[0]
        np.testing.assert_allclose(a, s, rtol=1e-9, atol=1e-11)


def test_nnls_L2_fixed_point_is_nonnegative_least_squares():
    """After many outer iterations the splitting reaches the bound-constrained minimiser."""
    from scipy.optimize import lsq_linear
    m = _tiny_model(seed=70, Nz=5, Nx=4, D=4, nT=40)
    A = _dense_of(m)
    g = rng(71)
    y = A @ np.maximum(g.standard_normal(A.shape[1]), 0) + 0.05 * g.standard_normal(A.shape[0])
    lam = 0.02
    rec, _, _ = m.rec_nn_with_L2_reg(y.reshape(1, m.D, m.nT), 10000, lam)
    Aa = np.vstack([A, np.sqrt(lam) * np.eye(A.shape[1])])
    ya = np.concatenate([y, np.zeros(A.shape[1])])
    ref = lsq_linear(Aa, ya, bounds=(0, np.inf), tol=1e-14, lsmr_tol="auto", max_iter=10000).x
    assert rel_l2(rec.ravel(), ref) < 1e-3


def test_nnls_l2_minimiser_matches_lsq_linear_and_kkt():
    """The FP64 minimiser (accelerated projected gradient) against scipy's bounded
    least squares on the stacked system [A; sqrt(lambda) I] x ~ [y; 0], x >= 0, with the
    tiny model's dense operator, and its KKT conditions (gradient >= 0 where x = 0,
    = 0 where x > 0); two sinograms with their own lambda in one call."""
    from scipy.optimize import lsq_linear
    m = _tiny_model(seed=31)                                   # sigma = 3
    s0 = np.linalg.svd(_dense_of(m), compute_uv=False)[0]
    m1 = O.Model(m.Nz, m.Nx, m.tt, m.nT, m.R, sigma=3.0 * s0 / 0.999, kz=m.kz, kx=m.kx, ir=m.ir)
    A1 = _dense_of(m1)                                         # ||A1|| = 0.999, as a normalised model
    g = rng(32)
    y = g.standard_normal((2, m.D * m.nT))
    lam = np.array([0.05, 0.3])
    x, kkt = O.nnls_l2_minimiser(m1, y.reshape(2, m.D, m.nT), lam, max_iter=5000, tol=1e-11)
    assert kkt <= 1e-11
    P = m.Nz * m.Nx
    for b in range(2):
        S = np.vstack([A1, np.sqrt(lam[b]) * np.eye(P)])
        want = lsq_linear(S, np.concatenate([y[b], np.zeros(P)]), bounds=(0, np.inf), tol=1e-14,
                          lsmr_tol="auto", max_iter=10000).x
        xb = x[b].ravel()
        active = want < 1e-12 * want.max()
        assert 0 < active.sum() < P                            # some bounds active, some not
        np.testing.assert_allclose(xb, want, atol=1e-7 * np.abs(want).max())
        grad = A1.T @ (A1 @ xb - y[b]) + lam[b] * xb
        assert np.all(grad[xb == 0] >= -1e-8) and np.abs(grad[xb > 0]).max() < 1e-8


def test_nnls_L2_warm_start_oracle():
    """Warm-start form (w0, v0) of nnls_L2_less: zero start = the cold solve; the KKT pair
    (x*, g(x*)/beta on the zero set) of the bound-constrained minimiser is a fixed point;
    a general start equals the exact-Krylov restatement started from it."""
    from scipy.optimize import lsq_linear
    m = _tiny_model(seed=80, Nz=5, Nx=4, D=4, nT=40)
    A = _dense_of(m)
    n = A.shape[1]
    g = rng(81)
    y = A @ np.maximum(g.standard_normal(n), 0) + 0.05 * g.standard_normal(A.shape[0])
    lam = 0.02
    cold, _, _ = m.rec_nn_with_L2_reg(y.reshape(1, m.D, m.nT), 7, lam)
    np.testing.assert_allclose(m.rec_nn_with_L2_reg_from(y.reshape(1, m.D, m.nT), 7, [lam]), cold,
                               rtol=1e-12, atol=1e-14)
    Aa = np.vstack([A, np.sqrt(lam) * np.eye(n)])
    ya = np.concatenate([y, np.zeros(n)])
    xs = lsq_linear(Aa, ya, bounds=(0, np.inf), tol=1e-15, lsmr_tol="auto", max_iter=20000).x
    xs[xs < 1e-12 * xs.max()] = 0.0
    gr = A.T @ (A @ xs - y) + lam * xs
    vs = np.where(xs > 0, 0.0, np.maximum(gr, 0.0) / 1e-3)
    out = m.rec_nn_with_L2_reg_from(y.reshape(1, m.D, m.nT), 5, [lam], xs.reshape(1, -1), vs.reshape(1, -1))
    assert rel_l2(out.ravel(), xs) < 1e-6
    # general start: the restatement with the same (w0, v0)
    w0 = np.maximum(g.standard_normal(n), 0)
    v0 = g.standard_normal(n)
    sb = np.sqrt(1e-3)
    Aaug = np.vstack([Aa, sb * np.eye(n)])
    w, v = w0.copy(), v0.copy()
    for _ in range(6):
        u = O.lsqr_krylov(Aaug, np.concatenate([ya, sb * (w + v)]), 2, w)
        w = np.maximum(u - v, 0)
        v = v + w - u
    out = m.rec_nn_with_L2_reg_from(y.reshape(1, m.D, m.nT), 6, [lam], w0.reshape(1, -1), v0.reshape(1, -1))
    assert rel_l2(out.ravel(), w) < 1e-8


# --- L1 split Bregman with MATLAB-cgs inner solves ---------------------------------
def _cgs_by_cg_polynomial(M, b, k):
    """CGS iterates for symmetric M with shadow residual r0 = b, x0 = 0, from the identity
    r_k(CGS) = phi_k(M)^2 r0 where phi_k is CG's residual polynomial (Sonneveld 1989):
    phi_k is recovered from textbook CG and applied once more."""
    r0 = b.copy()
    xs, rs = [], []
    x = np.zeros_like(b); r = r0.copy(); p = r.copy()
    krylov = [r0]
    for _ in range(k):
        krylov.append(M @ krylov[-1])
    for j in range(1, k + 1):
        a = (r @ r) / (p @ (M @ p))
        x = x + a * p
        rn = r - a * (M @ p)
        p = rn + ((rn @ rn) / (r @ r)) * p
        r = rn
        # phi_j coefficients: r_CG = sum_i c_i M^i r0 with c_0 = 1
        K = np.stack(krylov[1:j + 1], axis=1)
        c, *_ = np.linalg.lstsq(K, r - r0, rcond=None)
        Kr = [r]
        for _ in range(j):
            Kr.append(M @ Kr[-1])
        phi_r = r + np.stack(Kr[1:], axis=1) @ c      # phi_j(M) r_CG = phi_j(M)^2 r0
        rs.append(phi_r)
        xs.append(np.linalg.solve(M, r0 - phi_r))
    return xs, rs


def _cgs_matlab_return(M, b, k):
    xs, rs = _cgs_by_cg_polynomial(M, b, k)
    norms = [np.linalg.norm(r) for r in rs]
    imin = int(np.argmin([np.linalg.norm(b)] + norms)) - 1      # -1 -> x0 = 0
    xmin = np.zeros_like(b) if imin < 0 else xs[imin]
    rcomp = np.linalg.norm(b - M @ xmin)
    return xmin if rcomp <= norms[-1] else xs[-1]


def _nnls_L1_reference(A, y, u0, eta, nIter):
    """nnls_L1_eye_matrix_less.m:24-55 with lambda = 1 and the CGS polynomial form."""
    n = A.shape[1]
    M = A.T @ A + np.eye(n)
    aty = A.T @ y
    b = np.zeros(n); u = u0.copy()
    for _ in range(nIter):
        d = np.maximum(u - b, 0)
        bin_ = d + b + aty - eta
        u = _cgs_matlab_return(M, bin_, 2)
        b = b + d - u
    return u


def test_cgs_polynomial_identity_matches_recurrence():
    g = rng(90)
    Q = g.standard_normal((15, 15)); M = Q @ Q.T + np.eye(15)
    b = g.standard_normal(15)
    xs, rs = _cgs_by_cg_polynomial(M, b, 2)
    # one CGS step by its recurrence: x1 = alpha (2 r0 - alpha M r0)
    a = (b @ b) / (b @ (M @ b))
    np.testing.assert_allclose(xs[0], a * (2 * b - a * (M @ b)), rtol=1e-10)
    for x, r in zip(xs, rs):
        np.testing.assert_allclose(M @ x, b - r, rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("eta", [0.0, 0.05])
def test_nnls_L1_matches_cg_polynomial_restatement(eta):
    m = _tiny_model(seed=80, Nz=5, Nx=4, D=4, nT=40)
    A = _dense_of(m)
    g = rng(81)
    y = A @ np.maximum(g.standard_normal(A.shape[1]), 0) + 0.05 * g.standard_normal(A.shape[0])
    rec, cnt = m.rec_nn_with_L1_eye_reg(y.reshape(1, m.D, m.nT), 15, eta)
    want = _nnls_L1_reference(A, y, A.T @ y, eta, 15)
    assert rel_l2(rec.ravel(), want) < 1e-8
    # 1 start A'y + per outer iteration: A'y, 2 x 2 products of (A'A + I), 1 check product
    assert tuple(cnt) == (15 * 5, 1 + 15 * 6)


# --- L2 regularisers ------------------------------------------------------------
def test_del2_exact_on_cubics_edges_included():
    """del2 reproduces (1/4) Laplacian of any cubic exactly on a uniform unit grid, at the
    edges too (the second differences of a cubic are linear, so their extrapolation is
    exact)."""
    Nx, Nz = 7, 9
    x = np.arange(Nx, dtype=float)[:, None]
    z = np.arange(Nz, dtype=float)[None, :]
    p = 0.3 * x ** 3 - 2 * x ** 2 * z + 0.5 * z ** 3 + x * z - 4 * z + 1.0
    lap = (1.8 * x - 4 * z) + (-0 + 3.0 * z)          # d2/dx2 + d2/dz2 of p
    np.testing.assert_allclose(O.del2_matlab(p), lap / 4, atol=1e-12)


def test_del2_three_point_and_short_dimensions():
    """n = 3 along z: the middle second difference is used at both ends; n = 2 along x:
    zero contribution.  U (Nx = 2, Nz = 3) = [[1, 4, 2], [0, 1, 5]]:
    z second differences halved: (2 - 8 + 1)/2 = -2.5 and (5 - 2 + 0)/2 = 1.5, so
    del2 = [[-2.5]*3, [1.5]*3] / 2."""
    U = np.array([[1.0, 4.0, 2.0], [0.0, 1.0, 5.0]])
    want = np.array([[-2.5, -2.5, -2.5], [1.5, 1.5, 1.5]]) / 2
    np.testing.assert_allclose(O.del2_matlab(U), want, atol=1e-15)


def test_regs_oracle_with_identity_matches_the_C_oracle():
    """The numpy stacked-regulariser loop with RegL2 = identity equals the C oracle's
    rec_nn_with_L2_reg (pinned above) -- same stacking, same lsqr recurrences."""
    m = _tiny_model(seed=41)
    y = rng(42).standard_normal((2, m.D, m.nT))
    rec, lc = O.rec_nn_with_L2_reg_regs(m, y, 6, [(0.3, lambda a: a, lambda a: a)])
    ro, lco, _ = m.rec_nn_with_L2_reg(y, 6, 0.3)
    assert rel_l2(rec, ro) < 1e-12
    assert abs(lc[0] - lco[0]) < 1e-9 * lco[0] and abs(lc[1] - lco[1]) < 1e-9 * lco[1]
