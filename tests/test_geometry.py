"""Travel-time tables (MSOTmodel.m:350-382): single speed of sound (straight paths) and
dual speed of sound (Fermat minimum over the membrane crossing), and the acoustic cache.

The oracle's straight-path table is pinned in test_oracle.py; its Fermat minimisation is
pinned here by Snell's law for a horizontal membrane, solved as a quartic in the crossing."""
import numpy as np
import pytest

import oracle as O
from helpers import rng


def _snell_crossing(xp, zp, xq, zq, m, cb, ca):
    """Crossing x of the refracted ray from p (below, z > m) to q (above, z < m) across the
    line z = m: u / (cb sqrt(u^2 + a^2)) = (L - u) / (ca sqrt((L - u)^2 + b^2)), u = x - xp,
    L = xq - xp, a = zp - m, b = m - zq, squared into a quartic in u."""
    a, b, L = zp - m, m - zq, xq - xp
    # ca^2 u^2 ((L - u)^2 + b^2) - cb^2 (L - u)^2 (u^2 + a^2) = 0
    p1 = np.polymul([ca ** 2, 0, 0], np.polyadd(np.polymul([-1, L], [-1, L]), [b ** 2]))
    p2 = np.polymul([cb ** 2], np.polymul(np.polymul([-1, L], [-1, L]), [1, 0, a ** 2]))
    roots = np.roots(np.polysub(p1, p2))
    ok = [r.real for r in roots if abs(r.imag) < 1e-9 and (r.real * L >= -1e-15) and abs(r.real) <= abs(L) + 1e-15]
    tt = lambda u: np.hypot(u, a) / cb + np.hypot(L - u, b) / ca
    u = min(ok, key=tt)
    return xp + u, tt(u)


def test_fermat_oracle_obeys_snell():
    m, cb, ca = -0.01, 1500.0, 1397.0
    mem = lambda x: m * np.ones_like(x)
    X = np.array([[0.0, 0.004], [-0.006, 0.01]])
    Z = np.array([[0.0, 0.003], [-0.002, 0.012]])
    xd, zd = O.arc_detectors(6, 0.06, 145.0)
    t, xi = O.travel_times_fermat(X, Z, xd, zd, cb, ca, mem, 0.06)
    for d in range(6):
        for k, (xp, zp) in enumerate(zip(X.ravel(), Z.ravel())):
            assert zp > m and zd[d] < m
            xs, ts = _snell_crossing(xp, zp, xd[d], zd[d], m, cb, ca)
            # scipy's bounded search stops at ~sqrt(eps) |x| in x; the time is stationary there
            assert abs(xi[d, k] - xs) < 1e-8 and abs(t[d, k] - ts) < 1e-15 + 1e-12 * ts


def test_acoustic_cache_roundtrip_and_collision(tmp_path, monkeypatch):
    """acoustic_file_<hash>: written on the first call, read on the second (no
    recomputation), and a stored file whose hashed fields differ is refused."""
    import os
    from mbrec import model as M
    from mbrec.discretization import Discretization
    from mbrec.probe import Probe
    calls = []

    def fake(disc, probe, c, single_sos=True):
        calls.append(1)
        D = probe.detector.numOfTransducers
        return np.full((D, disc.Nx, disc.Nz), 1e-5), np.full((D, disc.Nx, disc.Nz), np.nan)
    monkeypatch.setattr(M, "acoustic_data", fake)
    disc = Discretization([-3e-4, 3e-4], 7)
    probe = Probe("arc256_r40_174", True, False, num_detectors=8)
    t1, _ = M.load_or_compute_acoustic_data(disc, probe, 1500.0, True, str(tmp_path))
    t2, _ = M.load_or_compute_acoustic_data(disc, probe, 1500.0, True, str(tmp_path))
    assert len(calls) == 1 and np.array_equal(t1, t2)
    (f,) = [p for p in os.listdir(tmp_path) if p.startswith("acoustic_file_")]
    # a different medium: a different hash, so a new file
    M.load_or_compute_acoustic_data(disc, probe, 1510.0, True, str(tmp_path))
    assert len(calls) == 2 and len(os.listdir(tmp_path)) == 2
    # tamper with the stored fields of the first file: the hash check refuses it
    z = dict(np.load(tmp_path / f))
    z["hashed_Medium_speedOfSound"] = np.float64(1.0)
    np.savez(tmp_path / f, **z)
    with pytest.raises(AssertionError, match="Hash collision"):
        M.load_or_compute_acoustic_data(disc, probe, 1500.0, True, str(tmp_path))


@pytest.mark.gpu
def test_device_travel_times_single_and_dual(require_gpu):
    """Device tables against the oracle: straight paths for a single speed of sound (400^2 /
    256 detectors, the benchmark geometry) and the Fermat times and crossings for the
    example probe's dual speed of sound (a coarse grid)."""
    from mbrec.discretization import Discretization
    from mbrec.model import acoustic_data, travel_time_indices
    from mbrec.probe import Probe
    print("travel times", flush=True)
    disc = Discretization([-0.01995, 0.01995], 400)
    probe = Probe("arc256_r40_174", True, False)
    X, Z = disc.region["xPositionsFov"].T, disc.region["zPositionsFov"].T
    xd, zd = O.arc_detectors(256, 0.040, 174.0)
    tt = travel_time_indices(disc, probe, 1500.0, 3).reshape(256, -1)
    np.testing.assert_allclose(tt, O.travel_times(X, Z, xd, zd, 1500.0, 4e7, 0, 3), rtol=0, atol=1e-8)
    disc2 = Discretization([-0.012, 0.012, -0.015, 0.02], [9, 7])   # membrane at z = -0.01
    probe2 = Probe("example_system", True, False, num_detectors=12)
    t, xi = acoustic_data(disc2, probe2, 1540.0, single_sos=False)
    X2, Z2 = disc2.region["xPositionsFov"].T, disc2.region["zPositionsFov"].T
    to, xo = O.travel_times_fermat(X2, Z2, probe2.detector.xPositionsOfTransducers,
                                   probe2.detector.zPositionsOfTransducers, 1540.0, 1397.0,
                                   probe2.membrane, probe2.detector.radius)
    t, xi = t.reshape(12, -1), xi.reshape(12, -1)
    assert np.sum(np.isnan(xo)) > 0 and np.sum(~np.isnan(xo)) > 0     # both branches present
    np.testing.assert_array_equal(np.isnan(xi), np.isnan(xo))
    np.testing.assert_allclose(t * 4e7, to * 4e7, rtol=0, atol=1e-7)      # in samples
    ok = ~np.isnan(xo)
    np.testing.assert_allclose(xi[ok], xo[ok], rtol=0, atol=2e-8)
