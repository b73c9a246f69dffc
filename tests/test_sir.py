"""SIR (SPMR) tables: the generator that replaces Field II, the polar grid and the
polarSPMR file.

CPU tests pin the numpy generator in the oracle: one sub-element at the element centre
gives the N-shape itself; the tap sum of an untrimmed entry is the N-shape's sum; the
trimmed window brackets exactly the non-zero samples.  GPU tests compare the device
generator and grid with it, and the file round trip."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2

C_, FS = 1500.0, 4e7


def _pixel_model(sig=5e-5):
    return lambda r: np.exp(-np.asarray(r) ** 2 / (2 * sig ** 2))


def test_one_sub_element_gives_the_nshape():
    pm = _pixel_model()
    tab, se = O.spmr_table_numpy([0.03, 0.05], [0.0, 10.0, 40.0], 0.5e-3, 1, C_, FS, pm, trim=0.0)
    kn = np.arange(-45, 46)
    nsh = pm(C_ * kn / FS) * (-kn / FS)
    nsh[np.abs(nsh) <= 1e-12 * np.abs(nsh).max()] = 0.0
    for row in tab:
        np.testing.assert_allclose(row, nsh, atol=1e-15 * np.abs(nsh).max())


def test_tap_sum_is_conserved_and_window_rule():
    pm = _pixel_model()
    kn = np.arange(-45, 46)
    nsh = pm(C_ * kn / FS) * (-kn / FS)
    nsh[np.abs(nsh) <= 1e-12 * np.abs(nsh).max()] = 0.0
    dists, angs = np.linspace(0.02, 0.06, 5), np.linspace(0, 60, 7)
    raw, _ = O.spmr_table_numpy(dists, angs, 0.5e-3, 8, C_, FS, pm, trim=0.0)
    # every sub-element's hat weights sum to 1 / n_sub: the entry sums to the N-shape's sum
    # times the hat mass, when the N-shape shifted by the hats stays on the grid
    np.testing.assert_allclose(raw.sum(axis=1), nsh.sum(), atol=1e-12 * np.abs(nsh).sum())
    tab, se = O.spmr_table_numpy(dists, angs, 0.5e-3, 8, C_, FS, pm)
    for row, (a, b) in zip(tab, se):
        nz = np.nonzero(row)[0]
        assert (a, b) == (nz[0] + 1, nz[-1] + 1)
        assert np.all(np.abs(row[nz]) >= 1e-4 * np.abs(row).max())


@pytest.fixture
def _gpu(require_gpu):
    import mbrec
    mbrec.lib()


@pytest.mark.gpu
def test_device_table_matches_oracle(_gpu):
    from mbrec.sir import spmr_table
    print("SIR table", flush=True)
    pm = _pixel_model()
    dists, angs = np.linspace(0.02, 0.062, 60), np.linspace(0, 85, 40)
    tab, se = spmr_table(dists, angs, 0.5e-3, 8, C_, FS, pm)
    to, so = O.spmr_table_numpy(dists, angs, 0.5e-3, 8, C_, FS, pm)
    assert rel_l2(tab, to) < 1e-12
    np.testing.assert_array_equal(se, so)


@pytest.mark.gpu
def test_device_grid_and_indices_match_oracle(_gpu):
    from mbrec.discretization import Discretization
    from mbrec.sir import device_indices, required_grid
    print("SIR grid", flush=True)
    disc = Discretization([-0.01, 0.01], 81)
    X, Z = disc.region["xPositionsFov"].T.ravel(), disc.region["zPositionsFov"].T.ravel()
    xd, zd = O.arc_detectors(64, 0.040, 174.0)
    dmin, dmax, dd, amax, da = required_grid(X, Z, xd, zd, 2.5e-4, C_)
    dist, ang = O.spmr_polar(X, Z, xd, zd)
    # MSOTmodel.m:599-616
    dd_o = 2.5e-4 * 1450.0 / C_
    assert abs(dd - dd_o) < 1e-18 and abs(dmin - (dist.min() - dd_o)) < 1e-15
    assert abs(dmax - (dist.max() + dd_o)) < 1e-15
    da_o = np.arctan2(dd_o, dmax) / np.pi * 180
    assert abs(da - da_o) < 1e-14 and abs(amax - np.ceil(ang.max() / np.pi * 180 * da_o) / da_o) < 1e-9
    n_angle = int(np.ceil(amax / da)) + 1
    got = device_indices(X, Z, xd, zd, dmin, dd, da, n_angle)
    np.testing.assert_array_equal(got, O.spmr_indices(dist, ang, dmin, dd, da, n_angle))


@pytest.mark.gpu
def test_polar_spmr_file_roundtrip(_gpu, tmp_path):
    """First model: no file, the table is generated and written; second model with the same
    grid: the file is read (same table and indices); a model needing a larger grid: the
    union of both grids is generated and the file rewritten."""
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe
    print("polarSPMR file", flush=True)
    probe = Probe("arc256_r40_174", True, False, num_detectors=32)
    d1 = Discretization([-0.004, 0.004], 41)
    m1 = MSOTmodel(probe, d1, C_, True, True, 0, 1e5, 12e6, 1.0, spmr_dir=str(tmp_path))
    assert m1.Data.SPMRGrid["loaded"] is False
    files = list(tmp_path.iterdir())
    assert len(files) == 1 and files[0].name.startswith("polarSPMR_arc256_r40_174_c_1500_pixelModel_")
    m2 = MSOTmodel(probe, d1, C_, True, True, 0, 1e5, 12e6, 1.0, spmr_dir=str(tmp_path))
    assert m2.Data.SPMRGrid["loaded"] is True
    np.testing.assert_array_equal(m2.Data.SPMR, m1.Data.SPMR)
    np.testing.assert_array_equal(m2.Data.SPMRIndices, m1.Data.SPMRIndices)
    np.testing.assert_array_equal(m2.Data.SPMRTimestepIndexStartAndEnd, m1.Data.SPMRTimestepIndexStartAndEnd)
    d3 = Discretization([-0.008, 0.008], 81)      # farther pixels: a larger distance / angle range
    m3 = MSOTmodel(probe, d3, C_, True, True, 0, 1e5, 12e6, 1.0, spmr_dir=str(tmp_path))
    assert m3.Data.SPMRGrid["loaded"] is False
    m4 = MSOTmodel(probe, d1, C_, True, True, 0, 1e5, 12e6, 1.0, spmr_dir=str(tmp_path))
    assert m4.Data.SPMRGrid["loaded"] is True            # the rewritten file covers the small grid too
