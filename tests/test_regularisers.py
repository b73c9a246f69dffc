"""L2 terms of rec_nn_with_L2_reg beyond the identity: MATLAB del2 (the reference's
laplacian_per_wavelength, used as its own transpose) and caller-supplied operators.

The oracle (oracle.del2_matlab, oracle.rec_nn_with_L2_reg_regs) is pinned in
tests/test_oracle.py (cubic exactness of del2, identity case = the C oracle)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, rng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu(require_gpu):
    import mbrec
    mbrec.lib()
    yield


@pytest.mark.parametrize("Nx,Nz", [(20, 24), (3, 7), (2, 9), (5, 3), (1, 6)])
def test_del2_device_matches_oracle(Nx, Nz):
    import torch
    from mbrec.solvers import laplacian_per_wavelength
    x = rng(Nx * 31 + Nz).standard_normal((3, Nx, Nz))
    got = laplacian_per_wavelength(torch.tensor(x, dtype=torch.float32).cuda()).cpu().numpy()
    want = np.stack([O.del2_matlab(xi) for xi in x])
    np.testing.assert_allclose(got, want, atol=1e-5 * np.abs(x).max())


def _model():
    from test_gpu_operators import _geometry_model, _oracle_of
    m = _geometry_model(24, 20, 16, 300, 1.25e-4, sir=False, sigma=3.0)
    return m, _oracle_of(m)


def test_identity_plus_laplacian_matches_oracle():
    """reconstruction_overview_script.m:46-54: RegL2 = identity with lambdaL2 and RegL2_2 =
    laplacian_per_wavelength (also as RegL2T_2) with lambdaL2_2, two wavelengths stacked."""
    from mbrec.phantoms import add_noise, disks
    from mbrec.solvers import identity, laplacian_per_wavelength, rec_nn_with_L2_reg
    print("identity + laplacian", flush=True)
    m, orc = _model()
    ph = np.stack([disks(m.Discretization, n=4, seed=s) for s in (3, 4)])
    y = add_noise(orc.forward(ph), 0.01, seed=5)
    lam1, lam2 = 0.05, 0.5
    rec, e_img, e_r1, e_r2, *_ = rec_nn_with_L2_reg(m, y, 20, lam1, identity, identity, lam2,
                                                    laplacian_per_wavelength, laplacian_per_wavelength)
    ro, lc = O.rec_nn_with_L2_reg_regs(orc, y, 20, [(lam1, lambda a: a, lambda a: a),
                                                    (lam2, O.del2_matlab, O.del2_matlab)])
    assert rel_l2(rec, ro) < 1e-3
    for got, want in zip((e_img, e_r1, e_r2), lc):
        assert abs(got - want) <= 2e-3 * abs(want)
    # the laplacian term changes the result (it is not silently dropped)
    r1, *_ = rec_nn_with_L2_reg(m, y, 20, lam1, identity, identity)
    assert rel_l2(r1, ro) > 1e-2


def test_callback_terms_match_oracle():
    """Caller-supplied operators through the C-ABI callback: R = a one-pixel shift along z
    with its true adjoint as RegL2T, and a scaled identity as the second term."""
    import torch
    from mbrec.phantoms import add_noise, disks
    from mbrec.solvers import rec_nn_with_L2_reg
    print("callback terms", flush=True)
    m, orc = _model()
    ph = disks(m.Discretization, n=4, seed=6)[None]
    y = add_noise(orc.forward(ph), 0.01, seed=7)

    def shift(x):                       # (B, Nx, Nz) -> z shifted by one, zero filled
        out = torch.zeros_like(x)
        out[..., 1:] = x[..., :-1]
        return out

    def shift_t(x):
        out = torch.zeros_like(x)
        out[..., :-1] = x[..., 1:]
        return out
    lam1, lam2 = 0.2, 0.1
    rec, e_img, e_r1, e_r2, *_ = rec_nn_with_L2_reg(m, y, 15, lam1, shift, shift_t, lam2, lambda x: 2 * x,
                                                    lambda x: 2 * x)
    sh = lambda a: np.concatenate([np.zeros_like(a[:, :1]), a[:, :-1]], axis=1)
    sh_t = lambda a: np.concatenate([a[:, 1:], np.zeros_like(a[:, :1])], axis=1)
    ro, lc = O.rec_nn_with_L2_reg_regs(orc, y, 15, [(lam1, sh, sh_t), (lam2, lambda a: 2 * a, lambda a: 2 * a)])
    assert rel_l2(rec, ro) < 1e-3
    for got, want in zip((e_img, e_r1, e_r2), lc):
        assert abs(got - want) <= 2e-3 * abs(want)
