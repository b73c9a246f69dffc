"""TV path: split-Bregman TV prox, DCT Neumann solver, ITV, SpaRSA, and the device TV solver.

Pins: SB_ITV2 computes the isotropic-TV prox (checked against a dual projected-gradient
solution run to convergence, an independent algorithm); the DCT solve inverts I + B'B
exactly; ITV matches its definition.  SpaRSA's step, its Barzilai-Borwein rule and its
initial alpha = 1 are pinned by a trace worked by hand on a two-variable L1 problem, by the
one-step answer for A = I, and by convergence to minimisers computed by independent
algorithms (L-BFGS-B on a split L1 problem, Chambolle-Pock on a TV deblurring problem).
SpaRSA's MiniterA default (5) and its counter convention come from its documentation
(the package is not vendored): the iteration counts below test the rule as stated.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from helpers import rel_l2, rng


def _tv_prox_dual(g, mu, B, iters=40000):
    """min_u 1/2||u - g||^2 + mu sum_i |(B u)_i| via projected gradient on the dual."""
    n = g.size
    p = np.zeros(2 * n)
    tau = 0.12
    Bt = B.T.tocsr()
    for _ in range(iters):
        q = p + tau * (B @ (g - Bt @ p))
        nrm = np.sqrt(q[:n] ** 2 + q[n:] ** 2)
        f = np.maximum(1.0, nrm / mu)
        p = np.concatenate([q[:n] / f, q[n:] / f])
    return g - Bt @ p


def _soft(v, t):
    return np.sign(v) * np.maximum(np.abs(v) - t, 0.0)


def _l1(x):
    return float(np.abs(x).sum())


def test_sparsa_hand_trace_two_variables():
    """min 1/2||A x - y||^2 + ||x||_1 with A = diag(2, 1), y = (4, 1), x0 = 0, worked by hand:
    step 1 (alpha = 1): grad = A'(A x0 - y) = (-8, -1), soft((8, 1), 1) = (7, 0);
    BB: dx = (7, 0), A dx = (14, 0), alpha = 196 / 49 = 4;
    step 2: grad = A'((14, 0) - (4, 1)) = (20, -1), x - grad/4 = (2, 0.25),
    soft(., 1/4) = (1.75, 0) -- the minimiser (2(2x - 4) + 1 = 0 gives x = 1.75)."""
    A = np.diag([2.0, 1.0])
    y = np.array([4.0, 1.0])
    run = lambda maxiter: O.sparsa(y, lambda x: A @ x, lambda r: A.T @ r, 1.0, _soft, _l1, np.zeros(2),
                                   maxiter=maxiter, miniter=0, tolA=0.0)
    np.testing.assert_allclose(run(1), [7.0, 0.0], atol=1e-14)
    np.testing.assert_allclose(run(2), [1.75, 0.0], atol=1e-14)


def test_sparsa_identity_operator_one_step_is_the_prox():
    """With A = I the first step (alpha = 1) is psi(x0 - (x0 - y), tau) = psi(y, tau) for any x0."""
    N = 6
    B = O.diff_oper(N)
    y = rng(11).standard_normal(N * N)
    psi = lambda v, t: _tv_prox_dual(v, t, B, iters=3000)
    x1 = O.sparsa(y, lambda x: x, lambda r: r, 0.3, psi, lambda x: 0.0, rng(12).standard_normal(N * N),
                  maxiter=1, miniter=0, tolA=0.0)
    np.testing.assert_allclose(x1, psi(y, 0.3), atol=1e-12)


def test_sparsa_iteration_count_rule():
    """Steps taken: max(MiniterA, MaxiterA) when the objective criterion never fires; one
    step when it fires at once and MiniterA <= 1."""
    A = np.array([[1.0, 0.3], [0.2, 0.9]])
    y = np.array([1.0, -0.5])
    calls = []

    def psi(v, t):
        calls.append(1)
        return _soft(v, t)

    def steps(maxiter, miniter, tolA):
        calls.clear()
        O.sparsa(y, lambda x: A @ x, lambda r: A.T @ r, 0.01, psi, _l1, np.zeros(2), maxiter=maxiter,
                 miniter=miniter, tolA=tolA)
        return len(calls)

    assert steps(2, 5, 0.0) == 5          # the reference's call: MaxiterA 2, default MiniterA 5
    assert steps(2, 0, 0.0) == 2
    assert steps(7, 3, 0.0) == 7
    assert steps(10, 0, 1e9) == 1         # criterion 1 met after the first step
    assert steps(10, 4, 1e9) == 4         # ... but never before MiniterA steps


def test_sparsa_converges_to_the_lasso_minimiser():
    """Many SpaRSA steps on a well-conditioned L1 problem reach the minimiser that L-BFGS-B
    finds on the split form x = p - q, p, q >= 0 (an independent method)."""
    from scipy.optimize import minimize
    g = rng(13)
    A = np.eye(8)[:, :8] * 1.5 + 0.2 * g.standard_normal((12, 8))[:8]
    A = np.vstack([A, 0.3 * g.standard_normal((4, 8))])
    y = g.standard_normal(12)
    tau = 0.4
    x = O.sparsa(y, lambda v: A @ v, lambda r: A.T @ r, tau, _soft, _l1, np.zeros(8), maxiter=3000, miniter=0,
                 tolA=0.0)

    def fg(z):
        p, q = z[:8], z[8:]
        r = A @ (p - q) - y
        gr = A.T @ r
        return 0.5 * r @ r + tau * (p.sum() + q.sum()), np.concatenate([gr + tau, -gr + tau])
    res = minimize(fg, np.zeros(16), jac=True, method="L-BFGS-B", bounds=[(0, None)] * 16,
                   options=dict(ftol=1e-15, gtol=1e-12, maxiter=10000))
    xs = res.x[:8] - res.x[8:]
    assert np.any(np.abs(xs) < 1e-8) and np.any(np.abs(xs) > 1e-2)    # a sparse, non-trivial minimiser
    np.testing.assert_allclose(x, xs, atol=1e-6)


def test_sparsa_converges_to_the_tv_deblurring_minimiser():
    """SpaRSA with the isotropic-TV prox (psi) on a small deblurring problem reaches the
    minimiser of 1/2||A x - y||^2 + mu TV_B(x) computed by Chambolle-Pock (PDHG)."""
    N = 6
    n = N * N
    B = O.diff_oper(N)
    Bd = B.toarray()
    g = rng(14)
    A = np.eye(n) + 0.15 * g.standard_normal((n, n)) / np.sqrt(n)
    y = A @ np.kron(np.ones(N), (np.arange(N) > 2).astype(float)) + 0.05 * g.standard_normal(n)
    mu = 0.2
    tvb = lambda x: float(np.sum(np.sqrt((Bd @ x)[:n] ** 2 + (Bd @ x)[n:] ** 2)))
    psi = lambda v, t: _tv_prox_dual(v, t, B, iters=4000)
    x = O.sparsa(y, lambda v: A @ v, lambda r: A.T @ r, mu, psi, tvb, np.zeros(n), maxiter=150, miniter=0,
                 tolA=0.0)
    # Chambolle-Pock: K = B, F = mu * group-L2, G = 1/2||A x - y||^2 (prox by a dense solve)
    s_, t_ = 0.3, 0.3
    M = np.linalg.inv(np.eye(n) + t_ * A.T @ A)
    xc = np.zeros(n); xb = xc.copy(); p = np.zeros(2 * n)
    for _ in range(30000):
        q = p + s_ * (Bd @ xb)
        nrm = np.maximum(1.0, np.sqrt(q[:n] ** 2 + q[n:] ** 2) / mu)
        p = np.concatenate([q[:n] / nrm, q[n:] / nrm])
        xn = M @ (xc - t_ * (Bd.T @ p) + t_ * (A.T @ y))
        xb = 2 * xn - xc
        xc = xn
    assert rel_l2(x, xc) < 1e-4


def test_neumann_solver_inverts_I_plus_BtB():
    for N in (5, 12, 17):
        B = O.diff_oper(N)
        M = (sp.identity(N * N) + B.T @ B).toarray()
        r = rng(N).standard_normal(N * N)
        np.testing.assert_allclose(M @ O.neumann_solver(N)(r), r, atol=1e-12)


@pytest.mark.parametrize("exact", [False, True])
def test_sb_itv2_is_the_tv_prox(exact):
    N = 10
    B = O.diff_oper(N)
    g = rng(3).standard_normal(N * N) + np.kron(np.ones(N), np.linspace(0, 2, N))
    u, it = O.sb_itv2(g, 0.4, B, exact=exact)
    ref = _tv_prox_dual(g, 0.4, B)
    assert it > 1
    assert rel_l2(u, ref) < 5e-3


def test_itv_definition():
    X = rng(4).standard_normal((6, 7))             # (Nx, Nz)
    M = X.T                                        # MATLAB [Nz x Nx]
    gx = np.zeros_like(M); gy = np.zeros_like(M)
    gx[:, 1:-1] = (M[:, 2:] - M[:, :-2]) / 2; gx[:, 0] = M[:, 1] - M[:, 0]; gx[:, -1] = M[:, -1] - M[:, -2]
    gy[1:-1] = (M[2:] - M[:-2]) / 2; gy[0] = M[1] - M[0]; gy[-1] = M[-1] - M[-2]
    assert abs(O.itv(X) - np.sum(np.sqrt(gx ** 2 + gy ** 2))) < 1e-12


@pytest.mark.gpu
def test_tv_solver_matches_oracle(require_gpu):
    from test_gpu_operators import _geometry_model, _oracle_of
    from mbrec.phantoms import disks, add_noise
    from mbrec.solvers import rec_nn_with_TV_reg
    print("tv parity", flush=True)
    m = _geometry_model(24, 24, 24, 400, 1.25e-4, sir=False, sigma=3.0)
    orc = _oracle_of(m)
    ph = np.stack([disks(m.Discretization, n=4, seed=s) for s in (8, 9)])
    y = add_noise(orc.forward(ph), 0.01, seed=10)
    tau = 1e-3
    rec, e_img, e_reg, *_ = rec_nn_with_TV_reg(m, y, 6, tau, coupled=True)
    ro, lc = O.rec_nn_with_TV_reg(orc, y, 6, tau, exact_prox=True)
    assert rel_l2(rec, ro) < 1e-3
    # the reference's cgs-based prox differs from the exact solve by its 1e-5 tolerance
    ro_cgs, _ = O.rec_nn_with_TV_reg(orc, y, 6, tau, exact_prox=False)
    assert rel_l2(rec, ro_cgs) < 1e-2
    # MaxiterA and MiniterA passed explicitly: two SpaRSA steps per outer iteration
    rec2, *_ = rec_nn_with_TV_reg(m, y, 6, tau, coupled=False, inner_iterations=2, inner_min_iterations=0)
    for b in range(2):
        ro2, _ = O.rec_nn_with_TV_reg(orc, y[b:b + 1], 6, tau, exact_prox=True, inner=2, inner_min=0)
        assert rel_l2(rec2[b], ro2[0]) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("N,mu", [(7, 0.3), (12, 0.4), (25, 0.2), (64, 0.5), (400, 0.05), (600, 0.05)])
def test_sb_itv2_device_matches_oracle(require_gpu, N, mu):
    """The device split-Bregman TV prox (FFT-based DCT Neumann solve: radix 4/2/3/5 and a
    prime side) against the oracle's SB_ITV2 with the exact dense-DCT solve, same FP32
    input, two images with different mu in one call."""
    from mbrec.solvers import SB_ITV2
    print(f"sb_itv2 N={N}", flush=True)
    g = rng(N).standard_normal((2, N, N)).astype(np.float32)
    yy, xx = np.mgrid[0:N, 0:N]
    g += 3.0 * ((xx - N / 2) ** 2 + (yy - N / 3) ** 2 < (N / 4) ** 2)
    g = g.astype(np.float32)
    mus = np.array([mu, 2 * mu])
    u, it = SB_ITV2(g, mus)
    B = O.diff_oper(N)
    its = []
    for b in range(2):
        ref, k = O.sb_itv2(g[b].astype(np.float64).ravel(), mus[b], B, exact=True)
        its.append(k)
        assert rel_l2(u[b].ravel(), ref) < 1e-6, b
    assert it == max(its)
