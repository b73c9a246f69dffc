"""Parity at the five benchmark configurations' own sizes (GPU vs the FP64 oracle).

C1  200^2 / 125 um, no SIR, 256 x 2030: rec_nn_with_L2_reg, lambdaL2 = 1e-2 ||A'y||, 50
    iterations, within 1e-3 of the oracle's solve.
C2  400^2 / 100 um with SIR: forward and transpose within 1e-5 of the oracle, the device
    pair's adjoint identity within 1e-6; dmas_cf and backproject_waveeq on one seeded
    256 x 2030 float32 sinogram against the oracle (FP64 both sides).
C3  the 224-sinogram batch over 200^2 / 300^2 / 400^2 grids with crop lengths U{1500..2030}
    in one graph: one sinogram per grid within 1e-3 of the cropped-model oracle's solve,
    graph replay = direct launch bit for bit for all 224.
C4  400^2, batch 64, learned warm start: fewer outer iterations to 1e-3 than the cold start;
    the warm-started loop run to the iteration count
    at which the whole batch is within 1e-3 of the device's converged solution; two
    sinograms checked against the oracle's FP64 converged solution (1e-3, plus the FP32
    operator's own 1e-5 scale).
C5  600^2 / 60 um, SIR, 512 detectors: the device pair's adjoint identity at full size and
    one outer iteration of rec_nn_with_TV_reg on one sinogram against the oracle.
Sizes and seeds follow the benchmark; each test prints what it runs before it runs it.
"""
import time

import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, rng
from test_gpu_operators import _oracle_of

pytestmark = pytest.mark.gpu

C_, FS, COV, RAD = 1500.0, 4e7, 174.0, 0.040


@pytest.fixture(scope="module", autouse=True)
def _gpu(require_gpu):
    import mbrec
    mbrec.lib()
    yield


def _log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _model(N, h, sir, D=256, sigma=None):
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe
    half = (N - 1) / 2 * h
    probe = Probe("arc256_r40_174", True, False, num_detectors=D)
    return MSOTmodel(probe, Discretization([-half, half], N), C_, sir, True, 0, 1e5, 12e6, sigma)


def _noisy(model, ph, seed):
    import torch
    y = model.applyForward(torch.as_tensor(ph, dtype=torch.float32).cuda())
    g = torch.Generator(device="cuda").manual_seed(seed)
    return (y + 0.01 * y.abs().amax(dim=(1, 2), keepdim=True) * torch.randn(y.shape, generator=g,
                                                                            device="cuda")).contiguous()


def test_c1_plain_l2_200():
    from mbrec.phantoms import disks
    from mbrec.solvers import rec_nn_with_L2_reg
    _log("C1: 200^2 / 125 um, 256 x 2030, rec_nn_with_L2_reg 50 iterations")
    m = _model(200, 125e-6, False)
    orc = _oracle_of(m)
    ph = disks(m.Discretization, n=20, seed=1)[None]
    y = orc.forward(ph)
    y = y + 0.01 * np.abs(y).max() * rng(2).standard_normal(y.shape)
    lam = 1e-2 * np.linalg.norm(orc.transpose(y))
    rec, *_ = rec_nn_with_L2_reg(m, y, 50, lam)
    _log("C1: oracle solve")
    ro, _, _ = orc.rec_nn_with_L2_reg(y, 50, lam)
    e = rel_l2(rec, ro)
    _log(f"C1: rel.err vs oracle {e:.2e}")
    assert e < 1e-3


def test_c2_sir_operators_and_backprojection_400():
    import torch
    from mbrec.backprojection import backproject_waveeq, bp_kernel, dmas_cf
    _log("C2: 400^2 / 100 um with SIR: model set-up")
    m = _model(400, 1e-4, True)
    orc = _oracle_of(m)
    x = np.maximum(rng(21).standard_normal((1, 400, 400)), 0)
    s = rng(22).standard_normal((1, 256, 2030))
    _log("C2: device operators")
    y = m.applyForward(x)
    xt = m.applyTranspose(s)
    xd = torch.as_tensor(x, dtype=torch.float32).cuda()
    sd = torch.as_tensor(s, dtype=torch.float32).cuda()
    ip1 = float((m.applyForward(xd).double() * sd.double()).sum())
    ip2 = float((xd.double() * m.applyTranspose(sd).double()).sum())
    _log("C2: oracle operators")
    ey, ex = rel_l2(y, orc.forward(x)), rel_l2(xt, orc.transpose(s))
    ea = abs(ip1 - ip2) / max(abs(ip1), abs(ip2))
    _log(f"C2: forward {ey:.2e}, transpose {ex:.2e}, adjoint {ea:.2e}")
    assert ey < 1e-5 and ex < 1e-5 and ea < 1e-6
    # secondary paths on one seeded 256 x 2030 float32 sinogram (delay 0: cropped = 0)
    X, Z = m.Discretization.region["xPositionsFov"].T, m.Discretization.region["zPositionsFov"].T
    sig = rng(23).standard_normal((1, 256, 2030)).astype(np.float32)
    _log("C2: dmas_cf")
    got = dmas_cf(sig, X, Z, C_, FS, 0, COV, RAD)
    want = O.dmas_cf(sig.astype(np.float64), X, Z, C_, FS, 0, COV, RAD)
    for k in range(4):
        assert rel_l2(got[k], want[k]) < 1e-10, k
    _log("C2: backproject_waveeq (K from the device bp_kernel, Nt = 2030)")
    t = np.arange(1, 2031) / FS
    K = bp_kernel(t)
    psf = np.outer(np.hanning(5), np.hanning(3))
    got = backproject_waveeq(sig.astype(np.float64), X, Z, C_, FS, 0, COV, RAD, psf, K)
    want = O.backproject_waveeq(sig.astype(np.float64), X, Z, C_, FS, 0, COV, RAD, psf, K)
    e = rel_l2(got, want)
    _log(f"C2: backproject_waveeq rel.err {e:.2e}")
    assert e < 1e-8


def test_c3_mixed_batch_224():
    import torch
    from mbrec.batch import BatchSolver
    from mbrec.warmstart import phantom_batch
    g = np.random.default_rng(2024)
    n_wl, n_fr = 28, 8
    n = n_wl * n_fr
    grids = g.choice([200, 300, 400], size=n)
    counts = g.integers(1500, 2031, size=n)
    firsts = np.array([g.integers(0, 2030 - c + 1) for c in counts])
    spectrum = 0.6 + 0.4 * np.sin(np.linspace(0, np.pi, n_wl)) ** 2
    groups, ys, picks = [], [], []
    _log(f"C3: {n} sinograms, grids " + ", ".join(f"{N}^2 x {int(np.sum(grids == N))}" for N in (200, 300, 400)))
    for N in (200, 300, 400):
        idx = np.nonzero(grids == N)[0]
        m = _model(N, 1e-4, False)
        ph = phantom_batch(m.Discretization, len(idx), 7000 + N) * spectrum[idx // n_fr][:, None, None]
        y = _noisy(m, ph, N)
        lam = (1e-2 * torch.linalg.vector_norm(m.applyTranspose(y).reshape(len(idx), -1).double(), dim=1)).cpu()
        groups.append((m, [(int(firsts[i]), int(counts[i])) for i in idx]))
        ys.append((y, lam.numpy()))
    lam_all = np.concatenate([t[1] for t in ys])
    bs = BatchSolver(groups, 50, lam_all)
    for k, (y, _) in enumerate(ys):
        bs.data[k].copy_(y)
    _log("C3: graph launch")
    rec_g = [r.clone() for r in bs.run(graph=True)]
    rec_g2 = [r.clone() for r in bs.run(graph=True)]
    _log("C3: direct launch")
    rec_e = [r.clone() for r in bs.run(graph=False)]
    for a, b, c in zip(rec_g, rec_g2, rec_e):
        assert torch.equal(a, b) and torch.equal(a, c)
    for k, (m, wins) in enumerate(groups):
        f, c = wins[0]
        N = m.Discretization.Nx
        _log(f"C3: oracle solve of the first {N}^2 sinogram (window {f}..{f + c})")
        base = _oracle_of(m)
        orc = O.Model(base.Nz, base.Nx, base.tt - f, c, base.R, sigma=m.normalizationFactor, kz=base.kz, kx=base.kx,
                      ir=base.ir)
        yk = bs.data[k][0, :, f:f + c].double().cpu().numpy()
        ro = orc.rec_nn_with_L2_reg_from(yk[None], 50, ys[k][1][0])
        e = rel_l2(rec_g[k][0].cpu().numpy(), ro[0])
        _log(f"C3: {N}^2 rel.err vs the cropped-model oracle {e:.2e}")
        assert e < 1e-3
    bs.close()


def test_c4_learned_warm_start_400():
    import os
    import torch
    from mbrec.solvers import rec_nn_with_L2_reg_batch
    from mbrec.warmstart import UNet, iterations_to_tolerance, nnls_l2_minimiser, phantom_batch, warm_start_estimate
    weights = os.path.join(os.path.dirname(__file__), "..", "mb-rec-msot_amd", "mbrec", "weights", "unet_state_400.npz")
    if not os.path.exists(weights):
        pytest.skip("warm-start weights not committed yet")
    _log("C4: 400^2, batch 64, learned warm start")
    m = _model(400, 1e-4, False)
    B = 64
    y = _noisy(m, phantom_batch(m.Discretization, B, 1000), 7)
    lam = (1e-2 * torch.linalg.vector_norm(m.applyTranspose(y).reshape(B, -1).double(), dim=1)).cpu().numpy()
    ref, kkt = nnls_l2_minimiser(m, y, lam)
    net = UNet.load(weights)
    w0, v0, used = warm_start_estimate(net, m, y, lam)
    k, _ = iterations_to_tolerance(m, y, lam, ref, 1e-3, w0, v0)
    _log(f"C4: {k} outer iterations from the network's state ({int(used.sum())}/{B} items use it)")
    assert k is not None
    k_cold, _ = iterations_to_tolerance(m, y, lam, ref, 1e-3)
    _log(f"C4: cold start {k_cold} outer iterations: the network saves {k_cold - k} ({k_cold / k:.2f}x)")
    assert k_cold is not None and k < k_cold
    rec = rec_nn_with_L2_reg_batch(m, y, k, lam, init=w0, init_dual=v0)
    orc = _oracle_of(m)
    pick = [0, 37]
    _log("C4: oracle FP64 converged solution of two sinograms")
    xo, kkt_o = O.nnls_l2_minimiser(orc, y[pick].double().cpu().numpy(), lam[pick], tol=1e-8)
    for j, b in enumerate(pick):
        e = rel_l2(rec[b].cpu().numpy(), xo[j])
        e_ref = rel_l2(ref[b].cpu().numpy(), xo[j])
        _log(f"C4: sinogram {b}: learned solve vs FP64 solution {e:.2e} (device solution {e_ref:.2e})")
        assert e_ref < 1e-4 and e < 1e-3 + e_ref


def test_c5_tv_600_sir_512():
    import torch
    from mbrec.phantoms import vessel_tree
    from mbrec.solvers import rec_nn_with_TV_reg
    _log("C5: 600^2 / 60 um, SIR, 512 detectors: model set-up")
    m = _model(600, 60e-6, True, D=512)
    D, nT = m.shape_sinogram
    x = torch.as_tensor(np.maximum(rng(51).standard_normal((1, 600, 600)), 0), dtype=torch.float32).cuda()
    s = torch.as_tensor(rng(52).standard_normal((1, D, nT)), dtype=torch.float32).cuda()
    ip1 = float((m.applyForward(x).double() * s.double()).sum())
    ip2 = float((x.double() * m.applyTranspose(s).double()).sum())
    ea = abs(ip1 - ip2) / max(abs(ip1), abs(ip2))
    _log(f"C5: device adjoint identity {ea:.2e}")
    assert ea < 1e-6
    ph = vessel_tree(m.Discretization, seed=100)[None]
    y = _noisy(m, ph, 53)
    tau = 1e-3
    rec, *_ = rec_nn_with_TV_reg(m, y, 1, tau, coupled=True, inner_iterations=2, inner_min_iterations=0)
    _log("C5: oracle TV outer iteration (SpaRSA 2 steps, exact Neumann prox)")
    orc = _oracle_of(m)
    ro, _ = O.rec_nn_with_TV_reg(orc, y.double().cpu().numpy(), 1, tau, exact_prox=True, inner=2, inner_min=0)
    e = rel_l2(rec.cpu().numpy(), ro)
    _log(f"C5: rel.err vs oracle {e:.2e}")
    assert e < 1e-3
