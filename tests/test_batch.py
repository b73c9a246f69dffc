"""Batched NNLS-L2 over mixed grids and cropped sinograms in one graph launch.

Each sinogram of a batch is a window [first, first + count) of its model's
sample axis.  Its oracle is the reference model of the cropped acquisition:
travel times shifted by -first and count samples (MSOTmodel.m:350-382 with
numCroppedSamplesAtSinogramStart = first), solved by rec_nn_with_L2_reg with
its own lambdaL2 = 1e-2 ||A'y||.  Graph replays must equal the first replay and
the direct (uncaptured) launch bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_l2, rng

TOL_REC = 1e-3


def _geometry_model(Nz, Nx, D, nT, h, sigma):
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe
    disc = Discretization([-(Nx - 1) / 2 * h, (Nx - 1) / 2 * h, -(Nz - 1) / 2 * h, (Nz - 1) / 2 * h], [Nx, Nz])
    probe = Probe("arc256_r40_174", True, False, num_detectors=D)
    probe.DAC.numRecordedSamplesPerTransducer = nT
    probe.DAC.delayBeforeRecording = int(round(probe.detector.radius / 1500.0 * probe.DAC.frequency)) - nT // 2
    return MSOTmodel(probe, disc, 1500.0, False, True, 0, 1e5, 12e6, sigma)


def _cropped_oracle(model, first, count):
    d = model.Discretization
    sig_px = d.pixelSigma   # the model's own sigma = round(min(dx, dz) / 2 um)
    kz, kx = O.full_kernels(d.Nz, d.Nx, d.dz, d.dx, sig_px)
    D = model.Probe.detector.numOfTransducers
    tt = model.Data.absoluteTravelTimeIndices.reshape(D, -1) - first
    return O.Model(d.Nz, d.Nx, tt, count, model.Data.circularInterpolationRadiusInSamples,
                   sigma=model.normalizationFactor, kz=kz, kx=kx, ir=model.Data.IR)


@pytest.mark.gpu
def test_windowed_operators_match_cropped_oracle(require_gpu):
    """Forward / transpose on windows equal the oracle of the cropped model, including
    windows that cut the IR tails."""
    import torch
    from mbrec import _lib
    print("windowed operators", flush=True)
    m = _geometry_model(32, 28, 16, 360, 1.25e-4, 3.0)
    wins = np.array([[40, 320], [0, 250], [100, 150], [0, 360]], dtype=np.int32)
    B = len(wins)
    D, nT = m.shape_sinogram
    x = np.maximum(rng(31).standard_normal((B, 28, 32)), 0).astype(np.float32)
    s = rng(32).standard_normal((B, D, nT)).astype(np.float32)
    xd, sd = torch.as_tensor(x).cuda(), torch.as_tensor(s).cuda()
    yd = torch.empty((B, D, nT), device="cuda")
    ad = torch.empty((B, 28, 32), device="cuda")
    L = _lib.lib()
    _lib.check(L.mbrec_model_forward_window_dev(m.plan.h, xd.data_ptr(), yd.data_ptr(), B, wins.ctypes.data,
                                                _lib.stream_handle()))
    _lib.check(L.mbrec_model_transpose_window_dev(m.plan.h, sd.data_ptr(), ad.data_ptr(), B, wins.ctypes.data,
                                                  _lib.stream_handle()))
    y, a = yd.cpu().numpy(), ad.cpu().numpy()
    for b, (f, c) in enumerate(wins):
        orc = _cropped_oracle(m, f, c)
        yo = orc.forward(x[b].astype(np.float64))[0]
        assert rel_l2(y[b, :, f:f + c], yo) < 1e-5, (b, rel_l2(y[b, :, f:f + c], yo))
        assert not np.any(y[b, :, :f]) and not np.any(y[b, :, f + c:])
        ao = orc.transpose(s[b, :, f:f + c].astype(np.float64))[0]
        assert rel_l2(a[b], ao) < 1e-5, (b, rel_l2(a[b], ao))


@pytest.mark.gpu
def test_batch_graph_matches_oracle_on_cropped_sinograms(require_gpu):
    import torch
    from mbrec.batch import BatchSolver
    from mbrec.phantoms import disks
    print("batch graph parity", flush=True)
    m1 = _geometry_model(24, 24, 16, 320, 1.25e-4, 3.0)
    m2 = _geometry_model(32, 28, 16, 360, 1.25e-4, 3.0)
    wins = [[(0, 320), (20, 280), (5, 300)], [(40, 320), (0, 250)]]
    iters = 12
    ys, lams, orcs = [], [], []
    for gi, (m, w) in enumerate(zip((m1, m2), wins)):
        for k, (f, c) in enumerate(w):
            orc = _cropped_oracle(m, f, c)
            ph = disks(m.Discretization, n=5, seed=100 + 10 * gi + k)
            y = orc.forward(ph)
            y = y + 0.01 * np.abs(y).max() * rng(200 + 10 * gi + k).standard_normal(y.shape)
            ys.append(y[0])
            lams.append(1e-2 * np.linalg.norm(orc.transpose(y)))
            orcs.append(orc)
    bs = BatchSolver([(m1, wins[0]), (m2, wins[1])], iters, lams)
    i = 0
    for g, w in enumerate(wins):
        for k, (f, c) in enumerate(w):
            bs.data[g][k].fill_(12345.0)                      # outside the window: must be ignored
            bs.data[g][k, :, f:f + c] = torch.as_tensor(ys[i], dtype=torch.float32)
            i += 1
    r_graph = [t.clone() for t in bs.run(graph=True)]
    torch.cuda.synchronize()
    r_replay = [t.clone() for t in bs.run(graph=True)]
    torch.cuda.synchronize()
    r_eager = [t.clone() for t in bs.run(graph=False)]
    torch.cuda.synchronize()
    for a, b, c in zip(r_graph, r_replay, r_eager):
        assert torch.equal(a, b) and torch.equal(a, c)
    i = 0
    for g, w in enumerate(wins):
        for k in range(len(w)):
            want, _, _ = orcs[i].rec_nn_with_L2_reg(ys[i][None], iters, lams[i])
            got = r_graph[g][k].cpu().numpy()
            assert np.abs(want).max() > 0
            assert rel_l2(got, want[0]) < TOL_REC, (g, k, rel_l2(got, want[0]))
            i += 1
    bs.close()


@pytest.mark.gpu
def test_batch_pins_plans_and_checks_windows(require_gpu):
    from mbrec import MbrecError
    from mbrec.batch import BatchSolver
    m = _geometry_model(16, 16, 8, 200, 1.25e-4, 3.0)
    with pytest.raises(MbrecError) as ei:
        BatchSolver([(m, [(150, 100)])], 2, [0.1])
    assert ei.value.identifier == "mbrec:window"
    bs = BatchSolver([(m, [(0, 200), (10, 150)])], 2, [0.1, 0.2])
    with pytest.raises(MbrecError) as ei:
        m.plan.set_sigma(2.0)
    assert ei.value.identifier == "mbrec:planPinned"
    bs.close()
    m.plan.set_sigma(3.0)
