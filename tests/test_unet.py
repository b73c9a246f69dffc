"""Warm-start U-Net: the CPU oracle's pins and the device kernels' parity.

CPU: the oracle's convolution equals scipy.signal.correlate2d, pooling and
upsampling satisfy their adjoint identities, bfloat16 rounding reproduces
hand-derived values, and the oracle's backward pass equals central finite
differences of its forward pass (FP64, no rounding).

GPU: one matrix-core convolution against torch's FP32 conv2d; the full forward
pass and every layer's gradient against the oracle with bfloat16 emulated at
the device's storage points (tolerances in the tests); a few Adam steps lower
the loss.
"""
import numpy as np
import pytest

import unet_oracle as U
from helpers import rel_l2, rng


# --- CPU pins ------------------------------------------------------------------------
def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even mantissa (1); 1 + 3*2^-8 -> 1 + 2^-6;
    # 1/3 = 0x3EAAAAAB in float32 -> 0x3EAB = 0.333984375; 65504 -> 65536 (0x4780)
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 1.0 / 3.0, 65504.0])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.333984375, 65536.0])
    np.testing.assert_array_equal(U.bf16(x), want)


def test_conv3_matches_scipy_correlate():
    from scipy.signal import correlate2d
    g = rng(1)
    x = g.standard_normal((2, 7, 9, 3))
    w = g.standard_normal((4, 3, 3, 3))
    y = U.conv3(x, w)
    for n in range(2):
        for co in range(4):
            ref = sum(correlate2d(x[n, :, :, ci], w[co, :, :, ci], mode="same") for ci in range(3))
            np.testing.assert_allclose(y[n, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_conv3_input_gradient_is_adjoint():
    g = rng(2)
    x = g.standard_normal((1, 6, 8, 3))
    w = g.standard_normal((5, 3, 3, 3))
    yb = g.standard_normal((1, 6, 8, 5))
    lhs = np.sum(U.conv3(x, w) * yb)
    rhs = np.sum(x * U.conv3(yb, U.flip_t(w)))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    # filter gradient: d<conv(x, w), yb>/dw = wgrad(yb, x)
    dw = U.conv3_wgrad(yb, x)
    e = g.standard_normal(w.shape)
    assert abs(np.sum(U.conv3(x, e) * yb) - np.sum(dw * e)) < 1e-10 * np.abs(dw).sum()


def test_pool_and_upsample_adjoints():
    g = rng(3)
    x = g.standard_normal((2, 8, 8, 3))
    p = g.standard_normal((2, 4, 4, 3))
    assert abs(np.sum(U.pool(x) * p) - np.sum(x * U.pool_bwd(p))) < 1e-12
    assert abs(np.sum(U.up(p) * x) - np.sum(p * U.up_bwd(x))) < 1e-12
    np.testing.assert_allclose(U.pool(U.up(p)), p)


def test_layout_counts():
    lay, n = U.layout(32)
    assert len(lay) == 15
    # hand count for c = 32: filters 9 * 216096 + 32 (head), biases 1409
    assert sum(co * k * k * ci + co for ci, co, k, _, _ in lay) == 1944896 + 1409
    assert all(wo % 8 == 0 for _, _, _, wo, _ in lay) and n >= 1944896 + 1409


def test_oracle_gradient_matches_finite_differences():
    c, N, H, W = 2, 2, 8, 8
    th = U.he_init(c, 5)
    g = rng(6)
    x = g.uniform(0.0, 1.0, (N, H, W))
    t = g.uniform(0.0, 1.0, (N, H, W))
    loss, grad, _ = U.loss_and_grad(th, x, t, c)
    eps = 1e-6
    for k in range(4):
        v = g.standard_normal(th.size)
        lp = U.loss_and_grad(th + eps * v, x, t, c)[0]
        lm = U.loss_and_grad(th - eps * v, x, t, c)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad @ v) < 1e-6 * max(abs(fd), 1e-3), (k, fd, grad @ v)
    # per-layer: a direction confined to the first filter and to the head
    lay, _ = U.layout(c)
    for l in (0, 7, 14):
        ci, co, kk, wo, bo = lay[l]
        v = np.zeros(th.size)
        v[wo:bo + co] = g.standard_normal(bo + co - wo)
        fd = (U.loss_and_grad(th + eps * v, x, t, c)[0] - U.loss_and_grad(th - eps * v, x, t, c)[0]) / (2 * eps)
        assert abs(fd - grad @ v) < 1e-6 * max(abs(fd), 1e-3), (l, fd, grad @ v)


def test_quantised_oracle_is_close_to_fp64():
    c, N, H, W = 32, 1, 16, 16
    th = U.he_init(c, 7)
    x = rng(8).uniform(0, 1, (N, H, W))
    o64, _ = U.forward(th, x, c, quant=False)
    o16, _ = U.forward(th, x, c, quant=True)
    assert 1e-4 < rel_l2(o16, o64) < 5e-2   # bfloat16 storage changes the result, but only slightly


# --- GPU parity ----------------------------------------------------------------------
def _bf16_tensor(a):
    import torch
    return torch.as_tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)


@pytest.mark.gpu
@pytest.mark.parametrize("N,H,W,Cin,Cout,relu", [(2, 9, 13, 32, 64, 1), (1, 16, 16, 96, 32, 0),
                                                 (3, 5, 7, 64, 128, 1)])
def test_conv3x3_bf16_matches_torch_fp32(require_gpu, N, H, W, Cin, Cout, relu):
    import torch
    import torch.nn.functional as F
    from mbrec import _lib
    print(f"conv3x3_bf16 N={N} H={H} W={W} Cin={Cin} Cout={Cout} relu={relu}", flush=True)
    g = rng(11)
    x = _bf16_tensor(g.standard_normal((N, H, W, Cin)))
    w = _bf16_tensor(g.standard_normal((Cout, 3, 3, Cin)) / np.sqrt(9 * Cin))
    b = torch.as_tensor(g.standard_normal(Cout).astype(np.float32)) if relu else None
    ref = F.conv2d(x.float().permute(0, 3, 1, 2), w.float().permute(0, 3, 1, 2), b, padding=1).permute(0, 2, 3, 1)
    if relu:
        ref = torch.relu(ref)
    xd, wd = x.cuda().contiguous(), w.cuda().contiguous()
    bd = b.cuda() if relu else None
    yd = torch.empty((N, H, W, Cout), dtype=torch.bfloat16, device="cuda")
    _lib.check(_lib.lib().mbrec_conv3x3_bf16(xd.data_ptr(), wd.data_ptr(), bd.data_ptr() if relu else None,
                                             yd.data_ptr(), N, H, W, Cin, Cout, relu, _lib.stream_handle()))
    got = yd.float().cpu()
    # the output is rounded to bfloat16: elementwise error <= 2^-8 relative (+ FP32 sum order)
    err = (got - ref).abs()
    assert torch.all(err <= 2.0 ** -8 * ref.abs() + 1e-5), float(err.max())
    assert rel_l2(got.numpy(), ref.numpy()) < 4e-3


def _net_case():
    # 32 x 32 (4 x 4 at the bottleneck): at 16 x 16 the bottleneck sums run over so
    # few pixels that recomputing the oracle with FP32 instead of FP64 sums (which
    # flips a few bfloat16 roundings) already moves deep-layer gradients by ~3 %;
    # at 32 x 32 that noise floor is <= 0.25 % per layer.
    c, N, H, W = 32, 2, 32, 32
    g = rng(21)
    th = U.he_init(c, 22).astype(np.float32).astype(np.float64)
    x = g.uniform(0.0, 1.0, (N, H, W)).astype(np.float32)
    t = (0.5 * x + 0.1 * g.uniform(0, 1, (N, H, W))).astype(np.float32)
    return c, th, x, t


@pytest.mark.gpu
def test_unet_layout_matches_oracle(require_gpu):
    from mbrec.warmstart import UNet
    net = UNet(32, seed=1)
    lay, n = U.layout(32)
    assert net.num_params == n
    assert [tuple(r[1:]) for r in net.layout()] == [tuple(r) for r in lay]


@pytest.mark.gpu
def test_unet_forward_matches_oracle(require_gpu):
    from mbrec.warmstart import UNet
    print("unet forward parity", flush=True)
    c, th, x, t = _net_case()
    net = UNet(c)
    net.set_params(th)
    got = net.forward(x).cpu().numpy()
    want, _ = U.forward(th, x.astype(np.float64), c, quant=True)
    assert np.abs(want).max() > 0
    assert rel_l2(got, want) < 5e-3
    assert np.allclose(net.get_params(), th.astype(np.float32))


@pytest.mark.gpu
def test_unet_gradients_match_oracle(require_gpu):
    from mbrec.warmstart import UNet
    print("unet gradient parity", flush=True)
    c, th, x, t = _net_case()
    net = UNet(c)
    net.set_params(th)
    loss, grad = net.gradients(x, t)
    lw, gw, _ = U.loss_and_grad(th, x.astype(np.float64), t.astype(np.float64), c, quant=True)
    assert abs(loss - lw) < 1e-3 * lw
    lay, _ = U.layout(c)
    for (ci, co, k, wo, bo), name in zip(lay, U.LAYER_NAMES):
        gw_l, gg_l = gw[wo:bo + co], grad[wo:bo + co]
        assert np.abs(gw_l).max() > 0, name
        assert rel_l2(gg_l, gw_l) < 1.5e-2, (name, rel_l2(gg_l, gw_l))
    assert rel_l2(grad, gw) < 5e-3


@pytest.mark.gpu
def test_unet_relative_loss_weights_items(require_gpu):
    """Relative loss of a two-item batch = (w0 L0 + w1 L1) / 2 with w_b = mean(P) / P_b,
    P_b = mean((t_b / s_b)^2), s_b = rms(x_b); each L_b and its gradient from the plain
    loss of that item alone."""
    from mbrec.warmstart import UNet
    print("unet relative loss", flush=True)
    c, th, x, t = _net_case()
    t = t.copy()
    t[1] *= 3.0                     # item 1: nine times the normalised target power
    net = UNet(c)
    net.set_params(th)
    per = [net.gradients(x[b:b + 1], t[b:b + 1]) for b in range(2)]
    plain_loss, plain_grad = net.gradients(x, t)
    assert abs(plain_loss - 0.5 * (per[0][0] + per[1][0])) < 1e-4 * plain_loss
    xd, td = x.astype(np.float64), t.astype(np.float64)
    s = np.sqrt((xd ** 2).mean(axis=(1, 2)))
    P = ((td / s[:, None, None]) ** 2).mean(axis=(1, 2))
    w = P.mean() / P
    assert w[0] > 4 * w[1]
    net.set_relative_loss(True)
    loss, grad = net.gradients(x, t)
    want_loss = 0.5 * (w[0] * per[0][0] + w[1] * per[1][0])
    want_grad = 0.5 * (w[0] * per[0][1].astype(np.float64) + w[1] * per[1][1].astype(np.float64))
    assert abs(loss - want_loss) < 1e-3 * want_loss
    assert rel_l2(grad, want_grad) < 1e-2, rel_l2(grad, want_grad)
    assert rel_l2(plain_grad, want_grad) > 0.1
    net.set_relative_loss(False)
    loss2, grad2 = net.gradients(x, t)
    # weight gradients are summed with FP32 atomics: equal up to the order of the additions
    assert abs(loss2 - plain_loss) < 1e-6 * plain_loss
    assert rel_l2(grad2, plain_grad) < 1e-5


@pytest.mark.gpu
def test_unet_adam_steps_reduce_loss(require_gpu):
    from mbrec.warmstart import UNet
    print("unet training steps", flush=True)
    c, th, x, t = _net_case()
    net = UNet(c, seed=3)
    losses = [net.train_step(x, t, lr=1e-3) for _ in range(30)]
    assert losses[-1] < 0.5 * losses[0], losses[::5]
