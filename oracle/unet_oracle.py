"""CPU oracle of the warm-start U-Net (test infrastructure only: imported by
tests/ as the checker, never by the product path).

The network has no counterpart in the reference (its NNLS loops start from
zero, nnls_L2_less.m:28-29), so this file restates the architecture that the
HIP library implements and serves two purposes:

* ``quant=False``: a plain FP64 forward / backward pass.  Its backward pass is
  pinned by central finite differences of its own forward pass
  (tests/test_unet.py), its convolution by scipy.signal.correlate2d, and
  pooling / upsampling by their adjoint identities.
* ``quant=True``: the same pass with every tensor that the device stores in
  bfloat16 rounded to bfloat16 (round to nearest even) at the same point, so
  device forward outputs and gradients can be compared at ~1e-3 instead of the
  ~1e-2 that bfloat16 storage alone would allow.

Layout: images (N, H, W); activations (N, H, W, C) (channel-minor, as on the
device); filters [Cout][k][k][Cin]; all parameters in one vector whose layer
offsets are those of ``layout`` (each layer's filter, then its bias, padded to
a multiple of 8 values).
"""
from __future__ import annotations

import numpy as np

LAYER_NAMES = ["e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "ba", "bb",
               "d3a", "d3b", "d2a", "d2b", "d1a", "d1b", "head"]


def layer_shapes(c):
    """(cin, cout, k) of the 15 layers for base width c."""
    return [(1, c, 3), (c, c, 3), (c, 2 * c, 3), (2 * c, 2 * c, 3), (2 * c, 4 * c, 3), (4 * c, 4 * c, 3),
            (4 * c, 8 * c, 3), (8 * c, 8 * c, 3), (12 * c, 4 * c, 3), (4 * c, 4 * c, 3), (6 * c, 2 * c, 3),
            (2 * c, 2 * c, 3), (3 * c, c, 3), (c, c, 3), (c, 1, 1)]


def layout(c):
    """[(cin, cout, k, w_off, b_off)] and the parameter count."""
    out, off = [], 0
    for cin, cout, k in layer_shapes(c):
        w_off = off
        off += cout * k * k * cin
        b_off = off
        off += cout
        off = (off + 7) // 8 * 8
        out.append((cin, cout, k, w_off, b_off))
    return out, off


def unpack(theta, c):
    lay, n = layout(c)
    assert theta.size == n
    Ws, bs = [], []
    for cin, cout, k, wo, bo in lay:
        Ws.append(theta[wo:wo + cout * k * k * cin].reshape(cout, k, k, cin).astype(np.float64))
        bs.append(theta[bo:bo + cout].astype(np.float64))
    return Ws, bs


def pack(gW, gb, c):
    lay, n = layout(c)
    th = np.zeros(n)
    for (cin, cout, k, wo, bo), w, b in zip(lay, gW, gb):
        th[wo:wo + w.size] = w.ravel()
        th[bo:bo + cout] = b
    return th


def bf16(a):
    """Round to bfloat16 (via float32, round to nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3(x, w):
    """Zero-padded 3x3 correlation: y[n,h,w,co] = sum_{dy,dx,ci} x[n,h+dy-1,w+dx-1,ci] w[co,dy,dx,ci]."""
    N, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((N, H, W, w.shape[0]))
    for dy in range(3):
        for dx in range(3):
            y += np.tensordot(xp[:, dy:dy + H, dx:dx + W, :], w[:, dy, dx, :], axes=([3], [1]))
    return y


def conv3_wgrad(g, x):
    """dW[co,dy,dx,ci] = sum_p g[p,co] x[p + (dy-1, dx-1), ci] (zero outside)."""
    N, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((g.shape[3], 3, 3, x.shape[3]))
    for dy in range(3):
        for dx in range(3):
            dw[:, dy, dx, :] = np.tensordot(g, xp[:, dy:dy + H, dx:dx + W, :], axes=([0, 1, 2], [0, 1, 2]))
    return dw


def flip_t(w):
    """Filter of the input gradient: wt[ci, dy, dx, co] = w[co, 2-dy, 2-dx, ci]."""
    return np.ascontiguousarray(w[:, ::-1, ::-1, :].transpose(3, 1, 2, 0))


def pool(x):
    N, H, W, C = x.shape
    return 0.25 * x.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def pool_bwd(g):
    """Adjoint of pool (each input gets a quarter of its parent's gradient)."""
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(g):
    N, H, W, C = g.shape
    return g.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def relu(x):
    return np.maximum(x, 0.0)


def image_scale(x):
    """Per-image rms (float32, as the device computes it), floored at 1e-30."""
    s = np.sqrt(np.mean(np.asarray(x, np.float64) ** 2, axis=(1, 2)))
    return np.maximum(s, 1e-30).astype(np.float32).astype(np.float64)


def forward(theta, x, c, quant=False):
    """Returns (out, cache).  x (N, H, W); out = s * net(x / s)."""
    q = bf16 if quant else (lambda a: a)
    Ws, bs = unpack(np.asarray(theta, np.float64), c)
    Wq = [Ws[0]] + [q(w) for w in Ws[1:14]] + [Ws[14]]
    s = image_scale(x)
    xn = np.asarray(x, np.float64) / s[:, None, None]
    if quant:
        xn = xn.astype(np.float32).astype(np.float64)
    A = {}
    A["a1"] = q(relu(conv3(xn[..., None], Wq[0]) + bs[0]))
    A["s1"] = q(relu(conv3(A["a1"], Wq[1]) + bs[1]))
    A["p1"] = q(pool(A["s1"]))
    A["a2"] = q(relu(conv3(A["p1"], Wq[2]) + bs[2]))
    A["s2"] = q(relu(conv3(A["a2"], Wq[3]) + bs[3]))
    A["p2"] = q(pool(A["s2"]))
    A["a3"] = q(relu(conv3(A["p2"], Wq[4]) + bs[4]))
    A["s3"] = q(relu(conv3(A["a3"], Wq[5]) + bs[5]))
    A["p3"] = q(pool(A["s3"]))
    A["a4"] = q(relu(conv3(A["p3"], Wq[6]) + bs[6]))
    A["a4b"] = q(relu(conv3(A["a4"], Wq[7]) + bs[7]))
    A["cat3"] = np.concatenate([up(A["a4b"]), A["s3"]], axis=3)
    A["d3a"] = q(relu(conv3(A["cat3"], Wq[8]) + bs[8]))
    A["d3b"] = q(relu(conv3(A["d3a"], Wq[9]) + bs[9]))
    A["cat2"] = np.concatenate([up(A["d3b"]), A["s2"]], axis=3)
    A["d2a"] = q(relu(conv3(A["cat2"], Wq[10]) + bs[10]))
    A["d2b"] = q(relu(conv3(A["d2a"], Wq[11]) + bs[11]))
    A["cat1"] = np.concatenate([up(A["d2b"]), A["s1"]], axis=3)
    A["d1a"] = q(relu(conv3(A["cat1"], Wq[12]) + bs[12]))
    A["d1b"] = q(relu(conv3(A["d1a"], Wq[13]) + bs[13]))
    on = A["d1b"] @ Ws[14][:, 0, 0, :].T[:, 0] + bs[14][0]
    cache = dict(A=A, Ws=Ws, Wq=Wq, bs=bs, s=s, xn=xn, on=on, q=q, c=c)
    return on * s[:, None, None], cache


def loss_and_grad(theta, x, target, c, quant=False):
    """loss = mean((net(x/s) - target/s)^2) and its gradient (parameter vector)."""
    out, K = forward(theta, x, c, quant)
    A, Wq, Ws, q, s = K["A"], K["Wq"], K["Ws"], K["q"], K["s"]
    d = K["on"] - np.asarray(target, np.float64) / s[:, None, None]
    NP = d.size
    loss = np.sum(d * d) / NP
    gon = 2.0 * d / NP
    if quant:
        gon = gon.astype(np.float32).astype(np.float64)
    gW = [None] * 15
    gb = [None] * 15
    wh = Ws[14][:, 0, 0, :][0]                     # (c,)
    gW[14] = np.einsum("nhw,nhwc->c", gon, A["d1b"]).reshape(1, 1, 1, -1)
    gb[14] = np.array([gon.sum()])
    g = q(gon[..., None] * wh * (A["d1b"] > 0))

    def layer(l, g_pre, X, mask):
        """filter / bias gradient of layer l, and its (masked) input gradient."""
        gW[l] = conv3_wgrad(g_pre, X)
        gb[l] = g_pre.sum(axis=(0, 1, 2))
        gx = conv3(g_pre, flip_t(Wq[l]))
        if mask is not None:
            gx = gx * (mask > 0)
        return q(gx)

    C = c
    g = layer(13, g, A["d1a"], A["d1a"])
    gcat1 = layer(12, g, A["cat1"], None)
    g = q(up_bwd(gcat1[..., :2 * C]) * (A["d2b"] > 0))
    g = layer(11, g, A["d2a"], A["d2a"])
    gcat2 = layer(10, g, A["cat2"], None)
    g = q(up_bwd(gcat2[..., :4 * C]) * (A["d3b"] > 0))
    g = layer(9, g, A["d3a"], A["d3a"])
    gcat3 = layer(8, g, A["cat3"], None)
    g = q(up_bwd(gcat3[..., :8 * C]) * (A["a4b"] > 0))
    g = layer(7, g, A["a4"], A["a4"])
    gp3 = layer(6, g, A["p3"], None)
    g = q((gcat3[..., 8 * C:] + pool_bwd(gp3)) * (A["s3"] > 0))
    g = layer(5, g, A["a3"], A["a3"])
    gp2 = layer(4, g, A["p2"], None)
    g = q((gcat2[..., 4 * C:] + pool_bwd(gp2)) * (A["s2"] > 0))
    g = layer(3, g, A["a2"], A["a2"])
    gp1 = layer(2, g, A["p1"], None)
    g = q((gcat1[..., 2 * C:] + pool_bwd(gp1)) * (A["s1"] > 0))
    g = layer(1, g, A["a1"], A["a1"])
    gW[0] = conv3_wgrad(g, K["xn"][..., None])
    gb[0] = g.sum(axis=(0, 1, 2))
    return loss, pack(gW, gb, c), out


def he_init(c, seed):
    """A random parameter vector (He-normal filters, small random biases) for tests."""
    g = np.random.default_rng(seed)
    lay, n = layout(c)
    th = np.zeros(n)
    for cin, cout, k, wo, bo in lay:
        th[wo:wo + cout * k * k * cin] = g.standard_normal(cout * k * k * cin) * np.sqrt(2.0 / (k * k * cin))
        th[bo:bo + cout] = 0.05 * g.standard_normal(cout)
    return th
