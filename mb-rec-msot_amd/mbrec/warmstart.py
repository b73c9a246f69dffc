"""Learned warm start for the NNLS-L2 loop (no counterpart in the reference).

The reference's loops start from zero (nnls_L2_less.m:28-29).  Here a small
convolutional U-Net (libmbrec, BF16 matrix-core kernels with FP32
accumulation) maps the back-projected image A^T y to an initial estimate,
which seeds the outer iterate of ``rec_nn_with_L2_reg`` (hook at
rec_nn_with_L2_reg.m:41, where A^T y is already available).  The solver always
finishes from the network's output, so the result is an NNLS solution
whatever the network predicts; the plain cold-start solve stays both the
fallback and the judge of the warm start's error.

Training data are produced by the same plain solver from seeded phantoms
(``make_training_set``); the network is trained on the device by its own
forward / backward kernels (``UNet.train_step``).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib


class UNet:
    """Device U-Net (4 levels, base width 32 or 64) operating on images
    (B, Nx, Nz) -- the model's (B, P) image vectors viewed as 2-D arrays."""

    NAMES = ["e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "ba", "bb",
             "d3a", "d3b", "d2a", "d2b", "d1a", "d1b", "head"]

    def __init__(self, base_channels: int = 32, seed: int = 0):
        L = _lib.lib()
        h = C.c_void_p()
        _lib.check(L.mbrec_unet_create(C.byref(h), int(base_channels), C.c_ulonglong(int(seed))))
        self.h = h
        self.base_channels = int(base_channels)
        self.num_params = int(L.mbrec_unet_num_params(h))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().mbrec_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters --------------------------------------------------------------
    def layout(self):
        """[(name, cin, cout, k, w_off, b_off)] of the 15 layers."""
        L = _lib.lib()
        out = []
        for i, name in enumerate(self.NAMES):
            ci, co, k = C.c_int32(), C.c_int32(), C.c_int32()
            wo, bo = C.c_size_t(), C.c_size_t()
            _lib.check(L.mbrec_unet_layout(self.h, i, C.byref(ci), C.byref(co), C.byref(k), C.byref(wo),
                                           C.byref(bo)))
            out.append((name, ci.value, co.value, k.value, wo.value, bo.value))
        return out

    def get_params(self) -> np.ndarray:
        th = np.empty(self.num_params, dtype=np.float32)
        _lib.check(_lib.lib().mbrec_unet_get_params(self.h, th.ctypes.data, th.size))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {th.size}")
        _lib.check(_lib.lib().mbrec_unet_set_params(self.h, th.ctypes.data, th.size))

    def save(self, path):
        np.savez(path, theta=self.get_params(), base_channels=np.int64(self.base_channels))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        net = cls(int(z["base_channels"]))
        net.set_params(z["theta"])
        return net

    # compute -------------------------------------------------------------------
    @staticmethod
    def _img(x):
        import torch
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32)).cuda() if isinstance(x, np.ndarray) else x
        if t.dtype != torch.float32:
            raise TypeError("images must be float32")
        if t.dim() == 2:
            t = t[None]
        return t.contiguous()

    def forward(self, x):
        """x: (B, H, W) float32 device tensor (or numpy) -> (B, H, W) estimate."""
        import torch
        t = self._img(x)
        B, H, W = t.shape
        out = torch.empty_like(t)
        _lib.check(_lib.lib().mbrec_unet_forward(self.h, t.data_ptr(), out.data_ptr(), B, H, W, _lib.stream_handle()))
        return out

    def gradients(self, x, target):
        """(loss, gradient vector) of mean((net(x/s) - target/s)^2)."""
        t, g = self._img(x), self._img(target)
        if t.shape != g.shape:
            raise ValueError("input and target shapes differ")
        B, H, W = t.shape
        grad = np.empty(self.num_params, dtype=np.float32)
        loss = C.c_double()
        _lib.check(_lib.lib().mbrec_unet_gradients(self.h, t.data_ptr(), g.data_ptr(), B, H, W, grad.ctypes.data,
                                                   C.byref(loss), _lib.stream_handle()))
        return float(loss.value), grad

    def train_step(self, x, target, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, want_loss=True):
        t, g = self._img(x), self._img(target)
        B, H, W = t.shape
        loss = C.c_double()
        _lib.check(_lib.lib().mbrec_unet_train_step(self.h, t.data_ptr(), g.data_ptr(), B, H, W, float(lr),
                                                    float(beta1), float(beta2), float(eps),
                                                    C.byref(loss) if want_loss else None, _lib.stream_handle()))
        return float(loss.value) if want_loss else None


# ---------------------------------------------------------------------------
# training data, training loop, warm-started solve
# ---------------------------------------------------------------------------
def phantom_batch(disc, n, seed, kinds=("disks", "ellipses", "vessels")):
    """n seeded phantoms (n, Nx, Nz) cycling through the phantom kinds."""
    from . import phantoms
    gen = {"disks": lambda s: phantoms.disks(disc, n=20, seed=s),
           "ellipses": lambda s: phantoms.ellipses(disc, n=12, seed=s),
           "vessels": lambda s: phantoms.vessel_tree(disc, seed=s)}
    return np.stack([gen[kinds[i % len(kinds)]](seed + i) for i in range(n)]).astype(np.float32)


def make_training_set(model, n, seed, num_iter, noise=0.01, lambda_rel=1e-2, batch=16, kinds=None):
    """(inputs A^T y, targets) as (n, Nx, Nz) float32 device tensors.

    Each phantom x gives y = A x + noise (Gaussian, ``noise`` * max|A x|); the
    target is the plain cold-start NNLS-L2 solution after ``num_iter`` outer
    iterations with lambda = lambda_rel * ||A^T y|| (per sinogram)."""
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    disc = model.Discretization
    Nx, Nz = model.shape_image
    kinds = kinds or ("disks", "ellipses", "vessels")
    ins, tgs = [], []
    g = torch.Generator(device="cuda").manual_seed(int(seed))
    for b0 in range(0, n, batch):
        nb = min(batch, n - b0)
        ph = torch.as_tensor(phantom_batch(disc, nb, seed + b0, kinds)).cuda()
        y = model.applyForward(ph)
        y = (y + noise * y.abs().amax(dim=(1, 2), keepdim=True) *
             torch.randn(y.shape, generator=g, device="cuda")).contiguous()
        aty = model.applyTranspose(y).reshape(nb, -1)
        lam = lambda_rel * torch.linalg.vector_norm(aty.double(), dim=1)
        rec = rec_nn_with_L2_reg_batch(model, y, num_iter, lam.cpu().numpy())
        ins.append(aty.reshape(nb, Nx, Nz))
        tgs.append(rec.reshape(nb, Nx, Nz))
    return torch.cat(ins), torch.cat(tgs)


def train(net, inputs, targets, epochs=10, batch=8, lr=1e-3, seed=0, log=None):
    """Adam over shuffled mini-batches; returns the per-epoch mean losses."""
    import torch
    n = inputs.shape[0]
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(n)
        tot, cnt = 0.0, 0
        t0 = time.perf_counter()
        for b0 in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(order[b0:b0 + batch], device=inputs.device)
            loss = net.train_step(inputs[idx].contiguous(), targets[idx].contiguous(), lr=lr)
            tot += loss
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.5g} ({time.perf_counter() - t0:.1f} s)")
    return hist


BETA = 1e-3   # nnls_L2_less.m:34


def warm_start_estimate(net, model, y, lambdaL2):
    """Warm start (w0, v0) for nnls_L2_less from sinograms y (B, D, nT).

    w0 = max(net(A'y), 0).  v0 is the scaled dual that makes w0 a fixed point of
    the outer iteration if w0 were the solution: with g = A'(A w0 - y) + lambdaL2 w0
    (the gradient of the subproblem's data + L2 term, halved), the u-step optimality
    at u = w = x gives v = g / beta, and the w-step requires v = 0 where x > 0 and
    v >= 0 where x = 0; so v0 = max(g, 0) / beta on the zero set of w0, 0 elsewhere.
    Returns (w0, v0) as (B, Nx, Nz) device tensors."""
    import torch
    B = y.shape[0]
    Nx, Nz = model.shape_image
    aty = model.applyTranspose(y).reshape(B, Nx, Nz)
    w0 = torch.clamp(net.forward(aty), min=0.0)
    lam = torch.as_tensor(np.asarray(lambdaL2, dtype=np.float32), device=w0.device).reshape(-1, 1, 1)
    r = model.applyForward(w0) - y
    g = model.applyTranspose(r).reshape(B, Nx, Nz) + lam * w0
    v0 = torch.where(w0 > 0, torch.zeros_like(g), torch.clamp(g, min=0.0) / BETA)
    return w0.contiguous(), v0.contiguous()
