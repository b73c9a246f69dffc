"""Learned warm start for the NNLS-L2 loop (no counterpart in the reference).

The reference's loops start from zero (nnls_L2_less.m:28-29).  Here a small
convolutional U-Net (libmbrec, BF16 matrix-core kernels with FP32
accumulation) maps the back-projected image A^T y to an initial estimate,
which seeds the outer iterate of ``rec_nn_with_L2_reg`` (hook at
rec_nn_with_L2_reg.m:41, where A^T y is already available).  The solver always
finishes from the network's output, so the result is an NNLS solution
whatever the network predicts; the plain cold-start solve stays both the
fallback and the arbiter of the warm start's error.

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

    def set_relative_loss(self, on=True):
        """Weight every item's squared error by the inverse of its normalised target power
        (mean over the batch of the powers / its own), so that items count by their relative
        error; off: the plain mean."""
        _lib.check(_lib.lib().mbrec_unet_set_relative_loss(self.h, 1 if on else 0))

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


BETA = 1e-3   # nnls_L2_less.m:34 (rec_nn_with_L2_reg.m:34)


def _lam_tensor(lambdaL2, B, device):
    import torch
    return torch.as_tensor(np.asarray(lambdaL2, dtype=np.float64), device=device).reshape(B, 1, 1)


def objective_gradient(model, y, lam, x):
    """grad of F(x) = 1/2||A x - y||^2 + lambdaL2/2 ||x||^2 (the objective of
    rec_nn_with_L2_reg.m with RegL2 = identity) for images x (B, Nx, Nz), FP64 result."""
    B = y.shape[0]
    r = model.applyForward(x.float().contiguous()) - y
    return model.applyTranspose(r).reshape(x.shape).double() + _lam_tensor(lam, B, x.device) * x.double()


def objective(model, y, lam, x):
    """F(x) per item (FP64), as above."""
    import torch
    B = y.shape[0]
    r = (model.applyForward(x.float().contiguous()) - y).double()
    lam_t = torch.as_tensor(np.asarray(lam, dtype=np.float64), device=x.device)
    return 0.5 * (r.reshape(B, -1) ** 2).sum(dim=1) + 0.5 * lam_t * (x.double().reshape(B, -1) ** 2).sum(dim=1)


def nnls_l2_minimiser(model, y, lam, max_iter=400, tol=2e-7, log=None):
    """The minimiser of F over x >= 0 -- the fixed point nnls_L2_less converges to
    (rec_nn_with_L2_reg.m:39-41, RegL2 = identity; F is strictly convex for lambdaL2 > 0,
    so the minimiser is unique) -- by accelerated projected gradient (Beck & Teboulle's
    FISTA with O'Donoghue & Candes' gradient restart), step 1/L with L = 1.02 + lambdaL2
    >= ||A||^2 + lambdaL2 for the sigma_max-normalised model.  Vectors FP64, operators
    FP32.  Runs until the KKT residual ||x - max(x - grad F, 0)|| / ||x|| <= tol for every
    item (about 150 iterations at 400^2) or max_iter.  Returns (x (B, Nx, Nz) FP64, kkt)."""
    import torch
    B = y.shape[0]
    Nx, Nz = model.shape_image
    dev = y.device
    L = 1.02 + _lam_tensor(lam, B, dev)
    x = torch.zeros((B, Nx, Nz), dtype=torch.float64, device=dev)
    z = x.clone()
    t = torch.ones((B, 1, 1), dtype=torch.float64, device=dev)
    kkt = float("inf")
    for k in range(max_iter):
        g = objective_gradient(model, y, lam, z)
        xn = torch.clamp(z - g / L, min=0.0)
        restart = ((z - xn) * (xn - x)).sum(dim=(1, 2), keepdim=True) > 0
        tn = (1 + torch.sqrt(1 + 4 * t * t)) / 2
        mom = torch.where(restart, torch.zeros_like(t), (t - 1) / tn)
        t = torch.where(restart, torch.ones_like(t), tn)
        z = xn + mom * (xn - x)
        x = xn
        if k % 10 == 9:
            g = objective_gradient(model, y, lam, x)
            pg = (x - torch.clamp(x - g, min=0.0)).reshape(B, -1)
            kkt = (torch.linalg.vector_norm(pg, dim=1) /
                   torch.linalg.vector_norm(x.reshape(B, -1), dim=1).clamp_min(1e-300)).max().item()
            if log and k % 50 == 49:
                log(f"  minimiser: {k + 1} iterations, KKT residual {kkt:.2e}")
            if kkt <= tol:
                break
    return x, kkt


def stationary_state(model, y, lam, x):
    """The state s = w - v of nnls_L2_less at which it is stationary for the minimiser x:
    after every outer iteration w = max(s, 0) and v = max(-s, 0) with s = u - v
    (nnls_L2_less.m:68-72); at the fixed point u = w = x and the u-step's optimality
    (grad F(x) = beta v, :56 with the stacked operator) gives v = grad F(x) / beta on
    {x = 0}, 0 elsewhere.  Returns s (B, Nx, Nz) FP64."""
    import torch
    g = objective_gradient(model, y, lam, x)
    return torch.where(x > 0, x, -torch.clamp(g, min=0.0) / BETA)


def make_training_set(model, n, seed, batch=64, noise=0.01, lambda_rel=1e-2, kinds=None, log=None):
    """(inputs A^T y, targets s*) as (n, Nx, Nz) float32 device tensors.

    Each seeded phantom x gives y = A x + noise (Gaussian, ``noise`` * max|A x|) and
    lambdaL2 = lambda_rel * ||A^T y||; the target is the stationary state s* of
    nnls_L2_less for the converged solution (nnls_l2_minimiser, KKT residual <= 2e-7):
    the network predicts the whole state of the loop -- the image where it is positive
    and the scaled dual v = grad F / beta where it is zero."""
    import torch
    disc = model.Discretization
    Nx, Nz = model.shape_image
    kinds = kinds or ("disks", "ellipses", "vessels")
    ins = torch.empty((n, Nx, Nz), dtype=torch.float32, device="cuda")
    tgs = torch.empty_like(ins)
    g = torch.Generator(device="cuda").manual_seed(int(seed))
    t0 = time.perf_counter()
    for b0 in range(0, n, batch):
        nb = min(batch, n - b0)
        ph = torch.as_tensor(phantom_batch(disc, nb, seed + b0, kinds)).cuda()
        y = model.applyForward(ph)
        y = (y + noise * y.abs().amax(dim=(1, 2), keepdim=True) *
             torch.randn(y.shape, generator=g, device="cuda")).contiguous()
        aty = model.applyTranspose(y).reshape(nb, Nx, Nz)
        lam = (lambda_rel * torch.linalg.vector_norm(aty.reshape(nb, -1).double(), dim=1)).cpu().numpy()
        x, _ = nnls_l2_minimiser(model, y, lam)
        ins[b0:b0 + nb] = aty
        tgs[b0:b0 + nb] = stationary_state(model, y, lam, x).float()
        if log:
            log(f"training set: {b0 + nb}/{n} ({time.perf_counter() - t0:.1f} s)")
    return ins, tgs


def train(net, inputs, targets, epochs=10, batch=8, lr=1e-3, lr_final=None, flip=False, seed=0, log=None):
    """Adam over shuffled mini-batches; the learning rate decays geometrically from lr to
    lr_final (default lr) over the epochs; flip=True mirrors each mini-batch along x with
    probability 1/2 (the arc probe is mirror-symmetric about the z axis, so A'A commutes
    with the mirror).  Returns the per-epoch mean losses."""
    import torch
    n = inputs.shape[0]
    rng = np.random.default_rng(seed)
    hist = []
    lr_final = lr if lr_final is None else lr_final
    for ep in range(epochs):
        lr_ep = lr * (lr_final / lr) ** (ep / max(epochs - 1, 1))
        order = rng.permutation(n)
        tot, cnt = 0.0, 0
        t0 = time.perf_counter()
        for b0 in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(order[b0:b0 + batch], device=inputs.device)
            xin, xt = inputs[idx], targets[idx]
            if flip and rng.uniform() < 0.5:
                xin, xt = torch.flip(xin, dims=(1,)), torch.flip(xt, dims=(1,))
            loss = net.train_step(xin.contiguous(), xt.contiguous(), lr=lr_ep)
            tot += loss
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.5g} lr {lr_ep:.2e} ({time.perf_counter() - t0:.1f} s)")
    return hist


def warm_start_estimate(net, model, y, lambdaL2):
    """Warm start (w0, v0, used) of nnls_L2_less from sinograms y (B, D, nT).

    The network maps A'y to the loop's state s0; w0 = max(s0, 0) and v0 = max(-s0, 0)
    (the form every outer iteration leaves them in, nnls_L2_less.m:68-72).  Fallback to
    the cold start (w0 = v0 = 0, nnls_L2_less.m:28-29), per item: when s0 is not finite,
    or when w0 scores a larger objective F than the cold start's w = 0.  ``used`` (bool
    array [B]) says which items start from the network's state."""
    import torch
    B = y.shape[0]
    Nx, Nz = model.shape_image
    aty = model.applyTranspose(y).reshape(B, Nx, Nz)
    s0 = net.forward(aty)
    finite = torch.isfinite(s0).reshape(B, -1).all(dim=1)
    s0 = torch.where(finite.reshape(B, 1, 1), s0, torch.zeros_like(s0))
    w0 = torch.clamp(s0, min=0.0)
    v0 = torch.clamp(-s0, min=0.0)
    f_warm = objective(model, y, lambdaL2, w0)
    f_cold = 0.5 * (y.double().reshape(B, -1) ** 2).sum(dim=1)
    used = finite & (f_warm <= f_cold)
    keep = used.reshape(B, 1, 1)
    w0 = torch.where(keep, w0, torch.zeros_like(w0))
    v0 = torch.where(keep, v0, torch.zeros_like(v0))
    return w0.contiguous(), v0.contiguous(), used.cpu().numpy()


def arbitrate_warm_start(model, y, lambdaL2, w0, v0, used, k=10):
    """The plain cold start as the arbiter of a warm start: both run k outer iterations of
    nnls_L2_less; per item the run whose last iteration moved the state s = w - v less is
    kept (for a contraction the distance to the fixed point scales with the step).
    Returns (w, v, used) after k iterations -- passing w, v back as init / init_dual
    continues that item's iteration sequence -- with used False where the cold run won."""
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    B = y.shape[0]

    def run(wa, va):
        w1, v1 = rec_nn_with_L2_reg_batch(model, y, k - 1, lambdaL2, init=wa, init_dual=va, return_dual=True)
        w2, v2 = rec_nn_with_L2_reg_batch(model, y, 1, lambdaL2, init=w1, init_dual=v1, return_dual=True)
        s1, s2 = (w1 - v1).double().reshape(B, -1), (w2 - v2).double().reshape(B, -1)
        return w2, v2, torch.linalg.vector_norm(s2 - s1, dim=1)

    ww, vw, sw = run(w0, v0)
    wc, vc, sc = run(None, None)
    keep = torch.as_tensor(used, device=ww.device) & torch.isfinite(sw) & (sw <= sc)
    sel = keep.reshape(B, 1, 1)
    return (torch.where(sel, ww, wc).contiguous(), torch.where(sel, vw, vc).contiguous(),
            keep.cpu().numpy())


def rec_nn_with_L2_reg_learned(net, model, y, num_iter, lambdaL2, arbiter_iters=10):
    """B independent rec_nn_with_L2_reg.m solves (RegL2 = identity) of num_iter outer
    iterations, each started from the network's state (warm_start_estimate) instead of
    zero, with the cold start as fallback and arbiter (arbitrate_warm_start over the first
    arbiter_iters iterations; 0 skips the arbitration).  Returns (rec (B, Nx, Nz) device tensor,
    used)."""
    from .solvers import rec_nn_with_L2_reg_batch
    w0, v0, used = warm_start_estimate(net, model, y, lambdaL2)
    k = min(int(arbiter_iters), int(num_iter))
    if k >= 2:
        w0, v0, used = arbitrate_warm_start(model, y, lambdaL2, w0, v0, used, k)
        num_iter -= k
    return rec_nn_with_L2_reg_batch(model, y, num_iter, lambdaL2, init=w0, init_dual=v0), used


def iterations_to_tolerance(model, y, lam, ref, tol, w0=None, v0=None, chunk=25, max_iter=3000):
    """Outer iterations nnls_L2_less needs, from (w0, v0) (None: cold start), until every
    item is within ``tol`` relative L2 of ``ref``; the loop runs in chunks of ``chunk``
    iterations, each continuing from the previous chunk's full state (w, v), so the
    count is exact to a multiple of ``chunk``.  Returns (k or None, {k: worst error})."""
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    B = y.shape[0]
    nref = torch.linalg.vector_norm(ref.reshape(B, -1).double(), dim=1)
    w, v = w0, v0
    errs = {}
    k = 0
    while k < max_iter:
        w, v = rec_nn_with_L2_reg_batch(model, y, chunk, lam, init=w, init_dual=v, return_dual=True)
        k += chunk
        e = (torch.linalg.vector_norm((w.double() - ref.double()).reshape(B, -1), dim=1) / nref).max().item()
        errs[k] = e
        if e <= tol:
            return k, errs
    return None, errs
