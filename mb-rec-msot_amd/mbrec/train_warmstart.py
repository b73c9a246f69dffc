"""Train the warm-start U-Net on the stationary states of the plain NNLS-L2 loop and
measure what it gains.

    python -m mbrec.train_warmstart --out weights.npz [--train 4096 --epochs 30 ...]

Setup (400 x 400 grid at 100 um, 256-element 40 mm / 174 deg arc, 2030 samples, no
SIR, gauspuls EIR, band 0.1-12 MHz): seeded disk / ellipse / vessel-tree phantoms, 1 %
Gaussian noise, lambdaL2 = 1e-2 ||A'y|| per sinogram.  Targets are the states s* at
which nnls_L2_less is stationary for the converged solution (warmstart.make_training_set).
After training, a held-out batch is solved from cold and from the network's state until
every item is within --tol of the converged solution; the outer iterations and the
times each takes are reported.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_model(N=400, h=1e-4, sir=False, sigma=None):
    from .discretization import Discretization
    from .model import MSOTmodel
    from .probe import Probe
    half = (N - 1) / 2 * h
    return MSOTmodel(Probe("arc256_r40_174", True, False), Discretization([-half, half], N), 1500.0, sir, True, 0,
                     1e5, 12e6, sigma, sigma_iterations=40)


def test_batch(model, n, seed, noise=0.01):
    """(y, lambda, phantoms) for n held-out sinograms."""
    import torch
    from .warmstart import phantom_batch
    ph = torch.as_tensor(phantom_batch(model.Discretization, n, seed)).cuda()
    y = model.applyForward(ph)
    g = torch.Generator(device="cuda").manual_seed(seed)
    y = (y + noise * y.abs().amax(dim=(1, 2), keepdim=True) * torch.randn(y.shape, generator=g, device="cuda"))
    y = y.contiguous()
    aty = model.applyTranspose(y).reshape(n, -1)
    lam = (1e-2 * torch.linalg.vector_norm(aty.double(), dim=1)).cpu().numpy()
    return y, lam, ph


def time_solve(model, y, lam, k, w0=None, v0=None, reps=2):
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    rec_nn_with_L2_reg_batch(model, y, k, lam, init=w0, init_dual=v0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        rec_nn_with_L2_reg_batch(model, y, k, lam, init=w0, init_dual=v0)
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--grid", type=int, default=400)
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--chunk", type=int, default=25)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--relative-loss", action="store_true",
                    help="weight each item by the inverse of its normalised target power")
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    import torch
    from .warmstart import (UNet, iterations_to_tolerance, make_training_set, nnls_l2_minimiser,
                            stationary_state, train, warm_start_estimate)
    t0 = time.perf_counter()
    model = build_model(a.grid)
    log(f"model: sigma {model.normalizationFactor:.6g} ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    xin, xt = make_training_set(model, a.train, 10_000 + 100_003 * a.seed, batch=64, log=log)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"training set: {a.train} stationary states in {t_data:.1f} s")
    net = UNet.load(a.init) if a.init else UNet(a.base, seed=a.seed)
    if a.relative_loss:
        net.set_relative_loss(True)
    t0 = time.perf_counter()
    hist = train(net, xin, xt, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, flip=True,
                 seed=a.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    net.save(a.out)
    log(f"trained {a.epochs} epochs in {t_train:.1f} s -> {a.out}")
    del xin, xt
    # evaluation on a held-out batch
    y, lam, _ = test_batch(model, a.test, 90_000 + a.seed)
    t0 = time.perf_counter()
    ref, kkt = nnls_l2_minimiser(model, y, lam)
    s_ref = stationary_state(model, y, lam, ref)
    log(f"reference minimiser: KKT residual {kkt:.2e} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w0, v0, used = warm_start_estimate(net, model, y, lam)
    torch.cuda.synchronize()
    t_net = time.perf_counter() - t0
    s0 = (w0 - v0).double()
    rel = lambda a_, b_: (torch.linalg.vector_norm((a_ - b_).reshape(a.test, -1), dim=1) /
                          torch.linalg.vector_norm(b_.reshape(a.test, -1), dim=1))
    e_state = rel(s0, s_ref)
    e_img = rel(w0.double(), ref)
    log(f"network: state error median {e_state.median():.3f} max {e_state.max():.3f}; image error median "
        f"{e_img.median():.3f}; {int(used.sum())}/{a.test} items use it; {t_net * 1e3:.1f} ms")
    k_cold, errs_cold = iterations_to_tolerance(model, y, lam, ref, a.tol, chunk=a.chunk)
    log(f"cold: {k_cold} outer iterations to {a.tol:g}")
    k_warm, errs_warm = iterations_to_tolerance(model, y, lam, ref, a.tol, w0, v0, chunk=a.chunk)
    log(f"learned: {k_warm} outer iterations to {a.tol:g}")
    rep = {"train": a.train, "epochs": a.epochs, "base_channels": a.base, "relative_loss": a.relative_loss,
           "loss": hist, "t_data_s": t_data,
           "t_train_s": t_train, "tol": a.tol, "chunk": a.chunk, "ref_kkt": kkt,
           "state_err_median": float(e_state.median()), "state_err_max": float(e_state.max()),
           "image_err_median": float(e_img.median()), "used": int(used.sum()),
           "k_cold": k_cold, "k_learned": k_warm, "errs_cold": errs_cold, "errs_learned": errs_warm,
           "t_net_s": t_net}
    if k_cold and k_warm:
        tc = time_solve(model, y, lam, k_cold)
        tw = time_solve(model, y, lam, k_warm, w0, v0) + t_net
        rep.update(t_cold_s=tc, t_learned_s=tw, speedup=tc / tw)
        log(f"times: cold {tc:.2f} s, learned {tw:.2f} s -> {tc / tw:.2f}x")
    log(json.dumps({k: v for k, v in rep.items() if k not in ("errs_cold", "errs_learned", "loss")}))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
