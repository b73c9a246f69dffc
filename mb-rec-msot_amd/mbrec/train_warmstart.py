"""Train the warm-start U-Net on plain NNLS-L2 solutions and measure what it gains.

    python -m mbrec.train_warmstart --out weights.npz [--train 512 --epochs 20 ...]

Setup (400 x 400 grid at 100 um, 256-element 40 mm / 174 deg arc, 2030 samples,
no SIR, gauspuls EIR, band 0.1-12 MHz): seeded disk / ellipse / vessel-tree
phantoms, 1 % Gaussian noise, lambdaL2 = 1e-2 ||A'y|| per sinogram.  Targets
are plain cold-start solves of --target-iters outer iterations.  After
training, a held-out batch is solved from cold and from the network's
estimate, and the iterations each needs to come within --tol (relative L2) of
a long plain solve (--ref-iters) are reported with the times they take.
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


def iterations_to_tol(model, y, lam, ref, init, tol, grid):
    """Smallest k in grid whose solve is within tol of ref for every sinogram (init =
    None or (w0, v0)); the per-k worst relative error."""
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    B = y.shape[0]
    errs = {}
    nref = torch.linalg.vector_norm(ref.reshape(B, -1).double(), dim=1)
    w0, v0 = init if init is not None else (None, None)
    for k in grid:
        r = rec_nn_with_L2_reg_batch(model, y, k, lam, init=w0, init_dual=v0)
        e = (torch.linalg.vector_norm((r - ref).reshape(B, -1).double(), dim=1) / nref).max().item()
        errs[k] = e
        log(f"  {'cold' if init is None else ('warm' if w0 is not None and v0 is not None else 'primal')} "
            f"{k} iterations: worst rel. error {e:.3e}")
        if e <= tol:
            return k, errs
    return None, errs


def time_solve(model, y, lam, k, init=None, reps=2):
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    w0, v0 = init if init is not None else (None, None)
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
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--target-iters", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--ref-iters", type=int, default=400)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--report", default=None)
    ap.add_argument("--grid-iters", default="10,20,30,50,75,100,150,200,300,400,600,800,1200,1600")
    a = ap.parse_args(argv)
    import torch
    from .warmstart import UNet, make_training_set, train, warm_start_estimate
    t0 = time.perf_counter()
    model = build_model(a.grid)
    log(f"model: sigma {model.normalizationFactor:.6g} ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    xin, xt = make_training_set(model, a.train, 10_000 + a.seed, a.target_iters, batch=32)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"training set: {a.train} solves of {a.target_iters} iterations in {t_data:.1f} s")
    net = UNet(32, seed=a.seed)
    t0 = time.perf_counter()
    hist = train(net, xin, xt, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    net.save(a.out)
    log(f"trained {a.epochs} epochs in {t_train:.1f} s -> {a.out}")
    del xin, xt
    # evaluation on a held-out batch
    y, lam, ph = test_batch(model, a.test, 90_000 + a.seed)
    from .solvers import rec_nn_with_L2_reg_batch
    log(f"reference: {a.ref_iters} iterations of the plain solve ...")
    t0 = time.perf_counter()
    ref = rec_nn_with_L2_reg_batch(model, y, a.ref_iters, lam)
    torch.cuda.synchronize()
    log(f"reference done in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    est = warm_start_estimate(net, model, y, lam)
    torch.cuda.synchronize()
    t_net = time.perf_counter() - t0
    e_net = (torch.linalg.vector_norm((est[0].reshape(a.test, -1) - ref.reshape(a.test, -1)).double(), dim=1) /
             torch.linalg.vector_norm(ref.reshape(a.test, -1).double(), dim=1))
    grid = [int(k) for k in a.grid_iters.split(",")]
    k_cold, errs_cold = iterations_to_tol(model, y, lam, ref, None, a.tol, grid)
    k_primal, errs_primal = iterations_to_tol(model, y, lam, ref, (est[0], None), a.tol, grid)
    k_warm, errs_warm = iterations_to_tol(model, y, lam, ref, est, a.tol, grid)
    rep = {"train": a.train, "target_iters": a.target_iters, "epochs": a.epochs, "loss": hist,
           "t_data_s": t_data, "t_train_s": t_train, "ref_iters": a.ref_iters, "tol": a.tol,
           "net_err_vs_ref_median": float(e_net.median()), "net_err_vs_ref_max": float(e_net.max()),
           "k_cold": k_cold, "k_warm": k_warm, "k_primal_only": k_primal, "errs_cold": errs_cold,
           "errs_warm": errs_warm, "errs_primal_only": errs_primal,
           "t_net_s": t_net}
    if k_cold and k_warm:
        tc = time_solve(model, y, lam, k_cold)
        tw = time_solve(model, y, lam, k_warm, init=est) + t_net
        rep.update(t_cold_s=tc, t_warm_s=tw, speedup=tc / tw)
    log(json.dumps(rep))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
