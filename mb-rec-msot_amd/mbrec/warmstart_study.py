"""How accurate a warm start of nnls_L2_less must be: the loop started from perturbed
exact states.

    python -m mbrec.warmstart_study [--batch 16] [--report study.json]

Setup as mbrec.train_warmstart (400 x 400 at 100 um, 256 ch x 2030, no SIR, lambdaL2 =
1e-2 ||A'y||, held-out seeds 90000..).  The minimiser x* (nnls_l2_minimiser) and its
stationary state s* (stationary_state) are computed; the loop is then started from
  - the cold start (w = v = 0, nnls_L2_less.m:28-29),
  - the exact state s*,
  - the exact image x* with v = 0 (what seeding only the image gives),
  - s* (1 + eps n) and, for comparison, x* (1 + eps n) with its KKT dual,
with n a smooth field (Gaussian noise averaged over 9 x 9 pixels, unit standard deviation),
and s* (1 + eps w) with w white Gaussian noise (which error scales the loop removes slowly),
and the worst relative L2 error to x* is reported after k outer iterations.
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--seed", type=int, default=90_000)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    import torch
    from .solvers import rec_nn_with_L2_reg_batch
    from .train_warmstart import build_model, test_batch
    from .warmstart import BETA, nnls_l2_minimiser, objective_gradient, stationary_state
    t0 = time.perf_counter()
    model = build_model(400)
    B = a.batch
    y, lam, _ = test_batch(model, B, a.seed)
    xs, kkt = nnls_l2_minimiser(model, y, lam, max_iter=1000, tol=1e-7)
    s = stationary_state(model, y, lam, xs)
    nx = torch.linalg.vector_norm(xs.reshape(B, -1), dim=1)
    nv = torch.linalg.vector_norm(torch.clamp(-s, min=0.0).reshape(B, -1), dim=1)
    log(f"minimiser: KKT residual {kkt:.2e} ({time.perf_counter() - t0:.1f} s); |x*| median {nx.median():.3g}, "
        f"|v*| median {nv.median():.3g}, active fraction {(xs > 0).double().mean():.3f}")

    def err(r):
        return (torch.linalg.vector_norm((r.double() - xs).reshape(B, -1), dim=1) / nx).max().item()

    def curve(name, w0, v0, grid):
        out = {}
        for k in grid:
            out[k] = err(rec_nn_with_L2_reg_batch(model, y, k, lam, init=w0, init_dual=v0))
        log(f"{name}: " + " ".join(f"{k}:{e:.2e}" for k, e in out.items()))
        return out

    f32 = lambda t: t.float().contiguous()
    res = {"kkt": kkt, "x_norm_median": float(nx.median()), "v_norm_median": float(nv.median())}
    grid = [10, 50, 100, 200, 300, 400, 600, 800, 1200]
    res["cold"] = curve("cold", None, None, grid)
    res["exact_state"] = curve("exact s*", f32(torch.clamp(s, min=0.0)), f32(torch.clamp(-s, min=0.0)), [1, 10, 100])
    res["exact_image_zero_dual"] = curve("x*, v = 0", f32(xs), None, [10, 100, 400, 1200])
    gen = torch.Generator(device="cuda").manual_seed(5)
    Nx, Nz = model.shape_image
    for eps in (0.3, 0.1, 0.03, 0.01):
        n = torch.randn(B, 1, Nx, Nz, generator=gen, device="cuda", dtype=torch.float64)
        n = torch.nn.functional.conv2d(n, torch.full((1, 1, 9, 9), 1.0 / 81, device="cuda", dtype=torch.float64),
                                       padding=4)[:, 0]
        n = n / n.reshape(B, -1).std(dim=1).view(B, 1, 1)
        s0 = s * (1 + eps * n)
        res[f"state_eps_{eps}"] = curve(f"s* (1 + {eps} n)", f32(torch.clamp(s0, min=0.0)),
                                        f32(torch.clamp(-s0, min=0.0)), grid)
        wn = torch.randn(B, Nx, Nz, generator=gen, device="cuda", dtype=torch.float64)
        s0 = s * (1 + eps * wn)
        res[f"state_white_eps_{eps}"] = curve(f"s* (1 + {eps} white)", f32(torch.clamp(s0, min=0.0)),
                                              f32(torch.clamp(-s0, min=0.0)), grid)
        x0 = torch.clamp(xs * (1 + eps * n), min=0.0)
        g0 = objective_gradient(model, y, lam, x0)
        v0 = torch.where(x0 <= 0, torch.clamp(g0, min=0.0) / BETA, torch.zeros_like(g0))
        res[f"image_kkt_eps_{eps}"] = curve(f"x* (1 + {eps} n) + KKT dual", f32(x0), f32(v0), grid)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
