"""Benchmark: NNLS-L2 reconstructions per second at 400^2 px x 256 channels on one MI355X,
learned warm start vs plain cold start, each to 1e-3 relative L2 of the converged solution.

Workload: a batch of B = 64 independent rec_nn_with_L2_reg solves (RegL2 = identity,
nnls_L2_less with 2 LSQR inner iterations, lambda = 1, beta = 1e-3) on the 400 x 400 /
100 um grid without SIR, 256-element 40 mm / 174 deg arc, fs 40 MHz, 2030 samples,
c 1500 m/s, delay 0, synthetic gauspuls EIR (fc 4 MHz, 100 %), band 0.1-12 MHz,
sigma_max by power iteration.  Phantoms: seeded disks / ellipses / vessel trees, 1 %
Gaussian noise, lambdaL2 = 1e-2 ||A'y|| per sinogram.  All inputs are resident in HBM
before the timed region.

The reference solution is the converged minimiser (the fixed point of the plain loop,
KKT residual ~1e-7).  The outer iterations each form needs for every item of the batch
to come within 1e-3 of it are found first (untimed, in chunks that continue from the
loop's full state); then one "step" = one batch solve: for the learned form the
network's state estimate from A'y (with its cold-start fallback) plus the warm-started
loop, for the plain form the loop from zero.  value = B / (time per learned step).

roofline: the dominant operator kernel (by time per outer iteration), its algorithmic
bytes and flops per launch over its average duration (HIP events on the launch stream),
labelled by the binding term.  cpu_baseline: the FP64 C oracle (same loops and OpenMP
decomposition as the reference MEX files) on a bounded sample of operator applications,
scaled to one plain solve of the same iteration count.

Other modes: --fixed-iters N (and --sir) time a fixed-iteration plain solve; --mixed the
224-sinogram mixed-grid batch; --tv600 the TV configuration; --secondary the secondary
paths; --mex the per-call time of the MEX drop-ins; --pair-b1 the one-sinogram operator
pair's launch choices; --copy-bw the box's device copy rate.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "mb-rec-msot_amd"), ROOT]

METRIC = "NNLS recons/sec 400²px×256ch at rel.err≤1e-3 (learned vs plain); op HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--iters", type=int, default=50, help="outer iterations of --mixed")
    ap.add_argument("--fixed-iters", type=int, default=0,
                    help="time a fixed-iteration plain solve instead of the 1e-3 metric (--sir implies 50)")
    ap.add_argument("--chunk", type=int, default=25, help="granularity of the iterations-to-1e-3 search")
    ap.add_argument("--weights", default=None, help="warm-start network weights (.npz)")
    ap.add_argument("--mex", action="store_true", help="per-call time of the four MEX drop-ins at 400^2")
    ap.add_argument("--pair-b1", action="store_true",
                    help="one-sinogram operator pair at 400^2 DAS: every forward choice and tile-gather "
                         "variant timed (HIP events) and checked against the default; prints one JSON line")
    ap.add_argument("--copy-bw", action="store_true",
                    help="time a 1 GiB device-to-device copy (the box's practical HBM rate); prints one JSON line")
    ap.add_argument("--mex-reps", type=int, default=20)
    ap.add_argument("--grid", type=int, default=400)
    ap.add_argument("--sir", action="store_true", help="400x400 with SIR (the SIR parity configuration)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--launch", default=None, metavar="GATHER,FORWARD",
                    help="pin the gather variant and forward choice instead of autotuning (profiling runs)")
    ap.add_argument("--cpu-sample-reps", type=int, default=100)
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="end with a plain 256 MiB device copy (reference for the FETCH/WRITE_SIZE counters)")
    ap.add_argument("--secondary", action="store_true",
                    help="time the secondary paths (dmas_cf, bp_kernel, backproject_waveeq, detector "
                         "upsampling, TV and L1 solves) on the device and their oracle on the host; "
                         "prints one JSON line")
    ap.add_argument("--mixed", action="store_true",
                    help="the mixed batch: 224 sinograms (28 wavelengths x 8 frames) over 200^2 / 300^2 / "
                         "400^2 grids with crop lengths U{1500..2030}, one graph; prints one JSON line")
    ap.add_argument("--tv600", action="store_true",
                    help="rec_nn_with_TV_reg at 600^2 / 60 um with SIR: 32 vessel-tree sinograms of 256 "
                         "detectors upsampled to 512 by spline; prints one JSON line")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU: turn two rocprofv3 --pmc counter_collection.csv files of a "
                         "--traffic-calibration run into profiles/traffic.json and exit")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


CALIB_BYTES = 256 * 1024 * 1024


def traffic_from(fetch_csv, write_csv, N=400, B=64):
    """HBM bytes per launch of the dominant kernels from rocprofv3 FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch), each scaled by what the counter reads for the 256 MiB copy."""
    import csv
    import statistics

    def per_kernel(path):
        acc = {}
        for r in csv.DictReader(open(path)):
            k = (r["Kernel_Name"], r["Dispatch_Id"])
            acc[k] = acc.get(k, 0.0) + float(r["Counter_Value"]) * 1024.0
        by = {}
        for (name, _), v in acc.items():
            by.setdefault(name, []).append(v)
        return by

    f, w = per_kernel(fetch_csv), per_kernel(write_csv)
    copy = [n for n in f if "copyBuffer" in n]
    cal_f = CALIB_BYTES / max(f[copy[0]]) if copy else 1.0
    cal_w = CALIB_BYTES / max(w[copy[0]]) if copy else 1.0
    path = os.path.join(ROOT, "profiles", "traffic.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    out["calibration"] = {"copy_bytes": CALIB_BYTES, "fetch_factor": round(cal_f, 4), "write_factor": round(cal_w, 4)}
    # every operator kernel of the run (median over its dispatches), under the workload
    per = {}
    for name in f:
        if name in w and ("k_fwd" in name or "k_gather" in name or "k_sum_parts" in name or "k_img_to_slots" in name):
            fb, wb = statistics.median(f[name]) * cal_f, statistics.median(w[name]) * cal_w
            per[name.split("(")[0]] = {"dispatches": len(f[name]), "fetch_bytes": fb, "write_bytes": wb,
                                       "hbm_bytes_per_launch": fb + wb}
    out[f"kernels_B{B}_N{N}"] = per
    if B > 1:   # the kernel the solve launches most often of each kind (the tuned variant)
        for kind, pat in (("scatter", "k_fwd_pull"), ("gather", "k_gather")):
            names = [n for n in f if pat in n and n in w]
            if not names:
                continue
            name = max(names, key=lambda n: len(f[n]))
            fb, wb = statistics.median(f[name]) * cal_f, statistics.median(w[name]) * cal_w
            out[f"das_{kind}_B{B}_N{N}"] = {"kernel": name.split("(")[0], "dispatches": len(f[name]),
                                            "fetch_bytes": fb, "write_bytes": wb, "hbm_bytes_per_launch": fb + wb}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def secondary(a):
    """Secondary paths at the 400^2 / 256 x 2030 sizes: device time (HIP events on the
    launch stream, inputs resident) next to the FP64 oracle on the host cores."""
    import numpy as np
    import torch

    from mbrec import _lib
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.phantoms import add_noise, disks
    from mbrec.preprocess import spline_upsampling_matrix
    from mbrec.probe import Probe
    from mbrec.solvers import rec_nn_with_L1_eye_reg, rec_nn_with_TV_reg
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    L = _lib.lib()
    st = _lib.stream_handle
    C_, FS, COV, RAD = 1500.0, 4e7, 174.0, 0.040
    N, h, D, nT = a.grid, 1e-4, 256, 2030
    v = (np.arange(N) - (N - 1) / 2) * h
    Xm, Zm = np.meshgrid(v, v)
    X, Z = Xm.T.copy(), Zm.T.copy()   # (Nx, Nz)
    sig = np.random.default_rng(11).standard_normal((1, D, nT)).astype(np.float32)
    res = {}

    def dev_time(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def host_time(fn):
        t = time.perf_counter()
        out = fn()
        return (time.perf_counter() - t) * 1e3, out

    def entry(name, gpu_ms, cpu_ms, note, err=None):
        res[name] = {"gpu_ms": round(gpu_ms, 4), "cpu_ms": round(cpu_ms, 2), "cpu_cores": O.num_threads(),
                     "cpu_kind": "port", "speedup": round(cpu_ms / gpu_ms, 1), "workload": note}
        if err is not None:
            res[name]["rel_l2_vs_oracle"] = float(err)
        log(f"{name}: device {gpu_ms:.3f} ms, oracle {cpu_ms:.1f} ms ({note})")

    rel = lambda x, y: float(np.linalg.norm(np.ravel(x - y)) / max(np.linalg.norm(np.ravel(y)), 1e-300))
    # dmas_cf (C2 sinogram, W = 1)
    dx = torch.tensor(X.ravel()).cuda()
    dz = torch.tensor(Z.ravel()).cuda()
    ds = torch.tensor(sig).cuda()
    out = torch.empty((4, 1, N * N), dtype=torch.float64, device="cuda")
    g = dev_time(lambda: _lib.check(L.mbrec_dmas_cf_dev(dx.data_ptr(), dz.data_ptr(), N * N, ds.data_ptr(), 1, D,
                                                        nT, C_, FS, 0, COV, RAD, out.data_ptr(), st())))
    c, want = host_time(lambda: O.dmas_cf(sig.astype(np.float64), X, Z, C_, FS, 0, COV, RAD))
    got = out.cpu().numpy().reshape(4, 1, N, N)
    entry("dmas_cf", g, c, f"{N}x{N} px, {D} x {nT} sinogram, 4 outputs",
          max(rel(got[k], want[k]) for k in range(4)))
    # bp_kernel (Nt = nT)
    t = np.arange(1, nT + 1) / FS
    K = torch.empty((nT, nT), dtype=torch.float64, device="cuda")
    g = dev_time(lambda: _lib.check(L.mbrec_bp_kernel_dev(t.ctypes.data, nT, K.data_ptr(), st())), reps=2)
    c, Ko = host_time(lambda: O.bp_kernel(t))
    entry("bp_kernel", g, c, f"Nt = {nT}", rel(K.cpu().numpy().T, Ko))
    # backproject_waveeq (C2 sinogram, K above)
    psf = np.outer(np.hanning(5), np.hanning(3))
    psf_f = np.asfortranarray(psf).ravel(order="F").copy()
    xv, zv = np.ascontiguousarray(X[:, 0]), np.ascontiguousarray(Z[0, :])
    ds64 = torch.tensor(sig.astype(np.float64)).cuda()
    ob = torch.empty((1, N, N), dtype=torch.float64, device="cuda")
    g = dev_time(lambda: _lib.check(L.mbrec_backproject_waveeq_dev(
        xv.ctypes.data, zv.ctypes.data, N, N, ds64.data_ptr(), 1, D, nT, C_, FS, 0, COV, RAD, psf_f.ctypes.data,
        psf.shape[0], psf.shape[1], K.data_ptr(), nT, ob.data_ptr(), st())))
    c, want = host_time(lambda: O.backproject_waveeq(sig.astype(np.float64), X, Z, C_, FS, 0, COV, RAD, psf, Ko))
    entry("backproject_waveeq", g, c, f"{N}x{N} px, {D} x {nT} sinogram, K {nT}^2", rel(ob.cpu().numpy(), want))
    # detector upsampling 256 -> 512, 32 sinograms
    nb = 32
    U = spline_upsampling_matrix(D, 2)
    Ud = torch.tensor(np.asfortranarray(U).ravel(order="F").copy()).cuda()
    s32 = np.random.default_rng(12).standard_normal((nb, D, nT))
    sd = torch.tensor(s32).cuda()
    ou = torch.empty((nb, 2 * D, nT), dtype=torch.float64, device="cuda")
    g = dev_time(lambda: _lib.check(L.mbrec_upsample_detectors_dev(Ud.data_ptr(), 2 * D, D, sd.data_ptr(), nT, nb,
                                                                    ou.data_ptr(), st())))
    Uo = O.not_a_knot_spline_matrix(D, 2)
    c, want = host_time(lambda: np.einsum("ud,bdt->but", Uo, s32))
    entry("upsample_detectors", g, c, f"{nb} sinograms {D} -> {2 * D} detectors x {nT}", rel(ou.cpu().numpy(), want))
    # TV and L1 solves: 400^2 DAS model, one sinogram; the oracle on a bounded number of
    # outer iterations, scaled to the device's
    half = (N - 1) / 2 * h
    disc = Discretization([-half, half], N)
    model = MSOTmodel(Probe("arc256_r40_174", True, False), disc, 1500.0, False, True, 0, 1e5, 12e6, None,
                      sigma_iterations=40)
    kz, kx = O.full_kernels(N, N, disc.dz, disc.dx, disc.pixelSigma)
    orc = O.Model(N, N, model.Data.absoluteTravelTimeIndices.reshape(D, -1), nT,
                  model.Data.circularInterpolationRadiusInSamples, sigma=model.normalizationFactor, kz=kz, kx=kx,
                  ir=model.Data.IR)
    ph = disks(disc, n=20, seed=3)[None]
    y = add_noise(orc.forward(ph), 0.01, seed=4)
    it_gpu, it_cpu = 10, 2
    tau = 1e-3
    eta = 0.05 * float(np.abs(orc.transpose(y)).max())
    rec_nn_with_TV_reg(model, y, 1, tau)             # first calls: workspace and library set-up
    rec_nn_with_L1_eye_reg(model, y, 1, eta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec_nn_with_TV_reg(model, y, it_gpu, tau)
    torch.cuda.synchronize()
    g = (time.perf_counter() - t0) * 1e3 / it_gpu
    c, _ = host_time(lambda: O.rec_nn_with_TV_reg(orc, y, it_cpu, tau))
    entry("rec_nn_with_TV_reg", g, c / it_cpu, f"{N}x{N} DAS, 1 sinogram, per outer iteration "
          f"(device {it_gpu}, oracle {it_cpu} iterations, wall incl. host calls)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec_nn_with_L1_eye_reg(model, y, it_gpu, eta)
    torch.cuda.synchronize()
    g = (time.perf_counter() - t0) * 1e3 / it_gpu
    c, _ = host_time(lambda: orc.rec_nn_with_L1_eye_reg(y, it_cpu, eta))
    entry("rec_nn_with_L1_eye_reg", g, c / it_cpu, f"{N}x{N} DAS, 1 sinogram, per outer iteration "
          f"(device {it_gpu}, oracle {it_cpu} iterations, wall incl. host calls)")
    print(json.dumps({"secondary": res}), flush=True)


def tv600(a):
    """The TV configuration: 600^2 grid at 60 um, SIR on, 512 detectors (the 256-detector
    arc upsampled by the not-a-knot spline across detectors), a batch of 32 seeded vessel
    trees, 1 % noise; rec_nn_with_TV_reg (inner SpaRSA 2 steps, TV prox by split Bregman)
    for 5 outer iterations after a 1-iteration warm-up.  The 512-detector data are the
    model's sinograms at the even detectors (256 of them), upsampled back to 512."""
    import numpy as np
    import torch

    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.phantoms import vessel_tree
    from mbrec.preprocess import upsample_detectors_in_sinogram
    from mbrec.probe import Probe
    from mbrec.solvers import rec_nn_with_TV_reg

    N, h, B, iters = 600, 60e-6, 32, 5
    half = (N - 1) / 2 * h
    disc = Discretization([-half, half], N)
    t0 = time.time()
    model = MSOTmodel(Probe("arc256_r40_174", True, False, num_detectors=512), disc, 1500.0, True, True, 0, 1e5,
                      12e6, None, sigma_iterations=20, progress=log)
    D, nT = model.shape_sinogram
    t_setup = time.time() - t0
    tune = model.plan.autotune(B, 3)
    log(f"tv600: set-up {t_setup:.1f} s ({D} detectors, {nT} samples, tables {model.plan.device_bytes / 1e9:.2f} GB), "
        f"launch choice {tune}")
    ph = np.stack([vessel_tree(disc, seed=100 + b) for b in range(B)]).astype(np.float32)
    y = model.applyForward(torch.tensor(ph).cuda())
    y = y + 0.01 * y.abs().amax(dim=(1, 2), keepdim=True) * torch.randn(y.shape, device="cuda")
    y256 = y[:, 0::2, :].double().cpu().numpy()
    torch.cuda.synchronize()
    tu = time.perf_counter()
    y512 = upsample_detectors_in_sinogram(y256, 2)
    t_up = time.perf_counter() - tu
    yd = torch.tensor(y512, dtype=torch.float32).cuda()
    tau = 1e-3
    rec_nn_with_TV_reg(model, yd, 1, tau, coupled=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rec, e_img, e_reg, *_ = rec_nn_with_TV_reg(model, yd, iters, tau, coupled=False)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t1
    out = {"tv600": {"grid": N, "pixel_um": 60, "detectors": D, "samples": nT, "sir": True, "batch": B,
                     "outer_iterations": iters, "s_per_solve_batch": round(t_solve, 3),
                     "ms_per_outer_iteration": round(t_solve / iters * 1e3, 1),
                     "recon_iterations_per_s": round(B * iters / t_solve, 2),
                     "upsample_host_to_host_s": round(t_up, 3), "setup_s": round(t_setup, 1)}}
    log(f"tv600: {t_solve / iters * 1e3:.1f} ms per outer iteration for {B} sinograms")
    print(json.dumps(out), flush=True)


def mixed(a):
    """224 independent NNLS-L2 solves (50 outer iterations) in one graph launch: grids drawn
    from {200^2, 300^2, 400^2} at 100 um, each sinogram a window of U{1500..2030} samples
    of its grid's 2030-sample axis, per-sinogram lambdaL2 = 1e-2 ||A'y||, seeded phantoms
    scaled by a smooth per-wavelength spectrum, 1 % noise.  Also times each grid's group
    alone (one graph each, run one after the other) for comparison."""
    import numpy as np
    import torch

    from mbrec.batch import BatchSolver
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe
    from mbrec.warmstart import phantom_batch

    g = np.random.default_rng(2024)
    n_wl, n_fr = 28, 8
    n = n_wl * n_fr
    grids = g.choice([200, 300, 400], size=n)
    counts = g.integers(1500, 2031, size=n)
    firsts = np.array([g.integers(0, 2030 - c + 1) for c in counts])
    spectrum = 0.6 + 0.4 * np.sin(np.linspace(0, np.pi, n_wl)) ** 2
    h = 1e-4
    groups, tensors = [], []
    t0 = time.time()
    for N in (200, 300, 400):
        idx = np.nonzero(grids == N)[0]
        half = (N - 1) / 2 * h
        disc = Discretization([-half, half], N)
        model = MSOTmodel(Probe("arc256_r40_174", True, False), disc, 1500.0, False, True, 0, 1e5, 12e6, None,
                          sigma_iterations=40)
        model.plan.autotune(len(idx), 3)
        ph = phantom_batch(disc, len(idx), 7000 + N) * spectrum[idx // n_fr][:, None, None]
        x = torch.tensor(ph, dtype=torch.float32).cuda()
        y = model.applyForward(x)
        noise = torch.randn(y.shape, generator=torch.Generator(device="cuda").manual_seed(N), device="cuda")
        y = (y + 0.01 * y.abs().amax(dim=(1, 2), keepdim=True) * noise).contiguous()
        lam = (1e-2 * torch.linalg.vector_norm(model.applyTranspose(y).reshape(len(idx), -1).double(), dim=1)).cpu()
        groups.append((model, [(int(firsts[i]), int(counts[i])) for i in idx]))
        tensors.append((y, lam.numpy()))
    log(f"mixed batch set-up {time.time() - t0:.1f} s: " +
        ", ".join(f"{N}^2 x {int(np.sum(grids == N))}" for N in (200, 300, 400)))
    lam_all = np.concatenate([t[1] for t in tensors])

    def timed(bs, steps):
        bs.run(graph=True)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            bs.run(graph=True)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    bs = BatchSolver(groups, a.iters, lam_all)
    for k, (y, _) in enumerate(tensors):
        bs.data[k].copy_(y)
    ms = timed(bs, a.steps)
    bs.close()
    alone = []
    for (model, w), (y, lam) in zip(groups, tensors):
        b1 = BatchSolver([(model, w)], a.iters, lam)
        b1.data[0].copy_(y)
        alone.append(timed(b1, a.steps))
        b1.close()
    out = {"mixed_batch": {"sinograms": n, "grids": {f"{N}^2": int(np.sum(grids == N)) for N in (200, 300, 400)},
                           "crop_lengths": [int(counts.min()), int(counts.max())], "outer_iterations": a.iters,
                           "ms_per_batch": round(ms, 2), "recon_per_s": round(n / (ms / 1e3), 2),
                           "ms_groups_one_after_another": round(sum(alone), 2),
                           "ms_per_group_alone": [round(v, 2) for v in alone]}}
    log(f"mixed batch: {ms:.1f} ms per batch of {n} ({n / (ms / 1e3):.1f} recon/s); groups one after another: "
        f"{sum(alone):.1f} ms")
    print(json.dumps(out), flush=True)


def build_problem(a, use_sir, seed=1000):
    """The 400^2 / 256-channel model, B seeded phantoms, their noisy sinograms (device,
    resident) and the per-sinogram lambdaL2 = 1e-2 ||A'y||."""
    import numpy as np
    import torch

    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe
    from mbrec.warmstart import phantom_batch

    N, h = a.grid, 1e-4
    half = (N - 1) / 2 * h
    t0 = time.time()
    disc = Discretization([-half, half], N)
    model = MSOTmodel(Probe("arc256_r40_174", True, False), disc, 1500.0, use_sir, True, 0, 1e5, 12e6, None,
                      sigma_iterations=40)
    B = a.batch
    if a.launch:
        gv, fc = (int(v) for v in a.launch.split(","))
        model.plan.set_launch(B, gv, fc)
        tune = (gv, fc)
    else:
        tune = model.plan.autotune(B, 5)
    log(f"setup {time.time() - t0:.1f} s, sigma {model.normalizationFactor:.6g}, "
        f"device tables {model.plan.device_bytes / 1e6:.1f} MB, launch choice {tune}")
    ph = phantom_batch(disc, B, seed)
    x_true = torch.tensor(ph).cuda()
    y = model.applyForward(x_true)
    g = torch.Generator(device="cuda").manual_seed(7)
    noise = torch.randn(y.shape, generator=g, device="cuda")
    y = (y + 0.01 * y.abs().amax(dim=(1, 2), keepdim=True) * noise).contiguous()
    # lambdaL2 = 1e-2 ||A'y|| per sinogram: the configuration's choice (the reference's demo
    # uses a fixed lambdaL2, reconstruction_overview_script.m:49)
    lam = (1e-2 * torch.linalg.vector_norm(model.applyTranspose(y).reshape(B, -1).double(), dim=1)).cpu().numpy()
    return model, ph, x_true, y, lam, tune


def operator_roofline(model, B, use_sir, tune, reps=10):
    """Roofline object of the dominant operator kernel (by time per solve) for batches of B.
    Algorithmic bytes per application (FP32, one geometry): 4 P D travel times (+ 4 P D SIR
    indices + referenced SIR entries) + 4 B P image + 4 B D nT sinogram; algorithmic flops
    2 B per tap of every pair (about 2R taps per pair, x the SIR window with SIR).  The
    kernel's time is the average of reps launches, HIP events on the launch stream; the
    binding term is max(bytes / 8 TB/s, flops / 157.3 TFLOP/s)."""
    import numpy as np
    D, nT = model.shape_sinogram
    P = model.Discretization.numberOfPixels
    sir_bytes = 0
    taps = 2.0 * model.Data.circularInterpolationRadiusInSamples
    if use_sir:
        se = model.Data.SPMRTimestepIndexStartAndEnd
        ref = np.unique(model.Data.SPMRIndices) - 1
        sir_bytes = 4 * int(np.sum(np.maximum(se[ref, 1] - se[ref, 0] + 1, 0)))
        taps *= float(np.mean(np.maximum(se[ref, 1] - se[ref, 0] + 1, 0)))
    alg_bytes = 4 * P * D * (2 if use_sir else 1) + sir_bytes + 4 * B * P + 4 * B * D * nT
    alg_flops = 2.0 * B * P * D * taps
    stages = ("scatter", "bins_to_y", "fir", "z_to_bins", "gather")
    st = {k: model.plan.time_stage(k, B, reps) for k in stages}
    # launches per outer iteration: 3 forward and 2 transpose applications (LSQR's closing
    # transpose of its last inner iteration is dead and not launched)
    per_it = {"scatter": 3 * st["scatter"], "bins_to_y": 3 * st["bins_to_y"], "fir": 5 * st["fir"],
              "z_to_bins": 2 * st["z_to_bins"], "gather": 2 * st["gather"]}
    dom = "gather" if per_it["gather"] >= per_it["scatter"] else "scatter"
    t_dom = st[dom] * 1e-3
    t_bytes, t_flops = alg_bytes / (HBM_PEAK_GBS * 1e9), alg_flops / (FP32_PEAK_TFLOPS * 1e12)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    key = f"{'sir' if use_sir else 'das'}_{dom}_B{B}_N{model.Discretization.Nx}"
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(key, {}).get("hbm_bytes_per_launch")
    fwd_form = "band form" if use_sir and tune[1] == 1004 else "pull form"
    kname = {"gather": "transpose gather", "scatter": f"forward pairs ({fwd_form})"}[dom]
    if t_flops >= t_bytes:
        bound, achieved, peak, unit = "valu", alg_flops / t_dom / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s"
    else:
        bound, achieved, peak, unit = "hbm", alg_bytes / t_dom / 1e9, HBM_PEAK_GBS, "GB/s"
    roof = {"bound": bound, "achieved": round(achieved, 2), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": traffic,
            "kernel": f"{kname}, {'SIR' if use_sir else 'DAS'}, B={B}", "kernel_ms": round(st[dom], 4),
            "algorithmic_bytes": int(alg_bytes), "algorithmic_flops": alg_flops,
            "bytes_term": {"GB/s": round(alg_bytes / t_dom / 1e9, 1),
                           "frac": round(alg_bytes / t_dom / 1e9 / HBM_PEAK_GBS, 4)},
            "flops_term": {"TFLOP/s": round(alg_flops / t_dom / 1e12, 2),
                           "frac": round(alg_flops / t_dom / 1e12 / FP32_PEAK_TFLOPS, 4)},
            "stage_ms": {k: round(v, 4) for k, v in st.items()},
            "ms_per_outer_iteration_by_stage": {k: round(v, 3) for k, v in per_it.items()}}
    # the operator pair for one sinogram (travel-time stream per application)
    t1 = {k: model.plan.time_stage(k, 1, reps) for k in ("scatter", "gather")}
    b1 = 4 * P * D * (2 if use_sir else 1) + sir_bytes + 4 * P + 4 * D * nT
    roof["operator_pair_B1"] = {k: {"ms": round(v, 4), "GB/s": round(b1 / (v * 1e-3) / 1e9, 1),
                                    "frac": round(b1 / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                                for k, v in t1.items()}
    log("stages (ms, B=%d): %s; %s %s %.2f %s (%.1f %%); B=1 %s" % (
        B, " ".join(f"{k} {v:.3f}" for k, v in st.items()), dom, bound, achieved, unit, 100 * achieved / peak,
        " ".join(f"{k} {v:.3f} ms" for k, v in t1.items())))
    return roof


def cpu_solve_rate(model, ph, y, use_sir, reps, iters):
    """FP64 oracle (the MEX loops' order and OpenMP decomposition) on the host cores:
    reps forward and reps transpose applications of one sinogram, scaled to one solve of
    ``iters`` outer iterations at the reference's count of 3 A + 3 A' per iteration."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    D, nT = model.shape_sinogram
    disc = model.Discretization
    kz, kx = O.full_kernels(disc.Nz, disc.Nx, disc.dz, disc.dx, disc.pixelSigma)
    kw = dict(kz=kz, kx=kx, ir=model.Data.IR)
    if use_sir:
        kw.update(sidx=model.Data.SPMRIndices, sir=model.Data.SPMR, se=model.Data.SPMRTimestepIndexStartAndEnd)
    orc = O.Model(disc.Nz, disc.Nx, model.Data.absoluteTravelTimeIndices.reshape(D, -1), nT,
                  model.Data.circularInterpolationRadiusInSamples, sigma=model.normalizationFactor, **kw)
    x0 = ph[:1].astype(np.float64)
    y0 = y[:1].double().cpu().numpy()
    tc = time.perf_counter()
    for _ in range(reps):
        orc.forward(x0)
    t_f = (time.perf_counter() - tc) / reps
    tc = time.perf_counter()
    for _ in range(reps):
        orc.transpose(y0)
    t_a = (time.perf_counter() - tc) / reps
    t_solve = iters * 3 * (t_f + t_a)
    log(f"cpu baseline: A {t_f:.2f} s, A' {t_a:.2f} s -> {1 / t_solve:.3g} recon/s on {O.num_threads()} threads")
    return {"value": 1.0 / t_solve, "unit": "recon/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{reps} forward + {reps} transpose FP64 applications of the same {disc.Nx}x{disc.Nz} "
                      f"model on one sinogram (A: {t_f:.3f} s, A': {t_a:.3f} s), scaled to one solve of "
                      f"{iters} outer iterations at 3 A + 3 A' each"}


def events_ms(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    out = None
    for _ in range(steps):
        out = fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, out


def pair_b1(a, reps=50):
    """The one-sinogram operator pair at 400^2 / 256 x 2030 (DAS): the forward-pair stage
    for each one-sinogram forward choice (2, 3: record streaming; 5, 6, 7: tile form with 8,
    4, 2 detectors per workgroup, thread per row; 9, 10: its lean form; 11, 12, 13: the lean
    form over rows sorted by record count, 11 the default) and the gather for each tile-gather variant (44-54), HIP
    events on the launch stream, each result compared with the default choice's; then the
    model operators (prefilter, IR, 1/sigma included).  Algorithmic bytes per application:
    4 P D + 4 P + 4 D nT (166.5 MB)."""
    import torch
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe
    N, h = a.grid, 1e-4
    half = (N - 1) / 2 * h
    m = MSOTmodel(Probe("arc256_r40_174", True, False), Discretization([-half, half], N), 1500.0, False, True, 0,
                  1e5, 12e6, 1.0)
    D, nT = m.shape_sinogram
    alg = 4.0 * (m.Discretization.numberOfPixels * D + m.Discretization.numberOfPixels + D * nT)
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.rand((1, N, N), generator=g, dtype=torch.float32, device="cuda")
    sg = torch.randn((1, D, nT), generator=g, dtype=torch.float32, device="cuda")
    rel = lambda u, w: float(torch.linalg.vector_norm((u - w).double()) / torch.linalg.vector_norm(w.double()))
    out = {"forward": {}, "gather": {}}
    m.plan.set_launch(1, -1, 0)
    y_ref, x_ref = m.applyForward(x), m.applyTranspose(sg)
    for ch in (2, 3, 5, 6, 7, 9, 10, 11, 12, 13):
        m.plan.set_launch(1, -1, ch)
        ms = m.plan.time_stage("scatter", 1, reps)
        out["forward"][ch] = {"us": round(ms * 1e3, 2), "alg_GB/s": round(alg / (ms * 1e-3) / 1e9, 1),
                              "rel_diff_vs_default": rel(m.applyForward(x), y_ref)}
        log(f"forward choice {ch}: {ms * 1e3:.1f} us")
    for gv in range(44, 55):
        m.plan.set_launch(1, gv, 0)
        ms = m.plan.time_stage("gather", 1, reps)
        out["gather"][gv] = {"us": round(ms * 1e3, 2), "alg_GB/s": round(alg / (ms * 1e-3) / 1e9, 1),
                             "rel_diff_vs_default": rel(m.applyTranspose(sg), x_ref)}
        log(f"gather variant {gv}: {ms * 1e3:.1f} us")
    m.plan.set_launch(1, -1, 0)
    for name, fn, arg in (("model_forward_us", m.applyForward, x), ("model_transpose_us", m.applyTranspose, sg)):
        ms, _ = events_ms(lambda: fn(arg), reps, 5)
        out[name] = round(ms * 1e3, 2)
    m.plan.set_fir_kernel(True)   # the IR filter on the FP64 matrix cores, for comparison
    for name, fn, arg in (("model_forward_fir_mfma_us", m.applyForward, x),
                          ("model_transpose_fir_mfma_us", m.applyTranspose, sg)):
        ms, _ = events_ms(lambda: fn(arg), reps, 5)
        out[name] = round(ms * 1e3, 2)
    m.plan.set_fir_kernel(False)
    out.update({"workload": f"{N}x{N} px, {D} ch x {nT} samples, DAS, one sinogram",
                "algorithmic_bytes": alg, "hbm_peak_GB/s": HBM_PEAK_GBS})
    if a.traffic_calibration:
        xc = torch.empty(CALIB_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        torch.cuda.synchronize()
        yc = xc.clone()
        torch.cuda.synchronize()
        del xc, yc
    print(json.dumps(out), flush=True)


def copy_bw(reps=20):
    """Practical HBM rate: a 1 GiB device copy (read + write traffic) timed with HIP events."""
    import torch
    n = 1 << 28
    x = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    ms, _ = events_ms(lambda: y.copy_(x), reps, 3)
    print(json.dumps({"copy_bytes": 4 * n, "ms": round(ms, 4), "read_plus_write_GB/s": round(8 * n / (ms * 1e-3) / 1e9, 1),
                      "hbm_peak_GB/s": HBM_PEAK_GBS}), flush=True)


def metric_bench(a):
    """The headline metric: recon/s of batch-64 NNLS-L2 at 400^2 x 256 ch reaching 1e-3
    relative L2 of the converged solution, learned warm start vs plain cold start."""
    import numpy as np
    import torch

    from mbrec.solvers import rec_nn_with_L2_reg_batch
    from mbrec.warmstart import (UNet, iterations_to_tolerance, nnls_l2_minimiser, rec_nn_with_L2_reg_learned,
                                 warm_start_estimate)

    weights = a.weights or os.path.join(ROOT, "mb-rec-msot_amd", "mbrec", "weights", f"unet_state_{a.grid}.npz")
    if not os.path.exists(weights):
        raise SystemExit(f"warm-start weights {weights} missing (python -m mbrec.train_warmstart --out ...)")
    model, ph, x_true, y, lam, tune = build_problem(a, False)
    B = a.batch
    D, nT = model.shape_sinogram
    tol = 1e-3
    t0 = time.time()
    ref, kkt = nnls_l2_minimiser(model, y, lam)
    log(f"converged solution: KKT residual {kkt:.2e} ({time.time() - t0:.1f} s)")
    nref = torch.linalg.vector_norm(ref.reshape(B, -1), dim=1)
    worst = lambda r: (torch.linalg.vector_norm((r.double() - ref).reshape(B, -1), dim=1) / nref).max().item()
    net = UNet.load(weights)
    # iteration counts (untimed; chunks of 25 that continue from the full loop state)
    k_plain, errs_plain = iterations_to_tolerance(model, y, lam, ref, tol, chunk=a.chunk)
    log(f"plain cold start: {k_plain} outer iterations to {tol:g}")
    w0, v0, used = warm_start_estimate(net, model, y, lam)
    k_learned, errs_learned = iterations_to_tolerance(model, y, lam, ref, tol, w0, v0, chunk=a.chunk)
    log(f"learned warm start: {k_learned} outer iterations to {tol:g} ({int(used.sum())}/{B} items from the network)")
    if k_plain is None or k_learned is None:
        raise SystemExit("a solve did not reach the tolerance")

    arbitrated = {}

    def learned_step():
        # network, fallback test, cold start as arbiter over the first 10 iterations (its
        # own 10 cold iterations are extra work inside the timed step), then the rest
        rec, used_j = rec_nn_with_L2_reg_learned(net, model, y, k_learned, lam, arbiter_iters=10)
        arbitrated["used"] = used_j
        return rec

    def plain_step():
        return rec_nn_with_L2_reg_batch(model, y, k_plain, lam)

    t_wall = time.perf_counter()
    ms_learned, rec_l = events_ms(learned_step, a.steps, a.warmup)
    wall = (time.perf_counter() - t_wall) * 1e3
    ms_plain, rec_p = events_ms(plain_step, 1, 0)
    # the estimate alone (A'y, network, fallback test), part of every learned step above
    ms_net, _ = events_ms(lambda: warm_start_estimate(net, model, y, lam), 3, 1)
    e_learned, e_plain = worst(rec_l), worst(rec_p)
    value = B / (ms_learned / 1e3)
    plain_value = B / (ms_plain / 1e3)
    log(f"learned {ms_learned:.1f} ms per batch ({value:.2f} recon/s, worst rel.err {e_learned:.2e}); plain "
        f"{ms_plain:.1f} ms ({plain_value:.2f} recon/s, {e_plain:.2e}); speed-up {value / plain_value:.2f}x")
    roof = operator_roofline(model, B, False, tune)
    cpu = None if a.no_cpu_baseline else cpu_solve_rate(model, ph, y, False, a.cpu_sample_reps, k_plain)
    out = {"metric": METRIC, "value": round(value, 4), "unit": "recon/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(ms_learned, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"batch {B} independent rec_nn_with_L2_reg solves (RegL2 = identity, "
                                  f"lambdaL2 = 1e-2 ||A'y||), {a.grid}x{a.grid} px at 100 um, no SIR, {D} ch x {nT} "
                                  f"samples, each to rel.err <= {tol:g} of the converged solution; learned "
                                  f"warm start (U-Net state estimate, cold start as fallback and arbiter) vs "
                                  f"plain cold start",
                      "batch": B, "grid": a.grid, "detectors": D, "samples": nT, "sir": False, "tol": tol},
           "learned": {"recon_per_s": round(value, 4), "outer_iterations": k_learned, "worst_rel_err": e_learned,
                       "items_from_network": int(used.sum()),
                       "items_kept_by_arbiter": int(arbitrated["used"].sum()), "arbiter_iterations": 10,
                       "ms_per_batch": round(ms_learned, 2),
                       "wall_ms_per_batch": round(wall / max(a.steps + a.warmup, 1), 2),
                       "estimate_ms": round(ms_net, 2)},
           "plain": {"recon_per_s": round(plain_value, 4), "outer_iterations": k_plain, "worst_rel_err": e_plain,
                     "ms_per_batch": round(ms_plain, 2)},
           "speedup_learned_vs_plain": round(value / plain_value, 3),
           "reference_solution": {"method": "converged NNLS-L2 minimiser (accelerated projected gradient)",
                                  "kkt_residual": kkt},
           "convergence": {"plain": {str(k): v for k, v in errs_plain.items() if k % 100 == 0 or k == k_plain},
                           "learned": {str(k): v for k, v in errs_learned.items()}},
           "weights": os.path.relpath(weights, ROOT),
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def fixed_bench(a):
    """Fixed-iteration plain solve (--fixed-iters N): one captured graph of N outer
    iterations for the batch, recon/s = B / time per batch (the SIR configuration uses it)."""
    import numpy as np
    import torch

    from mbrec.batch import BatchSolver

    use_sir = a.sir
    model, ph, x_true, y, lam, tune = build_problem(a, use_sir)
    B = a.batch
    D, nT = model.shape_sinogram
    bs = BatchSolver([(model, [(0, nT)] * B)], a.fixed_iters, lam)
    bs.data[0].copy_(y)
    torch.cuda.synchronize()
    ms_step, res = events_ms(lambda: bs.run(graph=True), a.steps, a.warmup)
    rec = res[0]
    value = B / (ms_step / 1e3)
    log(f"{a.steps} steps: {ms_step:.1f} ms/step, {value:.2f} recon/s")
    roof = operator_roofline(model, B, use_sir, tune)
    cpu = None if a.no_cpu_baseline else cpu_solve_rate(model, ph, y, use_sir, a.cpu_sample_reps, a.fixed_iters)
    if a.traffic_calibration:
        xc = torch.empty(CALIB_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        torch.cuda.synchronize()
        yc = xc.clone()
        torch.cuda.synchronize()
        del xc, yc
    out = {"metric": f"NNLS-L2 recons/sec, {a.fixed_iters} outer iterations (plain cold start)",
           "value": round(value, 4), "unit": "recon/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"rec_nn_with_L2_reg {a.fixed_iters} it, {a.grid}x{a.grid} px at 100 um, "
                                  f"{'SIR' if use_sir else 'no SIR'}, {D} ch x {nT} samples, batch {B} "
                                  "independent seeded sinograms (plain cold start)",
                      "batch": B, "grid": a.grid, "detectors": D, "samples": nT, "sir": use_sir,
                      "outer_iterations": a.fixed_iters, "lambdaL2_median": float(np.median(lam))},
           "rel_err_vs_phantom": round((torch.linalg.vector_norm((rec - x_true).double()) /
                                        torch.linalg.vector_norm(x_true.double())).item(), 5),
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def mex_bench(a):
    """Per-call wall time of the four MEX drop-ins at 400^2 / 256 x 2030 (one wavelength,
    host FP64 in and out as MATLAB passes them), the tables cached after the first call,
    next to the FP64 oracle's loops (the reference MEX loops' order and OpenMP
    decomposition) on the host cores.  Prints one JSON line."""
    import numpy as np
    import torch

    from mbrec import mex
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from mbrec.discretization import Discretization
    from mbrec.model import MSOTmodel
    from mbrec.probe import Probe

    N, h = a.grid, 1e-4
    half = (N - 1) / 2 * h
    model = MSOTmodel(Probe("arc256_r40_174", True, False), Discretization([-half, half], N), 1500.0, True, True,
                      0, 1e5, 12e6, 1.0)
    D, nT = model.shape_sinogram
    P = N * N
    R = float(model.Data.circularInterpolationRadiusInSamples)
    # MATLAB layouts as F-ordered views: [P x D] tables, [Ls x Nsir] SIR table, [2 x Nsir] windows
    tt = model.Data.absoluteTravelTimeIndices.reshape(D, P).T
    sidx = np.ascontiguousarray(model.Data.SPMRIndices.reshape(D, P), dtype=np.int32).T
    sir = np.ascontiguousarray(model.Data.SPMR, dtype=np.float64).T
    se = np.ascontiguousarray(model.Data.SPMRTimestepIndexStartAndEnd, dtype=np.int32).T
    g = np.random.default_rng(5)
    img = np.asfortranarray(g.uniform(0, 1, (1, P)))
    sig = np.asfortranarray(g.standard_normal((1, nT, D)))
    nt = np.int32(nT)
    calls = {
        "forwardDelayAndSumMex": lambda: mex.forwardDelayAndSumMex(img, tt, nt, R),
        "transposeDelayAndSumMex": lambda: mex.transposeDelayAndSumMex(sig, tt, R),
        "forwardMex": lambda: mex.forwardMex(img, sidx, tt, sir, nt, se, R),
        "transposeMex": lambda: mex.transposeMex(sig, sidx, tt, sir, se, R),
    }
    orc_calls = {
        "forwardDelayAndSumMex": lambda: O.forward_das(img, tt, nT, R),
        "transposeDelayAndSumMex": lambda: O.transpose_das(sig, tt, R),
        "forwardMex": lambda: O.forward_sir(img, sidx, tt, sir, nT, se, R),
        "transposeMex": lambda: O.transpose_sir(sig, sidx, tt, sir, se, R),
    }
    res = {}
    for name, fn in calls.items():
        t0 = time.perf_counter()
        out = fn()                      # first call: plan and table upload
        first = (time.perf_counter() - t0) * 1e3
        ts = []
        for _ in range(a.mex_reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        want = orc_calls[name]()
        t_orc = (time.perf_counter() - t0) * 1e3
        err = float(np.linalg.norm((out - want).ravel()) / np.linalg.norm(want.ravel()))
        res[name] = {"first_call_ms": round(first, 1), "per_call_ms_median": round(float(np.median(ts)), 3),
                     "per_call_ms_min": round(float(np.min(ts)), 3), "oracle_ms": round(t_orc, 1),
                     "oracle_cores": O.num_threads(), "rel_l2_vs_oracle": err}
        log(f"{name}: first {first:.1f} ms, then {np.median(ts):.3f} ms per call (oracle {t_orc:.1f} ms), "
            f"rel.err {err:.2e}")
    print(json.dumps({"mex_drop_ins": {"grid": N, "detectors": D, "samples": nT, "wavelengths": 1,
                                       "reps": a.mex_reps, "calls": res}}), flush=True)


def main():
    a = parse()
    if a.traffic_from:
        traffic_from(*a.traffic_from, N=a.grid, B=a.batch)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(batches of independent solves would shard across GPUs with no exchange)", flush=True)
        sys.exit(2)
    if a.copy_bw:
        copy_bw()
    elif a.pair_b1:
        pair_b1(a)
    elif a.mex:
        mex_bench(a)
    elif a.mixed:
        mixed(a)
    elif a.tv600:
        tv600(a)
    elif a.secondary:
        secondary(a)
    elif a.fixed_iters or a.sir:
        a.fixed_iters = a.fixed_iters or 50
        fixed_bench(a)
    else:
        metric_bench(a)


if __name__ == "__main__":
    main()
