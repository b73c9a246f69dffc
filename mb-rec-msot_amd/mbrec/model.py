"""MSOTmodel: geometry tables and the normalised forward / transpose operators
(reconstruction_model_based/models/MSOTmodel.m).

Setup (travel times, interpolation radius, impulse response, SIR tables,
prefilter) runs on the host once; the tables are then uploaded into a device
plan of libmbrec and ``Funcs.applyForward`` / ``Funcs.applyTranspose`` run the
HIP operators (MSOTmodel.m:285-286).  Arrays use MATLAB memory order with
reversed axes: images (W, Nx, Nz), sinograms (W, D, nT), travel times (D, Nx, Nz).
"""
from __future__ import annotations

import ctypes as C
import time
from types import SimpleNamespace

import numpy as np

from . import _lib
from .discretization import Discretization
from .irf import impulse_response_with_sir, impulse_response_without_sir
from .probe import Probe
from .sir import sir_tables_for_model

MEMBRANE_SAMPLES = 8193   # samples of the membrane curve over [-R, R] handed to the device


def acoustic_data(disc: Discretization, probe: Probe, c: float, single_sos: bool = True):
    """calculate_and_save_acoustic_data_for_model_with_SPMR_or_singSoS (MSOTmodel.m:436-461)
    on the device (calculate_travel_time.m, refraction.m:10-24, fermat_principle.m:12-32):
    travel times t[d, x] in seconds -- the straight-path time |p_x - q_d| / c for a single
    speed of sound; for the dual speed of sound (c below the membrane,
    probe.coupling.speedOfSound above it) the Fermat minimum over the membrane crossing for
    pixels below the membrane -- and the crossing x (NaN where the path is straight).
    Returns (t, x_intersections) as (D, Nx, Nz) arrays."""
    import torch
    L = _lib.lib()
    _lib.check(L.mbrec_init(0))
    X = np.ascontiguousarray(disc.region["xPositionsFov"].T, dtype=np.float64)   # (Nx, Nz): x = iz + Nz*ix
    Z = np.ascontiguousarray(disc.region["zPositionsFov"].T, dtype=np.float64)
    xd = np.ascontiguousarray(probe.detector.xPositionsOfTransducers, dtype=np.float64)
    zd = np.ascontiguousarray(probe.detector.zPositionsOfTransducers, dtype=np.float64)
    D, P = xd.size, X.size
    Xd, Zd = torch.as_tensor(X.ravel()).cuda(), torch.as_tensor(Z.ravel()).cuda()
    t = torch.empty((D, P), dtype=torch.float64, device="cuda")
    xi = torch.empty_like(t)
    R = float(probe.detector.radius)
    if single_sos:
        mz, nm, x0, dxm, c_above, c_below = None, 0, 0.0, 1.0, float(c), float(c)
    else:
        if probe.membrane is None or probe.coupling.speedOfSound is None:
            raise ValueError("Dual speed of sound requires membrane model.")
        xs = np.linspace(-R, R, MEMBRANE_SAMPLES)
        mz = np.ascontiguousarray(probe.membrane(xs), dtype=np.float64)
        nm, x0, dxm = MEMBRANE_SAMPLES, -R, xs[1] - xs[0]
        c_above, c_below = float(probe.coupling.speedOfSound), float(c)
    _lib.check(L.mbrec_travel_times_dev(Xd.data_ptr(), Zd.data_ptr(), P, xd.ctypes.data, zd.ctypes.data, D, c_below,
                                        c_above, R, mz.ctypes.data if mz is not None else None, nm, x0, dxm, 1.0,
                                        0.0, 0.0, t.data_ptr(), xi.data_ptr(), _lib.stream_handle()))
    return t.cpu().numpy().reshape(D, *X.shape), xi.cpu().numpy().reshape(D, *X.shape)


def travel_time_indices(disc: Discretization, probe: Probe, c: float, ncrop: int, single_sos: bool = True,
                        cache_dir=None):
    """MSOTmodel.m:350-382: tau[d, x] = t * fs - delayBeforeRecording - ncrop, shape
    (D, Nx, Nz), t from ``acoustic_data`` or from the acoustic cache (``cache_dir``)."""
    t, _ = load_or_compute_acoustic_data(disc, probe, c, single_sos, cache_dir)
    return t * probe.DAC.frequency - probe.DAC.delayBeforeRecording - ncrop


def _acoustic_hash_fields(disc, probe, c, single_sos):
    """hash_to_load_and_save_acoustic_data (MSOTmodel.m:289-308): coupling speed of sound
    (the medium's for a single speed of sound), detector positions, medium, region,
    useSingleSos and the membrane evaluated along the region's first row (+inf for a
    single speed of sound)."""
    xrow = np.asarray(disc.region["xPositionsFov"], dtype=np.float64)[0, :]
    if single_sos:
        coupling, mem = float(c), np.full(xrow.shape, np.inf)
    else:
        coupling, mem = float(probe.coupling.speedOfSound), np.asarray(probe.membrane(xrow), dtype=np.float64)
    return {"coupling_speedOfSound": np.float64(coupling),
            "detector_x": np.asarray(probe.detector.xPositionsOfTransducers, dtype=np.float64),
            "detector_y": np.asarray(probe.detector.zPositionsOfTransducers, dtype=np.float64),
            "Medium_speedOfSound": np.float64(c),
            "region_xPositionsFov": np.asarray(disc.region["xPositionsFov"], dtype=np.float64),
            "region_zPositionsFov": np.asarray(disc.region["zPositionsFov"], dtype=np.float64),
            "useSingleSos": np.bool_(single_sos), "membrane": mem}


def acoustic_hash(fields) -> str:
    """MD5 over the fields' names, dtypes, shapes and bytes in name order (the reference uses
    DataHash of the same struct; the byte stream differs, so files do not interchange)."""
    import hashlib
    h = hashlib.md5()
    for k in sorted(fields):
        a = np.ascontiguousarray(fields[k])
        h.update(k.encode())
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def load_or_compute_acoustic_data(disc, probe, c, single_sos=True, cache_dir=None):
    """calculate_travel_time_indices' cache (MSOTmodel.m:351-377): with cache_dir, the
    travel times and membrane crossings are read from acoustic_file_<hash>.npz when it
    exists -- after checking that its stored hashed fields equal the current ones ('Hash
    collision: Delete outdated travel time files') -- and written there otherwise."""
    import os
    if cache_dir is None:
        return acoustic_data(disc, probe, c, single_sos)
    fields = _acoustic_hash_fields(disc, probe, c, single_sos)
    path = os.path.join(cache_dir, f"acoustic_file_{acoustic_hash(fields)}.npz")
    if os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        for k, v in fields.items():
            if not np.array_equal(z["hashed_" + k], v):
                raise AssertionError("Hash collision: Delete outdated travel time files")
        return z["travel_times"], z["x_membrane_intersections"]
    t, xi = acoustic_data(disc, probe, c, single_sos)
    os.makedirs(cache_dir, exist_ok=True)
    np.savez(path, travel_times=t, x_membrane_intersections=xi, **{"hashed_" + k: v for k, v in fields.items()})
    return t, xi


class DevicePlan:
    """Owns one libmbrec plan (device-resident tables)."""

    def __init__(self, Nz, Nx, D, nT, R, tt, sir=None, ir=None, prefilter=None, sigma=1.0):
        L = _lib.lib()
        _lib.check(L.mbrec_init(0))
        h = C.c_void_p()
        tt = np.ascontiguousarray(tt, dtype=np.float64)
        _lib.check(L.mbrec_plan_create(C.byref(h), int(Nz), int(Nx), int(D), int(nT), float(R),
                                       tt.ctypes.data))
        self.h = h
        self.Nz, self.Nx, self.D, self.nT, self.R = Nz, Nx, D, nT, R
        self.P = Nz * Nx
        if sir is not None:
            self.set_sir(*sir)
        if ir is not None:
            ir = np.ascontiguousarray(ir, dtype=np.float64)
            per_det = 1 if ir.ndim == 2 else 0
            Lir = ir.shape[-1]
            _lib.check(L.mbrec_plan_set_ir(h, ir.ctypes.data, int(Lir), per_det))
        if prefilter is not None:
            kz, kx = (np.ascontiguousarray(k, dtype=np.float64) for k in prefilter)
            _lib.check(L.mbrec_plan_set_prefilter(h, kz.ctypes.data, (len(kz) - 1) // 2,
                                                  kx.ctypes.data, (len(kx) - 1) // 2))
        self.set_sigma(sigma)

    def set_sir(self, sidx, table, se):
        """SIR tables (D, P) int32 1-based indices, (Nsir, Ls) table, (Nsir, 2) windows; a
        second call replaces the first (workspace and launch choices are rebuilt)."""
        sidx = np.ascontiguousarray(sidx, dtype=np.int32)
        table = np.ascontiguousarray(table, dtype=np.float64)
        se = np.ascontiguousarray(se, dtype=np.int32)
        _lib.check(_lib.lib().mbrec_plan_set_sir(self.h, sidx.ctypes.data, table.ctypes.data,
                                                 int(table.shape[1]), int(table.shape[0]), se.ctypes.data))

    def set_sigma(self, sigma):
        self.sigma = float(sigma)
        _lib.check(_lib.lib().mbrec_plan_set_sigma(self.h, self.sigma))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().mbrec_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return int(_lib.lib().mbrec_plan_device_bytes(self.h))

    @property
    def tau_frac_bits(self):
        return int(_lib.lib().mbrec_plan_tau_frac_bits(self.h))

    def autotune(self, B=1, reps=10):
        """Pick the fastest gather / scatter launch variants for a batch of B
        (project autotuner in libmbrec); returns (gather variant, scatter chunk)."""
        ch = (C.c_int32 * 2)()
        _lib.check(_lib.lib().mbrec_plan_autotune(self.h, int(B), int(reps), _lib.stream_handle(), ch))
        return int(ch[0]), int(ch[1])

    def set_launch(self, B, gather=-1, scatter=0):
        """Force a launch variant for batches of B (see mbrec_plan_set_launch)."""
        _lib.check(_lib.lib().mbrec_plan_set_launch(self.h, int(B), int(gather), int(scatter)))

    def set_fir_kernel(self, mfma):
        """IR filter kernel: FP64 matrix-core Toeplitz products (True) or the vector
        kernel (False, default); both sum in FP64."""
        _lib.check(_lib.lib().mbrec_plan_set_fir_kernel(self.h, 1 if mfma else 0))

    STAGES = {"scatter": 0, "bins_to_y": 1, "fir": 2, "z_to_bins": 3, "gather": 4}

    def time_stage(self, stage, B=1, reps=20):
        """Average device time (ms) of one operator stage, HIP events on the current stream."""
        ms = C.c_float()
        _lib.check(_lib.lib().mbrec_plan_time_stage(self.h, self.STAGES[stage], int(B), int(reps),
                                                    _lib.stream_handle(), C.byref(ms)))
        return float(ms.value)

    # device-tensor operators --------------------------------------------------
    def _run(self, fn, src, dst, B):
        _lib.check(fn(self.h, src.data_ptr(), dst.data_ptr(), int(B), _lib.stream_handle()))

    def forward(self, img, out=None, model=True):
        import torch
        B = img.numel() // self.P
        if out is None:
            out = torch.empty((B, self.D, self.nT), dtype=torch.float32, device=img.device)
        L = _lib.lib()
        self._run(L.mbrec_model_forward_dev if model else L.mbrec_mex_forward_dev,
                  img.contiguous(), out, B)
        return out

    def transpose(self, sino, out=None, model=True):
        import torch
        B = sino.numel() // (self.D * self.nT)
        if out is None:
            out = torch.empty((B, self.Nx, self.Nz), dtype=torch.float32, device=sino.device)
        L = _lib.lib()
        self._run(L.mbrec_model_transpose_dev if model else L.mbrec_mex_transpose_dev,
                  sino.contiguous(), out, B)
        return out


class MSOTmodel:
    """MSOTmodel(probe, discretization, speed_of_sound, include_sir, use_single_sos,
    num_cropped_samples, filt_cutoff_min, filt_cutoff_max, model_normalization_factor)
    (MSOTmodel.m:18-48)."""

    def __init__(self, probe: Probe, discretization: Discretization, speed_of_sound_below_membrane: float,
                 include_sir: bool, use_single_speed_of_sound: bool = True,
                 num_cropped_samples_at_sinogram_start: int = 0, filt_cutoff_min: float = 1e5,
                 filt_cutoff_max: float = 12e6, model_normalization_factor=None,
                 sigma_iterations: int = 50, progress=None, acoustic_cache_dir=None, spmr_dir=None):
        """progress: optional callable receiving a line after each set-up stage;
        acoustic_cache_dir: directory of the acoustic_file_<hash>.npz travel-time cache
        (MSOTmodel.m:351-377), None for no cache; spmr_dir: directory of the polarSPMR files
        (MSOTmodel.m:482-571), None to generate the SIR table every time."""
        say = progress or (lambda msg: None)
        t_start = time.time()
        self.Probe = probe
        self.Discretization = discretization
        self.Medium = SimpleNamespace(speedOfSound=float(speed_of_sound_below_membrane))
        self.includeSir = bool(include_sir)
        self.useSingleSpeedOfSound = bool(use_single_speed_of_sound)
        single = self.useSingleSpeedOfSound
        ncrop = int(num_cropped_samples_at_sinogram_start)
        self.DataPreprocessing = SimpleNamespace(numCroppedSamplesAtSinogramStart=ncrop,
                                                 filtCutoffMin=filt_cutoff_min,
                                                 filtCutoffMax=filt_cutoff_max)
        c = self.Medium.speedOfSound
        fs = probe.DAC.frequency
        nT = int(probe.DAC.numRecordedSamplesPerTransducer - ncrop)
        t_sec, x_mem = load_or_compute_acoustic_data(discretization, probe, c, single, acoustic_cache_dir)
        tt = t_sec * fs - probe.DAC.delayBeforeRecording - ncrop     # MSOTmodel.m:380-381
        Z = discretization.region["zPositionsFov"]
        R = (Z[1, 0] - Z[0, 0]) / (c / fs)                      # MSOTmodel.m:36-38
        self.Data = SimpleNamespace(numTimeSamplesPerTransducer=nT, absoluteTravelTimeIndices=tt,
                                    circularInterpolationRadiusInSamples=R)
        say(f"model: travel times ({time.time() - t_start:.1f} s)")
        if self.includeSir:
            ir = impulse_response_with_sir(probe.detector.EIR, fs, filt_cutoff_min, filt_cutoff_max)
            X = discretization.region["xPositionsFov"].T.ravel()
            Zf = discretization.region["zPositionsFov"].T.ravel()
            res = min(discretization.dx, discretization.dz)
            c_spmr = c
            if not single:
                # virtual pixel positions along the detector-side ray (MSOTmodel.m:582-593)
                D_ = probe.detector.numOfTransducers
                c_spmr = probe.coupling.speedOfSound
                xq = probe.detector.xPositionsOfTransducers[:, None]
                zq = probe.detector.zPositionsOfTransducers[:, None]
                xm = x_mem.reshape(D_, -1)
                zm = probe.membrane(np.nan_to_num(xm))
                nrm = np.sqrt((xm - xq) ** 2 + (zm - zq) ** 2)
                run = t_sec.reshape(D_, -1) * c_spmr
                X = np.where(np.isnan(xm), X[None, :], xq + (xm - xq) / nrm * run)
                Zf = np.where(np.isnan(xm), Zf[None, :], zq + (zm - zq) / nrm * run)
            table, se, sidx, info = sir_tables_for_model(
                X, Zf, probe.detector.xPositionsOfTransducers, probe.detector.zPositionsOfTransducers,
                res, c_spmr, fs, probe.detector.elementWidth, probe.detector.numSubElements,
                discretization.pixelModel, spmr_dir=spmr_dir, probe_id=probe.probeId)
            self.Data.SPMR = table
            self.Data.SPMRTimestepIndexStartAndEnd = se
            self.Data.SPMRIndices = sidx
            self.Data.SPMRGrid = info
            sir = (sidx, table, se)
            say(f"model: SIR tables {table.shape} ({time.time() - t_start:.1f} s)")
        else:
            ir = impulse_response_without_sir(probe.detector.EIR, fs, c, discretization.pixelModel,
                                              filt_cutoff_min, filt_cutoff_max)
            sir = None
        self.Data.IR = ir
        D = probe.detector.numOfTransducers
        self.plan = DevicePlan(discretization.Nz, discretization.Nx, D, nT, R,
                               tt.reshape(D, -1), sir=sir, ir=ir,
                               prefilter=(discretization.kz, discretization.kx), sigma=1.0)
        say(f"model: device plan ({time.time() - t_start:.1f} s)")
        if model_normalization_factor is None:
            model_normalization_factor = self.estimate_sigma(sigma_iterations)
            say(f"model: sigma_max ({time.time() - t_start:.1f} s)")
        self.normalizationFactor = float(model_normalization_factor)
        self.plan.set_sigma(self.normalizationFactor)
        self.Funcs = SimpleNamespace(applyForward=self.applyForward, applyTranspose=self.applyTranspose)

    # sizes
    @property
    def shape_image(self):
        return (self.Discretization.Nx, self.Discretization.Nz)

    @property
    def shape_sinogram(self):
        return (self.Probe.detector.numOfTransducers, self.Data.numTimeSamplesPerTransducer)

    def estimate_sigma(self, iters: int = 50, seed: int = 0, tol: float = 1e-10) -> float:
        """Largest singular value of the unnormalised operator (replaces svds(A, 1,
        'largest', MaxIterations 50, Tolerance 1e-8) at MSOTmodel.m:278) by Golub-Kahan-
        Lanczos bidiagonalisation with full reorthogonalisation: after j steps
        A V_j = U_j B_j with B_j upper bidiagonal, and the largest singular value of the
        small B_j converges to sigma_max from below, fast even when the second singular
        value is close (power iteration does not).  Stops when it changes by <= tol
        relative, or after ``iters`` steps.  Vectors and scalars FP64, operators FP32."""
        import torch
        self.plan.set_sigma(1.0)
        g = torch.Generator(device="cpu").manual_seed(seed)
        v = torch.rand((1, *self.shape_image), generator=g, dtype=torch.float64).cuda()
        v /= torch.linalg.vector_norm(v)
        fwd = lambda x: self.plan.forward(x.float().reshape(1, -1)).double()
        adj = lambda y: self.plan.transpose(y.float().reshape(1, -1)).double().reshape(v.shape)
        V, U, al, be = [v], [], [], []
        u = fwd(v)
        al.append(torch.linalg.vector_norm(u).item())
        u = u / al[-1]
        U.append(u)
        sig = al[0]
        for _ in range(iters):
            r = adj(U[-1]) - al[-1] * V[-1]
            for q in V:
                r -= (r * q).sum() * q
            be.append(torch.linalg.vector_norm(r).item())
            if be[-1] <= 1e-300:
                break
            V.append(r / be[-1])
            pv = fwd(V[-1]) - be[-1] * U[-1]
            for q in U:
                pv -= (pv * q).sum() * q
            al.append(torch.linalg.vector_norm(pv).item())
            if al[-1] <= 1e-300:
                break
            U.append(pv / al[-1])
            Bk = np.diag(al) + np.diag(be, 1)[: len(al), : len(al)]
            s_new = float(np.linalg.svd(Bk, compute_uv=False)[0])
            done = abs(s_new - sig) <= tol * s_new
            sig = s_new
            if done:
                break
        return sig

    # operators (MSOTmodel.m:285-286) --------------------------------------------
    def applyForward(self, p0):
        return self._apply(p0, forward=True)

    def applyTranspose(self, sig):
        return self._apply(sig, forward=False)

    def _apply(self, x, forward: bool):
        import torch
        host = isinstance(x, np.ndarray)
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32)).cuda() if host else x
        if t.dtype != torch.float32:
            raise TypeError("device tensors must be float32")
        t = t.contiguous()
        if forward:
            B = t.numel() // self.Discretization.numberOfPixels
            out = self.plan.forward(t.view(B, -1))
            out = out.view(B, *self.shape_sinogram)
        else:
            B = t.numel() // (self.shape_sinogram[0] * self.shape_sinogram[1])
            out = self.plan.transpose(t.view(B, -1))
            out = out.view(B, *self.shape_image)
        if host:
            out = out.cpu().numpy().astype(np.float64)
            if np.ndim(x) in (2,) or (np.ndim(x) == 1):
                out = out[0]
        return out


def define_model_for_reconstruction(field_of_view, number_of_grid_points_fov, device_probe_id,
                                    use_eir, use_indiv_eir, use_sir, use_single_speed_of_sound,
                                    speed_of_sound_tissue, num_cropped_samples_at_sinogram_start,
                                    filt_cutoff_min, filt_cutoff_max, model_normalization_factor=None,
                                    custom_pixel_model=None):
    """define_model_for_reconstruction.m:1-11."""
    disc = Discretization(field_of_view, number_of_grid_points_fov, custom_pixel_model)
    probe = Probe(device_probe_id, use_eir, use_indiv_eir)
    return MSOTmodel(probe, disc, speed_of_sound_tissue, use_sir, use_single_speed_of_sound,
                     num_cropped_samples_at_sinogram_start, filt_cutoff_min, filt_cutoff_max,
                     model_normalization_factor)
