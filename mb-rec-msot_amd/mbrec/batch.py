"""Batched NNLS-L2 over several geometries in one graph launch (libmbrec).

No counterpart in the reference, which reconstructs one acquisition per
rec_nn_with_L2_reg.m call.  A batch mixes grids (one model per grid) and
sinograms of uneven length (each a window of its model's sample axis, i.e. a
sinogram cropped at its start and/or end); every sinogram is an independent
solve with its own lambdaL2.  The first ``run`` captures all solves into one
HIP graph (one branch per model, so models of different sizes overlap); later
runs replay it.  Buffers are allocated here once and reused by every run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class BatchSolver:
    """groups: [(model, windows)] with windows a list of (first_sample, count)
    per sinogram on that model's sample axis.  ``data[g]`` (B_g, D, nT_g) and,
    with_init=True, ``init[g]`` / ``init_dual[g]`` (B_g, Nx, Nz: warm start of the
    outer iterate and of the scaled dual) are device tensors to fill before
    ``run``; ``rec[g]`` receives the reconstructions."""

    def __init__(self, groups, num_iter, lambdaL2, with_init=False):
        import torch
        plans, Bs, wins = [], [], []
        self.models = [m for m, _ in groups]
        self.data, self.rec, self.init, self.init_dual = [], [], [], []
        for m, w in groups:
            w = np.asarray(w, dtype=np.int32).reshape(-1, 2)
            D, nT = m.shape_sinogram
            B = w.shape[0]
            plans.append(m.plan.h.value)
            Bs.append(B)
            wins.append(w)
            self.data.append(torch.zeros((B, D, nT), dtype=torch.float32, device="cuda"))
            self.rec.append(torch.zeros((B, *m.shape_image), dtype=torch.float32, device="cuda"))
            if with_init:
                self.init.append(torch.zeros((B, *m.shape_image), dtype=torch.float32, device="cuda"))
                self.init_dual.append(torch.zeros((B, *m.shape_image), dtype=torch.float32, device="cuda"))
        self.windows = np.ascontiguousarray(np.concatenate(wins), dtype=np.int32)
        n = self.windows.shape[0]
        lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lambdaL2, dtype=np.float64), (n,)))
        G = len(groups)
        P = (C.c_void_p * G)(*plans)
        BB = (C.c_int * G)(*Bs)
        Dp = (C.c_void_p * G)(*[t.data_ptr() for t in self.data])
        Rp = (C.c_void_p * G)(*[t.data_ptr() for t in self.rec])
        Ip = (C.c_void_p * G)(*[t.data_ptr() for t in self.init]) if with_init else None
        Vp = (C.c_void_p * G)(*[t.data_ptr() for t in self.init_dual]) if with_init else None
        h = C.c_void_p()
        torch.cuda.synchronize()
        _lib.check(_lib.lib().mbrec_batch_create(C.byref(h), G, P, BB, Dp, Rp, Ip, Vp, self.windows.ctypes.data,
                                                 lam.ctypes.data, int(num_iter)))
        self.h = h
        self.sizes = Bs
        self.num_iter = int(num_iter)

    def set_lambda(self, lambdaL2):
        lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lambdaL2, dtype=np.float64), (len(self.windows),)))
        _lib.check(_lib.lib().mbrec_batch_set_lambda(self.h, lam.ctypes.data))

    def run(self, use_init=False, graph=True):
        _lib.check(_lib.lib().mbrec_batch_run(self.h, int(bool(use_init)), int(bool(graph)), _lib.stream_handle()))
        return self.rec

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().mbrec_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
