"""mbrec_model_transpose_noalloc_dev (no device allocation: its scratch is a file-scope
device array) against mbrec_model_transpose_dev, through the C-ABI: two calls in a row,
the second with a smaller input on another plan, each equal to the allocating entry
point's result (same kernels on the same plan: within 1e-6 relative, FP32), and a batch
too large for the scratch is refused with mbrec:scratchTooSmall."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_tt, rel_l2, rng

pytestmark = pytest.mark.gpu


def _plan(Nz, Nx, D, nT, R, seed):
    from mbrec.model import DevicePlan
    tt = random_tt(Nz * Nx, D, nT, R, seed=seed, spill=0.05)
    kz = np.array([0.02, -0.1, 1.2, -0.1, 0.02])
    ir = rng(seed + 1).standard_normal(33)
    return DevicePlan(Nz, Nx, D, nT, R, tt.T, ir=ir, prefilter=(kz, kz), sigma=2.0)


def test_noalloc_transpose_matches_allocating_entry_point(require_gpu):
    import torch
    from mbrec import _lib
    L = _lib.lib()
    fn = L.mbrec_model_transpose_noalloc_dev
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    cases = [(40, 36, 24, 400, 2.6666666666666665, 3, 81), (24, 20, 12, 200, 3.3333333333333335, 1, 82)]
    for Nz, Nx, D, nT, R, B, seed in cases:
        print(f"noalloc transpose Nz={Nz} Nx={Nx} D={D} nT={nT} B={B}", flush=True)
        plan = _plan(Nz, Nx, D, nT, R, seed)
        sg = torch.tensor(rng(seed + 2).standard_normal((B, D, nT)), dtype=torch.float32).cuda()
        want = torch.empty((B, Nz * Nx), dtype=torch.float32, device="cuda")
        got = torch.full_like(want, float("nan"))
        _lib.check(L.mbrec_model_transpose_dev(plan.h, sg.data_ptr(), want.data_ptr(), B, _lib.stream_handle()))
        rc = fn(plan.h, sg.data_ptr(), got.data_ptr(), B, _lib.stream_handle())
        assert rc == 0, L.mbrec_last_error().decode()
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(g))
        assert rel_l2(g, w) < 1e-6
        plan.close()


def test_noalloc_transpose_refuses_oversized_batch(require_gpu):
    import torch
    from mbrec import _lib
    L = _lib.lib()
    fn = L.mbrec_model_transpose_noalloc_dev
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    plan = _plan(40, 36, 24, 400, 2.6666666666666665, 83)
    B = 2048          # about 1.6 GB of scratch for this plan: beyond the 64 MiB array
    sg = torch.zeros((1,), dtype=torch.float32, device="cuda")
    rc = fn(plan.h, sg.data_ptr(), sg.data_ptr(), B, _lib.stream_handle())
    assert rc != 0 and L.mbrec_last_error_id().decode() == "mbrec:scratchTooSmall"
    plan.close()
