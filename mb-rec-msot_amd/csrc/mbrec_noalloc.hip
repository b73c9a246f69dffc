// mbrec_noalloc.hip -- the model transpose without any device allocation.
//
// mbrec_model_transpose_noalloc_dev runs the same kernels as mbrec_model_transpose_dev
// (MSOTmodel.m:224-257, 286), but its scratch -- the IR-filtered sinograms, the bins, the
// two prefilter images and the detector-group partial images -- lives in a file-scope
// device array of kScratchBytes instead of the plan's workspace, so the call allocates
// nothing (for callers under graph capture or in loops that must not allocate).  Every
// call on a device shares that device's array and it keeps the last call's contents:
// calls that overlap on one device race.
#include "mbrec_internal.h"

#include <mutex>

namespace mbrec {

static const size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per device
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

static const int kMaxDevices = 64;
static float *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// device address of this device's scratch (hipGetSymbolAddress once per device, with
// the device current); the array's name is a host-side stand-in and is never passed on
static int scratch_ptr(float **out)
{
    int dev = 0;
    MBREC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(MBREC_ERR_ARG, "mbrec:device", "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        MBREC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(MBREC_ERR_HIP, "mbrec:scratch", "no device address for the scratch array");
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return MBREC_OK;
}

__global__ void k_fill_zero(float *__restrict__ a, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0.0f;
}

static size_t round256(size_t nfloat) { return (nfloat + 63) / 64 * 64; }

// floats of scratch a transpose of B sinograms on plan p needs (the layout of run_noalloc)
static size_t scratch_floats(const Plan *p, int B)
{
    const size_t Bp = (size_t)batch_stride(B);
    return round256((size_t)B * p->D * p->nT) + round256(((size_t)p->D * p->geom.NB + 2) * Bp) +
           round256((size_t)2 * B * p->P) + round256((size_t)std::max(8 * B, 16) * p->P);
}

}  // namespace mbrec

using namespace mbrec;

extern "C" int mbrec_model_transpose_noalloc_dev(mbrec_plan h, const float *sino, float *img, int B,
                                                 mbrec_stream s)
{
    if (!h || !img || !sino || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "invalid arguments");
    const Plan *p = &h->p;
    const size_t need = scratch_floats(p, B);
    if (need * sizeof(float) > kScratchBytes)
        return set_error(MBREC_ERR_ARG, "mbrec:scratchTooSmall",
                         "model transpose of %d sinograms on this plan needs %zu bytes of scratch, more than the "
                         "%zu available without allocation",
                         B, need * sizeof(float), kScratchBytes);
    float *base;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    // a plan copy whose workspace points into the scratch (tables and launch choices shared)
    Plan q = *p;
    const size_t Bp = (size_t)batch_stride(B);
    float *at = base;
    q.d_sino_tmp = at;
    at += round256((size_t)B * p->D * p->nT);
    q.d_bins = at;
    at += round256(((size_t)p->D * p->geom.NB + 2) * Bp);
    q.d_img_tmp = at;
    at += round256((size_t)2 * B * p->P);
    q.d_gpart = at;
    q.bcap = B;
    q.d_hpart = q.d_imgI = q.d_img_slot = q.d_sband = nullptr;   // forward-only buffers: not used here
    // the two zero rows after the last detector's bins (pairs outside the bin range read
    // them); the previous call may have left bins of another geometry there
    hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(256), 0, (hipStream_t)s, q.d_bins + (size_t)p->D * p->geom.NB * Bp,
                       2 * Bp);
    MBREC_HIP(hipGetLastError());
    return op_transpose_run(&q, sino, img, B, true, (hipStream_t)s);
}
