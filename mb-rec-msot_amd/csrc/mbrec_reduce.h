// mbrec_reduce.h -- deterministic FP64 reductions shared by the device solvers.
//
// Every vector kernel of a solver runs on a fixed grid (kRedBlocks blocks per
// batch item) and writes one FP64 partial per block into part[q][b][blk];
// scalar kernels sum the partials of a group in a fixed order, so norms and
// inner products are reproducible from run to run.
#pragma once
#include "mbrec_internal.h"

namespace mbrec {

constexpr int kRedBlocks = 64;     // reduction blocks per batch item
constexpr int kMaxQ = 6;           // reduction slots

enum { Q_A = 0, Q_B = 1, Q_C = 2, Q_D = 3, Q_E = 4, Q_F = 5 };

// block-wide FP64 sum, written to part[q][b][blk]
__device__ __forceinline__ void block_sum_store(double v, double *part, int q, int b, int B)
{
    __shared__ double sh[8];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
        part[((size_t)q * B + b) * kRedBlocks + blockIdx.x] = t;
    }
    __syncthreads();
}

__device__ __forceinline__ double group_sum(const double *part, int q, int g, int B, int coupled)
{
    double t = 0.0;
    const int b0 = coupled ? 0 : g, b1 = coupled ? B : g + 1;
    for (int b = b0; b < b1; ++b)
        for (int i = 0; i < kRedBlocks; ++i) t += part[((size_t)q * B + b) * kRedBlocks + i];
    return t;
}


static inline dim3 vgrid(int B) { return dim3(kRedBlocks, B); }

}  // namespace mbrec
