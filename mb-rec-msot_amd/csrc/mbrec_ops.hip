// mbrec_ops.hip -- gfx950 kernels of the forward / transpose model operators.
//
// Pipeline (device layouts: images planar [B][P], sinograms planar [B][D][nT];
// internal bins batch-interleaved [D][NB][Bp], Bp = B rounded up to 4):
//
//   forward    img --prefilter (2 passes, 2nd writes [P][Bp])--> coeffs
//                  --k_scatter_bk (bucketed, LDS window atomics)--> windows [D][nbk][WL][Bp]
//                  --k_bins_to_y (window sums + polyphase hat filter)--> y [B][D][nT]
//                  --k_fir (IR convolution, 1/sigma)--> sino
//   transpose  sino --k_fir (IR correlation)--> z --k_z_to_bins (polyphase hat filter)--> bins
//                  --k_gather (one interpolation per pair, Kahan, 1/sigma)--> coeffs
//                  --prefilter^T--> img
//
// The bin formulation is explained in mbrec_internal.h; it is exact, so the
// composite equals MSOTmodel.m:189-257 (A = crop(IR * MEX(F p)) / sigma) up to
// FP32 rounding.  Sums that cancel (the band-pass IR filter, the prefilter)
// accumulate in FP64; the transpose pair sum uses Kahan compensation.
#include "mbrec_internal.h"

#include <algorithm>

namespace mbrec {

// ---------------------------------------------------------------------------
// fixed-point travel time -> (sample n + K, phase m, weight u)
// ---------------------------------------------------------------------------
struct Decoded {
    int nn;     // n + K  (valid iff 0 <= nn < nT + 2K)
    int m;      // phase 0..2
    float u;    // position inside the segment, [0, 1)
};

__device__ __forceinline__ Decoded decode_tau(int32_t q, const BinGeom &g)
{
    Decoded r;
    const int base = q >> g.F;   // arithmetic shift = floor
    const float fr = (float)(q & ((1 << g.F) - 1)) * __builtin_ldexpf(1.0f, -g.F);
    const int m = (fr >= g.phi1) + (fr >= g.phi2);
    // the three segment positions, then a select of values: selecting among the
    // parameters themselves lets the compiler turn the choice into a memory load
    // from the kernel-argument segment, a dependent load on every pair
    const float u0 = fr * g.iw0, u1 = (fr - g.phi1) * g.iw1, u2 = (fr - g.phi2) * g.iw2;
    r.nn = base + g.K;
    r.m = m;
    r.u = (m == 0) ? u0 : ((m == 1) ? u1 : u2);
    return r;
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Word 3 of a raw buffer resource on CDNA (32-bit data format, no swizzle);
// reads past num_records return 0.
constexpr int kBufferWord3 = 0x00020000;

// s_waitcnt immediate that waits until at most n vector-memory operations of
// the wave are outstanding (gfx9 encoding: vmcnt in bits 3:0 and 15:14,
// expcnt 6:4 and lgkmcnt 11:8 left at their maxima, i.e. not waited for)
__host__ __device__ constexpr int kWaitVm(int n) { return (n & 15) | (((n >> 4) & 3) << 14) | 0x70 | 0xF00; }

// Staging of n elements into LDS with U global loads in flight per thread
// before any is stored (a plain strided loop waits out one load latency per
// element).  ld(e) returns element e, st(e, v) stores it; NT threads take part.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int U, int NT, class Ld, class St>
__device__ __forceinline__ void stage_batched(int t, int n, Ld &&ld, St &&st)
{
    for (int e0 = t; e0 < n; e0 += NT * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + NT * u;
            v[u] = e < n ? ld(e) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + NT * u;
            if (e < n) st(e, v[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// planar [B][P] -> interleaved [P][Bp] (padding columns zeroed)
// ---------------------------------------------------------------------------
__global__ void k_to_interleaved(const float *__restrict__ in, float *__restrict__ out, int P, int B, int Bp)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)P * Bp) return;
    const int px = (int)(idx / Bp), b = (int)(idx - (size_t)px * Bp);
    out[idx] = b < B ? in[(size_t)b * P + px] : 0.0f;
}

// ---------------------------------------------------------------------------
// separable circular prefilter (Discretization.m:54-72): one pass per axis.
// forward: out[i] = sum_a k[a] in[(i - a) mod N];  transpose: in[(i + a) mod N].
// interleaved_out: write [P][Bp] instead of [B][P].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_filt_axis(
    const float *__restrict__ in, float *__restrict__ out, const float *__restrict__ k, int K,
    int Nz, int Nx, int B, int axis, int transp, int interleaved_out, int Bp)
{
    const size_t P = (size_t)Nz * Nx;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * B) return;
    const int b = (int)(idx / P);
    const int pxl = (int)(idx - (size_t)b * P);
    const int ix = pxl / Nz, iz = pxl - ix * Nz;
    const float *src = in + (size_t)b * P;
    const int N = axis == 0 ? Nz : Nx;
    const int i0 = axis == 0 ? iz : ix;
    const int sgn = transp ? 1 : -1;
    double acc = 0.0;
    for (int a = -K; a <= K; ++a) {
        int j = i0 + sgn * a;   // |a| <= K < N: one wrap at most
        j = j < 0 ? j + N : (j >= N ? j - N : j);
        const float v = axis == 0 ? src[(size_t)ix * Nz + j] : src[(size_t)j * Nz + iz];
        acc = fma((double)k[a + K], (double)v, acc);
    }
    if (interleaved_out) out[(size_t)pxl * Bp + b] = (float)acc;
    else out[idx] = (float)acc;
}

// ---------------------------------------------------------------------------
// separable circular prefilter, LDS-staged passes (sums in PfAcc, FP32: see below).
// z pass: COLS columns of Nz samples per workgroup, each staged once with K
// wrapped samples on either side.  Planar output [B][P] (float4 stores where
// aligned); interleaved output [P][Bp] (IL) takes the columns as consecutive
// sinograms of one x column, so its stores are 64-byte runs for 16 columns
// (columns b >= B are written as zero).
// x pass: one sinogram's ROWS consecutive z rows over all Nx columns per
// workgroup, planar in and out.
// forward: out[i] = sum_a k[a] in[(i - a) mod N];  transpose: in[(i + a) mod N].
// ---------------------------------------------------------------------------
constexpr int kFzCols = 16;
constexpr int kFxRows = 16;
// batches: 16 columns per z-pass workgroup (planar and interleaved output), 32 z rows per
// x-pass workgroup (128-byte runs per x; 34 us against 37 us for 16 rows at 64 x 400^2)
constexpr int kFzColsB = 16, kFzColsIL = 16, kFxRowsB = 32;

constexpr int kFiltR = 8;   // outputs per thread of the register-blocked prefilter passes
// prefilter sums: FP32 (the taps decay geometrically from a dominant centre tap, so an
// FP32 sum of the 2K + 1 terms is within a few ulps; FP64 sums cost 17 ms per 50-iteration
// batch-64 solve at 400^2 and changed no result beyond the operator tolerances)
using PfAcc = float;

// KT > 0: K = KT known at compile time, each thread forms kFiltR consecutive
// outputs of a column from one register window of kFiltR + 2 KT samples
// (staged columns carry kFiltR extra samples so the last window stays inside).
template <bool IL, int KT, int COLS = kFzCols>
__global__ __launch_bounds__(256) void k_filt_z(const float *__restrict__ in, float *__restrict__ out,
                                                const float *__restrict__ k, int K, int Nz, int Nx, int B, int Bp,
                                                int transp, int vst = 0)
{
    extern __shared__ float sm[];
    const int S = (Nz + 2 * K + (KT ? kFiltR : 0)) | 1;   // column stride in LDS
    PfAcc *tk = reinterpret_cast<PfAcc *>(sm);
    float *xs = sm + 2 * (2 * K + 1 + 1);
    const size_t P = (size_t)Nz * Nx;
    // columns of this workgroup: IL: (ix, b0 + c); planar: (b, ix0 + c) over the flat b * Nx + ix index
    int ixc = 0, b0 = 0;
    size_t col0 = 0;
    if (IL) {
        ixc = blockIdx.x;
        b0 = blockIdx.y * COLS;
    } else {
        col0 = (size_t)blockIdx.x * COLS;
    }
    for (int i = threadIdx.x; i < 2 * K + 1; i += blockDim.x) tk[i] = (PfAcc)k[i];
    const int W = Nz + 2 * K + (KT ? kFiltR : 0);
    // column c: sample i of the staged window holds in[(i - K) mod Nz] (0 past Nz + 2K, the
    // overrun of the last register window, whose outputs are not stored).  Column bases are
    // uniform; a column past the data (IL: sinogram >= B; planar: past B * Nx) reads column
    // 0 and stages zeros.  Thread t stages samples t, t + 256, ... of every column, its
    // wrapped source offset computed once per sample for all COLS columns.
    const float *colp[COLS];
    bool cok[COLS];
#pragma unroll
    for (int c = 0; c < COLS; ++c) {
        if (IL) {
            cok[c] = b0 + c < B;
            colp[c] = in + (cok[c] ? (size_t)(b0 + c) * P : 0) + (size_t)ixc * Nz;
        } else {
            cok[c] = col0 + c < (size_t)B * Nx;
            colp[c] = in + (cok[c] ? (col0 + c) * Nz : 0);
        }
    }
    for (int i = threadIdx.x; i < W; i += 256) {
        const bool pad = i >= Nz + 2 * K;
        int j = i - K;
        j = j < 0 ? j + Nz : (j >= Nz ? j - Nz : j);
        j = pad ? 0 : j;
        float v[COLS];
#pragma unroll
        for (int c = 0; c < COLS; ++c) v[c] = colp[c][(unsigned)j];
#pragma unroll
        for (int c = 0; c < COLS; ++c) xs[c * S + i] = (pad || !cok[c]) ? 0.0f : v[c];
    }
    __syncthreads();
    // out[i] = sum_a k[a + K] xs[i + K - sgn a]: forward reads xs[i + 2K - t], transpose xs[i + t] (t = a + K)
    if (KT) {   // planar: 256 / COLS threads per column, blocks of kFiltR outputs
        constexpr int NT = 2 * (KT ? KT : 1) + 1, NW = kFiltR + NT - 1;
        PfAcc tr[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) tr[t] = tk[t];
        constexpr int TPC = 256 / COLS;   // threads per column (planar)
        const int c = IL ? threadIdx.x % COLS : threadIdx.x / TPC, ib = IL ? threadIdx.x / COLS : threadIdx.x % TPC;
        const bool live = IL ? b0 + c < Bp : col0 + c < (size_t)B * Nx;
        if (!live) return;
        for (int iz0 = ib * kFiltR; iz0 < Nz; iz0 += (IL ? 256 / COLS : TPC) * kFiltR) {
            const float *x = xs + c * S + iz0;
            PfAcc w[NW], acc[kFiltR];
#pragma unroll
            for (int i = 0; i < NW; ++i) w[i] = (PfAcc)x[i];
#pragma unroll
            for (int q = 0; q < kFiltR; ++q) {
                acc[q] = 0;
                if (transp) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[q] = fma(tr[t], w[q + t], acc[q]);
                } else {
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[q] = fma(tr[t], w[q + NT - 1 - t], acc[q]);
                }
            }
            if (!IL && vst && iz0 + kFiltR <= Nz) {   // 16-byte aligned (checked by the launcher)
                float4 *o4 = reinterpret_cast<float4 *>(out + (col0 + c) * Nz + iz0);
#pragma unroll
                for (int q = 0; q < kFiltR; q += 4)
                    o4[q / 4] = make_float4((float)acc[q], (float)acc[q + 1], (float)acc[q + 2], (float)acc[q + 3]);
                continue;
            }
#pragma unroll
            for (int q = 0; q < kFiltR; ++q) {
                const int iz = iz0 + q;
                if (iz >= Nz) break;
                if (IL)
                    out[((size_t)ixc * Nz + iz) * Bp + b0 + c] = b0 + c < B ? (float)acc[q] : 0.0f;
                else
                    out[(col0 + c) * Nz + iz] = (float)acc[q];
            }
        }
        return;
    }
    if (IL) {
        const int c = threadIdx.x % COLS;
        const int b = b0 + c;
        if (b >= Bp) return;
        for (int iz = threadIdx.x / COLS; iz < Nz; iz += 256 / COLS) {
            const float *x = xs + c * S + iz;
            PfAcc acc = 0;
            if (transp)
                for (int t = 0; t <= 2 * K; ++t) acc = fma(tk[t], (PfAcc)x[t], acc);
            else
                for (int t = 0; t <= 2 * K; ++t) acc = fma(tk[t], (PfAcc)x[2 * K - t], acc);
            out[((size_t)ixc * Nz + iz) * Bp + b] = b < B ? (float)acc : 0.0f;
        }
    } else {
        for (int c = 0; c < COLS; ++c) {
            if (col0 + c >= (size_t)B * Nx) break;
            for (int iz = threadIdx.x; iz < Nz; iz += blockDim.x) {
                const float *x = xs + c * S + iz;
                PfAcc acc = 0;
                if (transp)
                    for (int t = 0; t <= 2 * K; ++t) acc = fma(tk[t], (PfAcc)x[t], acc);
                else
                    for (int t = 0; t <= 2 * K; ++t) acc = fma(tk[t], (PfAcc)x[2 * K - t], acc);
                out[(col0 + c) * Nz + iz] = (float)acc;
            }
        }
    }
}

template <int KT, int ROWS = kFxRows>
__global__ __launch_bounds__(256) void k_filt_x(const float *__restrict__ in, float *__restrict__ out,
                                                const float *__restrict__ k, int K, int Nz, int Nx, int transp)
{
    extern __shared__ float sm[];
    PfAcc *tk = reinterpret_cast<PfAcc *>(sm);
    float *xs = sm + 2 * (2 * K + 1 + 1);   // [Nx + 2K (+ kFiltR)][ROWS]
    const int iz0 = blockIdx.x * ROWS, b = blockIdx.y;
    const size_t P = (size_t)Nz * Nx;
    const float *src = in + (size_t)b * P;
    for (int i = threadIdx.x; i < 2 * K + 1; i += blockDim.x) tk[i] = (PfAcc)k[i];
    const int r = threadIdx.x % ROWS;
    const bool rv = iz0 + r < Nz;
    // element e = i * ROWS + r: row i of the wrapped x axis, z row iz0 + r
    stage_batched<16, 256>(
        threadIdx.x, (Nx + 2 * K + (KT ? kFiltR : 0)) * ROWS,
        [&](int e) {
            const int i = e / ROWS, rr = e % ROWS;
            if (i >= Nx + 2 * K) return 0.0f;   // window overrun of the last block (outputs not stored)
            int j = i - K;
            j = j < 0 ? j + Nx : (j >= Nx ? j - Nx : j);
            return iz0 + rr < Nz ? src[(size_t)j * Nz + iz0 + rr] : 0.0f;
        },
        [&](int e, float v) { xs[e] = v; });
    __syncthreads();
    if (!rv) return;
    float *dst = out + (size_t)b * P + iz0 + r;
    if (KT) {   // blocks of kFiltR consecutive x outputs per thread
        constexpr int NT = 2 * (KT ? KT : 1) + 1, NW = kFiltR + NT - 1;
        PfAcc tr[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) tr[t] = tk[t];
        for (int ix0 = (threadIdx.x / ROWS) * kFiltR; ix0 < Nx; ix0 += (256 / ROWS) * kFiltR) {
            const float *x = xs + ix0 * ROWS + r;
            PfAcc w[NW], acc[kFiltR];
#pragma unroll
            for (int i = 0; i < NW; ++i) w[i] = (PfAcc)x[i * ROWS];
#pragma unroll
            for (int q = 0; q < kFiltR; ++q) {
                acc[q] = 0;
                if (transp) {
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[q] = fma(tr[t], w[q + t], acc[q]);
                } else {
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[q] = fma(tr[t], w[q + NT - 1 - t], acc[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < kFiltR; ++q) {
                if (ix0 + q >= Nx) break;
                dst[(size_t)(ix0 + q) * Nz] = (float)acc[q];
            }
        }
        return;
    }
    for (int ix = threadIdx.x / ROWS; ix < Nx; ix += 256 / ROWS) {
        const float *x = xs + ix * ROWS + r;
        PfAcc acc = 0;
        if (transp)
            for (int t = 0; t <= 2 * K; ++t) acc = fma(tk[t], (PfAcc)x[t * ROWS], acc);
        else
            for (int t = 0; t <= 2 * K; ++t) acc = fma(tk[t], (PfAcc)x[(2 * K - t) * ROWS], acc);
        dst[(size_t)ix * Nz] = (float)acc;
    }
}

// ---------------------------------------------------------------------------
// forward scatter, bucketed: one workgroup per (detector, bucket, batch chunk).
// The bucket's records (pixel, travel time) are listed at plan time; every
// record deposits its two interpolation weights (times each SIR tap) into an
// LDS window of WL bins x Bc sinograms with LDS float atomics; the window is
// then written once.  Windows of neighbouring buckets overlap and are summed
// in a fixed order by k_bins_to_y.
// ---------------------------------------------------------------------------
template <bool SIR, bool VEC>
__global__ __launch_bounds__(256) void k_scatter_bk(
    const int2 *__restrict__ rec, const int32_t *__restrict__ rec_off, const int32_t *__restrict__ rec_k,
    const float *__restrict__ imgI, float *__restrict__ hpart, const float *__restrict__ sir,
    const int32_t *__restrict__ swin, int Ls, BinGeom g, int Wb, int nbk, int lo_off, int WL, int WLp,
    int Bp, int B, int Bc)
{
    extern __shared__ float L[];   // [Bc][WLp]
    const int d = blockIdx.x, kb = blockIdx.y, b0 = blockIdx.z * Bc;
    const int nb = min(Bc, B - b0);
    for (int i = threadIdx.x; i < Bc * WLp; i += blockDim.x) L[i] = 0.0f;
    __syncthreads();
    const int base = kb * Wb + lo_off;
    const int r0 = rec_off[(size_t)d * (nbk + 1) + kb], r1 = rec_off[(size_t)d * (nbk + 1) + kb + 1];
    const int span = g.nT + 2 * g.K;
    for (int r = r0 + (int)threadIdx.x; r < r1; r += blockDim.x) {
        const int2 e = rec[r];
        const Decoded dc = decode_tau(e.y, g);
        const int jl = 3 * dc.nn + dc.m - base;
        const float *c = imgI + (size_t)e.x * Bp + b0;
        const float u1 = dc.u, u0 = 1.0f - dc.u;
        if (!SIR) {
            if ((unsigned)jl >= (unsigned)(WL - 1)) continue;   // cannot happen for a consistent plan
            if (VEC) {
                for (int b = 0; b < nb; b += 4) {
                    const float4 v = *reinterpret_cast<const float4 *>(c + b);
                    const float cv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (b + q < nb) {
                            atomicAdd(&L[(b + q) * WLp + jl], cv[q] * u0);
                            atomicAdd(&L[(b + q) * WLp + jl + 1], cv[q] * u1);
                        }
                    }
                }
            } else {
                for (int b = 0; b < nb; ++b) {
                    const float cv = c[b];
                    atomicAdd(&L[b * WLp + jl], cv * u0);
                    atomicAdd(&L[b * WLp + jl + 1], cv * u1);
                }
            }
        } else {
            const int k = rec_k[r];
            const int s0 = swin[2 * k], len = swin[2 * k + 1];
            const float *hk = sir + (size_t)k * Ls;
            const int half = Ls / 2;
            const float c0 = c[0];   // the first sinogram's value, kept for the common single-sinogram case
            // taps in runs of 8: the run's SIR weights are loaded before any is deposited
            for (int cs = 0; cs < len; cs += 8) {
                float wt[8];
                int jj[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int s = s0 + cs + t;
                    const int nn = dc.nn + s - half;
                    jj[t] = jl + 3 * (s - half);
                    // taps past the window, outside the bin range or (cannot happen for a
                    // consistent plan) outside the LDS window deposit nothing
                    const bool ok = cs + t < len && (unsigned)nn < (unsigned)span &&
                                    (unsigned)jj[t] < (unsigned)(WL - 1);
                    wt[t] = ok ? hk[s] : 0.0f;
                    if (!ok) jj[t] = -1;
                }
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    if (jj[t] < 0) continue;
                    const float w0 = wt[t] * u0, w1 = wt[t] * u1;
                    atomicAdd(&L[jj[t]], c0 * w0);
                    atomicAdd(&L[jj[t] + 1], c0 * w1);
                    for (int b = 1; b < nb; ++b) {
                        const float cv = c[b];
                        atomicAdd(&L[b * WLp + jj[t]], cv * w0);
                        atomicAdd(&L[b * WLp + jj[t] + 1], cv * w1);
                    }
                }
            }
        }
    }
    __syncthreads();
    float *dst = hpart + (((size_t)d * nbk + kb) * WL) * Bp + b0;
    for (int i = threadIdx.x; i < WL * nb; i += blockDim.x) {
        const int jl = i / nb, b = i - jl * nb;
        dst[(size_t)jl * Bp + b] = L[b * WLp + jl];
    }
}

// ---------------------------------------------------------------------------
// window sums + polyphase hat filter (M^T): y[b][d][r] = sum_m sum_i w[m][i] bins(3(r-kmin-i+K)+m)
// One workgroup per (time tile, detector, batch chunk).
// ---------------------------------------------------------------------------
constexpr int kYTile = 128;

// DIRECT: read the bins [D][NB][Bp] themselves (pull forward) instead of the
// overlapping bucket windows
// NW > 0: taps per phase known at compile time (the taps in registers instead of indexed in
// the kernel arguments at run time; the same sums in the same order)
template <bool DIRECT, int NW = 0>
__global__ __launch_bounds__(256) void k_bins_to_y(
    const float *__restrict__ hpart, float *__restrict__ y, HatTaps taps, BinGeom g, int Wb, int nbk,
    int lo_off, int WL, int Bp, int D, int B, int Bc, float scale, const int2 *__restrict__ win)
{
    extern __shared__ float sm[];
    const int d = blockIdx.y, b0 = blockIdx.z * Bc;
    const int nb = min(Bc, B - b0);
    const int nT = g.nT;
    const int t0 = blockIdx.x * kYTile, t1 = min(nT, t0 + kYTile);
    const int kmax = taps.kmin + taps.nw - 1;
    const int jlo = 3 * (t0 - kmax + g.K);
    const int jhi = 3 * (t1 - 1 - taps.kmin + g.K) + 2;
    const int nj = jhi - jlo + 1;
    const int njp = nj | 1;
    // load summed bins for the tile: bs[b][j - jlo]; window k covers bins [k*Wb + lo_off, k*Wb + lo_off + WL)
    for (int i = threadIdx.x; i < nj * nb; i += blockDim.x) {
        const int jj = i / nb, b = i - jj * nb;
        float acc = 0.0f;
        if (DIRECT) {
            acc = hpart[((size_t)d * g.NB + jlo + jj) * Bp + b0 + b];
        } else {
            const int rel = jlo + jj - lo_off;          // >= 0: jlo >= 3 and lo_off <= 0
            const int ka = rel >= WL - 1 ? (rel - (WL - 1) + Wb - 1) / Wb : 0;
            const int kz = min(nbk - 1, rel / Wb);
            for (int k = ka; k <= kz; ++k)
                acc += hpart[(((size_t)d * nbk + k) * WL + (rel - k * Wb)) * Bp + b0 + b];
        }
        sm[b * njp + jj] = acc;
    }
    __syncthreads();
    const int nr = t1 - t0;
    float wr[3 * (NW ? NW : 1)];
    if constexpr (NW > 0) {
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int q = 0; q < NW; ++q) wr[m * NW + q] = taps.w[m][q];
    }
    for (int i = threadIdx.x; i < nr * nb; i += blockDim.x) {
        const int b = i / nr, r = t0 + (i - b * nr);
        const float *bs = sm + b * njp;
        float acc = 0.0f;
        if constexpr (NW > 0) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int jb = 3 * (r - taps.kmin + g.K) + m - jlo;
#pragma unroll
                for (int q = 0; q < NW; ++q) acc = fmaf(wr[m * NW + q], bs[jb - 3 * q], acc);
            }
        } else {
            for (int m = 0; m < 3; ++m) {
                const int jb = 3 * (r - taps.kmin + g.K) + m - jlo;
                for (int q = 0; q < taps.nw; ++q) acc = fmaf(taps.w[m][q], bs[jb - 3 * q], acc);
            }
        }
        if (win) {
            const int2 wb = win[b0 + b];
            if (r < wb.x || r >= wb.y) acc = 0.0f;
        }
        y[((size_t)(b0 + b) * D + d) * nT + r] = scale * acc;
    }
}

// ---------------------------------------------------------------------------
// FIR along sinogram rows (MSOTmodel.m:259-269): out[t] = scale * sum_i taps[i] x[t + off - i],
// x = 0 outside [0, nT).  Convolution: taps = IR, correlation: taps = flip(IR).
// Each thread owns 4 consecutive outputs with a sliding register window; FP64 sums.
// ---------------------------------------------------------------------------
constexpr int kFirTile = 1024;   // outputs per workgroup (256 threads x 4)

__global__ __launch_bounds__(256) void k_fir(const float *__restrict__ x, float *__restrict__ out,
                                             const float *__restrict__ taps, int L, int per_det, int off,
                                             int D, int nT, float scale, const int2 *__restrict__ win)
{
    extern __shared__ float sm[];
    const int row = blockIdx.y;                   // b * D + d
    const int d = row % D;
    const int t0 = blockIdx.x * kFirTile;
    const int xbase = t0 + off - (L - 1);         // first input sample needed
    const int nx = kFirTile + L + 2;
    float *xs = sm;
    float *ts = sm + nx;
    const float *xr = x + (size_t)row * nT;
    int lo = 0, hi = nT;   // samples outside [lo, hi) are zero on input and output
    if (win) {
        const int2 wb = win[row / D];
        lo = max(0, wb.x);
        hi = min(nT, wb.y);
    }
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
        const int t = xbase + i;
        xs[i] = (t >= lo && t < hi) ? xr[t] : 0.0f;
    }
    const float *tp = taps + (per_det ? (size_t)d * L : 0);
    for (int i = threadIdx.x; i < L; i += blockDim.x) ts[i] = tp[i];
    __syncthreads();
    // output t = t0 + 4*tid + k needs x[t + off - i] = xs[4*tid + k + L - 1 - i]
    const int o = 4 * threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    float v0 = xs[o + L - 1], v1 = xs[o + L], v2 = xs[o + L + 1], v3 = xs[o + L + 2];
    for (int i = 0; i < L; ++i) {
        const double w = (double)ts[i];
        a0 = fma(w, (double)v0, a0);
        a1 = fma(w, (double)v1, a1);
        a2 = fma(w, (double)v2, a2);
        a3 = fma(w, (double)v3, a3);
        v3 = v2; v2 = v1; v1 = v0;
        v0 = (i + 1 < L) ? xs[o + L - 2 - i] : 0.0f;
    }
    float *orow = out + (size_t)row * nT;
    const int t = t0 + o;
    if (t < nT) orow[t] = (t >= lo && t < hi) ? scale * (float)a0 : 0.0f;
    if (t + 1 < nT) orow[t + 1] = (t + 1 >= lo && t + 1 < hi) ? scale * (float)a1 : 0.0f;
    if (t + 2 < nT) orow[t + 2] = (t + 2 >= lo && t + 2 < hi) ? scale * (float)a2 : 0.0f;
    if (t + 3 < nT) orow[t + 3] = (t + 3 >= lo && t + 3 < hi) ? scale * (float)a3 : 0.0f;
}

// FIR, register-blocked: each thread owns 16 consecutive outputs; taps are
// consumed in groups of 16 (zero-padded to Lp = 16 * ceil(L / 16)) against a
// 31-sample register window, so every staged sample feeds 16 FMAs.  A half
// workgroup (128 threads) covers a tile of kFirSpan outputs of one row; the
// two halves take consecutive (row, tile) pairs.  Outputs go back through LDS
// for contiguous stores; the taps are FP64 scalar operands.  Same sums as k_fir
// (FP64, taps ascending).
constexpr int kFirOut = 16;
constexpr int kFirSpan = 128 * kFirOut;

__host__ __device__ constexpr int fir16_padded_taps(int L) { return (L + 15) & ~15; }
// LDS window per half workgroup: 16 x NBLK floats, element e at (e % 16) * NBLK + e / 16
// (NBLK odd), so 16-float runs owned by consecutive threads sit in distinct banks
__host__ __device__ constexpr int fir16_nblk(int Lp) { return ((kFirSpan + Lp + 32) / 16) | 1; }
__host__ __device__ constexpr int fir16_window(int Lp) { return 16 * fir16_nblk(Lp); }

// LPT > 0: the padded tap count Lp known at compile time (the tap loop unrolls fully, the
// register window needs no moves and the LDS offsets are immediates); LPT = 0: any L.
// SPLIT: for few rows (one sinogram: 256 tiles for the whole filter) both halves of the
// workgroup take the same tile, half 0 the first ceil(Lp / 32) groups of 16 taps and half
// 1 the rest, and the two FP64 partial sums are added at the end (one tile per workgroup,
// twice the workgroups; the sum of each half in tap order, then the two halves).
template <int LPT, bool SPLIT = false>
__global__ __launch_bounds__(256) void k_fir16(const float *__restrict__ x, float *__restrict__ out,
                                               const double *__restrict__ taps64, int L, int per_det, int off, int D,
                                               int nT, int rows, float scale, const int2 *__restrict__ win)
{
    extern __shared__ float sm[];
    const int Lp = LPT ? LPT : fir16_padded_taps(L), nx = fir16_window(Lp), NBLK = fir16_nblk(Lp);
    const int hf = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7), lt = threadIdx.x & 127;
    float *xs = SPLIT ? sm : sm + (size_t)hf * nx;
    const int ntile = (nT + kFirSpan - 1) / kFirSpan;
    const int tile = SPLIT ? (int)blockIdx.x : (int)blockIdx.x * 2 + hf;
    const bool live = tile < rows * ntile;
    const int row = live ? tile / ntile : 0;
    const int t0 = live ? (tile - row * ntile) * kFirSpan : 0;
    int lo = 0, hi = nT;   // samples outside [lo, hi) are zero on input and output
    if (win) {
        const int2 wb = win[row / D];
        lo = max(0, wb.x);
        hi = min(nT, wb.y);
    }
    // xs[m] = x[xbase + m], xbase = t0 + off - Lp - 15
    const int xbase = t0 + off - Lp - 15;
    const float *xr = x + (size_t)row * nT;
    auto ldx = [&](int i) {
        const int t = xbase + i;
        return (live && t >= lo && t < hi) ? xr[t] : 0.0f;
    };
    auto stx = [&](int i, float v) { xs[(i & 15) * NBLK + (i >> 4)] = v; };
    if constexpr (SPLIT)
        stage_batched<8, 256>((int)threadIdx.x, nx, ldx, stx);
    else
        stage_batched<8, 128>(lt, nx, ldx, stx);
    // the row's taps, FP64 and zero-padded to Lp, read as scalars (uniform per wave)
    const double *tk = taps64 + (per_det ? (size_t)(row % D) * Lp : 0);
    __syncthreads();
    // output t0 + 16 lt + k at tap i reads element 16 lt + k - i + Lp + 15.  Taps go in
    // groups of 8: for the group starting at tap g the window is W[w] = element
    // 16 lt + Lp + 8 - g + w (w < 23), read as W[k - q + 7]; the next group's window is
    // this one moved up by 8 with 8 new elements below (lanes 0..7 of a block after a
    // group at an even multiple of 8, lanes 8..15 of the block before after an odd one).
    const int blk = lt + Lp / 16;
    // tap groups [i0a, i0b) of this thread (all of them unless SPLIT); the register window
    // at the start of the group at i0 is W[w] = element (w + 8, blk - i0 / 16) for w < 8 and
    // (w - 8, blk - i0 / 16 + 1) for w >= 8 (the state the shifts below leave)
    const int Sp = 16 * ((Lp / 16 + 1) / 2);
    const int i0a = SPLIT && hf ? Sp : 0;
    double W[23], acc[kFirOut];
#pragma unroll
    for (int w = 0; w < 8; ++w) W[w] = (double)xs[(w + 8) * NBLK + blk - i0a / 16];
#pragma unroll
    for (int w = 8; w < 23; ++w) W[w] = (double)xs[(w - 8) * NBLK + blk - i0a / 16 + 1];
#pragma unroll
    for (int k = 0; k < kFirOut; ++k) acc[k] = 0.0;
    auto group = [&](int g) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double w = tk[g + q];
#pragma unroll
            for (int k = 0; k < kFirOut; ++k) acc[k] = fma(w, W[k - q + 7], acc[k]);
        }
    };
    auto shift = [&](int bn, int lane0) {
#pragma unroll
        for (int w = 22; w >= 8; --w) W[w] = W[w - 8];
#pragma unroll
        for (int w = 0; w < 8; ++w) W[w] = (double)xs[(lane0 + w) * NBLK + bn];
    };
    auto step = [&](int i0) {
        group(i0);
        shift(blk - i0 / 16, 0);
        group(i0 + 8);
        if (i0 + 16 < Lp) shift(blk - i0 / 16 - 1, 8);
    };
    if constexpr (SPLIT && LPT > 0) {
        constexpr int S = 16 * ((LPT / 16 + 1) / 2);
        if (hf == 0) {
#pragma unroll
            for (int i0 = 0; i0 < S; i0 += 16) step(i0);
        } else {
#pragma unroll
            for (int i0 = S; i0 < LPT; i0 += 16) step(i0);
        }
    } else if constexpr (SPLIT) {
        for (int i0 = i0a; i0 < (hf ? Lp : Sp); i0 += 16) step(i0);
    } else if constexpr (LPT > 0) {
#pragma unroll
        for (int i0 = 0; i0 < LPT; i0 += 16) step(i0);
    } else {
        for (int i0 = 0; i0 < Lp; i0 += 16) step(i0);   // Lp is a multiple of 16
    }
    __syncthreads();   // xs is reused for the outputs (same layout)
    if constexpr (SPLIT) {   // half 1's partial sums through LDS (past xs) into half 0's
        double *ex = reinterpret_cast<double *>(sm + nx);
        if (hf) {
#pragma unroll
            for (int k = 0; k < kFirOut; ++k) ex[k * 128 + lt] = acc[k];
        }
        __syncthreads();
        if (!hf) {
#pragma unroll
            for (int k = 0; k < kFirOut; ++k) {
                const int t = t0 + 16 * lt + k;
                xs[k * NBLK + lt] = (t >= lo && t < hi) ? scale * (float)(acc[k] + ex[k * 128 + lt]) : 0.0f;
            }
        }
        __syncthreads();
        if (!live) return;
        float *orow = out + (size_t)row * nT;
        for (int i = threadIdx.x; i < kFirSpan && t0 + i < nT; i += 256) orow[t0 + i] = xs[(i & 15) * NBLK + (i >> 4)];
        return;
    }
#pragma unroll
    for (int k = 0; k < kFirOut; ++k) {
        const int t = t0 + 16 * lt + k;
        xs[k * NBLK + lt] = (t >= lo && t < hi) ? scale * (float)acc[k] : 0.0f;
    }
    __syncthreads();
    if (!live) return;
    float *orow = out + (size_t)row * nT;
    for (int i = lt; i < kFirSpan && t0 + i < nT; i += 128) orow[t0 + i] = xs[(i & 15) * NBLK + (i >> 4)];
}

// FIR on the matrix cores (shared IR, FP64 sums of the exact FP32 products): for a block
// of 16 outputs t = tb .. tb + 15 of 16 rows, y[t] = sum_i taps[i] x[t + off - i] is the
// product of the 16 x K Toeplitz matrix T[a][j] = taps[a + L - 1 - j] (zero outside [0, L))
// with the K x 16 window X[j][r] = x_r[tb + off - (L - 1) + j], K = 4 kFmNS >= L + 15, done
// as kFmNS v_mfma_f64_16x16x4_f64 steps.  T's fragments are the same for every block and
// stay in registers (lane l: row l % 16, k-column l / 16 of each step); a workgroup stages
// 16 rows x kFmW samples in LDS (zero outside the row's window and [0, nT)), each of its 4
// waves computes kFmBlk blocks, and the outputs go back through LDS for row-contiguous
// stores.
constexpr int kFmNS = 54;                  // k steps: IRs of up to 4 * 54 - 15 = 201 taps
constexpr int kFmBlk = 4;                  // 16-output blocks per wave
constexpr int kFmTT = 4 * kFmBlk * 16;     // outputs per row and workgroup
constexpr int kFmW = kFmTT + 4 * kFmNS;    // staged samples per row
constexpr int kFmWp = 516;                 // LDS row stride: = 4 mod 64, >= kFmW
static_assert(kFmWp >= kFmW && kFmWp % 64 == 4, "LDS row stride");
typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_fir_mfma(const float *__restrict__ x, float *__restrict__ out,
                                                  const double *__restrict__ taps, int L, int off, int D, int nT,
                                                  int rows, float scale, const int2 *__restrict__ win)
{
    __shared__ float xs[16 * kFmWp];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = blockIdx.y * 16, t0 = blockIdx.x * kFmTT;
    const int sbase = t0 + off - (L - 1);   // sample at staged index 0
    constexpr float kInvW = 1.0f / (float)kFmW;
    stage_batched<8, 256>(
        threadIdx.x, 16 * kFmW,
        [&](int e) {
            int r, m;
            divmod_small(e, kFmW, kInvW, r, m);
            const int row = r0 + r;
            if (row >= rows) return 0.0f;
            int lo = 0, hi = nT;   // samples outside [lo, hi) are zero on input and output
            if (win) {
                const int2 wb = win[row / D];
                lo = max(0, wb.x);
                hi = min(nT, wb.y);
            }
            const int t = sbase + m;
            return (t >= lo && t < hi) ? x[(size_t)row * nT + t] : 0.0f;
        },
        [&](int e, float v) {
            int r, m;
            divmod_small(e, kFmW, kInvW, r, m);
            xs[r * kFmWp + m] = v;
        });
    const int a = lane & 15, q = lane >> 4;
    double A[kFmNS];
#pragma unroll
    for (int s = 0; s < kFmNS; ++s) {
        const int i = a + L - 1 - (4 * s + q);
        A[s] = (i >= 0 && i < L) ? taps[i] : 0.0;
    }
    __syncthreads();
    f64x4 acc[kFmBlk];
#pragma unroll
    for (int b = 0; b < kFmBlk; ++b) acc[b] = f64x4{0.0, 0.0, 0.0, 0.0};
    // lane l supplies X[4 s + l / 16][row l % 16] of block b: staged index 16 (kFmBlk w + b) + 4 s + l / 16
    const float *xb = xs + a * kFmWp + q + wv * (kFmBlk * 16);
#pragma unroll
    for (int s = 0; s < kFmNS; ++s) {
#pragma unroll
        for (int b = 0; b < kFmBlk; ++b)
            acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[s], (double)xb[16 * b + 4 * s], acc[b], 0, 0, 0);
    }
    // xs is reused for the outputs; the f64 16x16 result layout: lane l, register v holds
    // D[4 v + l / 16][l % 16], i.e. y[row l % 16][tb + 4 v + l / 16]
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kFmBlk; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            xs[a * kFmWp + wv * (kFmBlk * 16) + 16 * b + 4 * v + q] = scale * (float)acc[b][v];
    __syncthreads();
    constexpr float kInvT = 1.0f / (float)kFmTT;
    for (int e = threadIdx.x; e < 16 * kFmTT; e += 256) {
        int r, m;
        divmod_small(e, kFmTT, kInvT, r, m);
        const int row = r0 + r, t = t0 + m;
        if (row >= rows || t >= nT) continue;
        int lo = 0, hi = nT;
        if (win) {
            const int2 wb = win[row / D];
            lo = max(0, wb.x);
            hi = min(nT, wb.y);
        }
        out[(size_t)row * nT + t] = (t >= lo && t < hi) ? xs[r * kFmWp + m] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// samples -> bins (M): bins[d][j][b] = sum_i w[m][i] z[b][d][n + kmin + i], z = 0 outside [0, nT).
// One workgroup per (bin tile, detector, batch chunk).
// ---------------------------------------------------------------------------
constexpr int kBinTile = 384;   // 128 samples x 3 phases

__global__ __launch_bounds__(256) void k_z_to_bins(const float *__restrict__ z, float *__restrict__ bins,
                                                   HatTaps taps, BinGeom g, int D, int B, int Bp, int Bc,
                                                   const int2 *__restrict__ win)
{
    extern __shared__ float sm[];
    const int d = blockIdx.y, b0 = blockIdx.z * Bc;
    const int nb = min(Bc, B - b0);
    const int nT = g.nT, NB = g.NB;
    const int j0 = blockIdx.x * kBinTile, j1 = min(NB, j0 + kBinTile);
    const int nlo = j0 / 3 - g.K, nhi = (j1 - 1) / 3 - g.K;
    const int zlo = nlo + taps.kmin, zhi = nhi + taps.kmin + taps.nw - 1;
    const int nz = zhi - zlo + 1;
    const int nzp = nz | 1;
    for (int i = threadIdx.x; i < nz * nb; i += blockDim.x) {
        const int b = i / nz, q = i - b * nz;
        const int t = zlo + q;
        int lo = 0, hi = nT;
        if (win) {
            const int2 wb = win[b0 + b];
            lo = max(0, wb.x);
            hi = min(nT, wb.y);
        }
        sm[b * nzp + q] = (t >= lo && t < hi) ? z[((size_t)(b0 + b) * D + d) * nT + t] : 0.0f;
    }
    __syncthreads();
    float *brow = bins + (size_t)d * NB * Bp + b0;
    const int njt = j1 - j0;
    for (int i = threadIdx.x; i < njt * nb; i += blockDim.x) {
        const int jj = i / nb, b = i - jj * nb;
        const int j = j0 + jj;
        const int n = j / 3 - g.K, m = j - 3 * (j / 3);
        const float *zs = sm + b * nzp + (n + taps.kmin - zlo);
        float acc = 0.0f;
        for (int q = 0; q < taps.nw; ++q) acc = fmaf(taps.w[m][q], zs[q], acc);
        brow[(size_t)j * Bp + b] = acc;
    }
}

// batch-chunk forms of the two hat-filter kernels (Bp >= 16): a workgroup
// takes BC consecutive sinograms (lane b of the chunk = sinogram b0 + b), so
// bin-row stores and loads are BC-float runs; integer work per element is
// compile-time (no runtime divisions), and the taps sit in LDS.  Same sums in
// the same order as k_z_to_bins / k_bins_to_y.
template <int BC, int NW>
__global__ __launch_bounds__(256) void k_z_to_bins_bc(const float *__restrict__ z, float *__restrict__ bins,
                                                      HatTaps taps, BinGeom g, int D, int B, int Bp,
                                                      const int2 *__restrict__ win)
{
    extern __shared__ float sm[];
    float *tw = sm;                        // [3][nw]
    float *zsm = sm + 3 * MBREC_MAX_TAPS;  // [BC][nzp]
    const int d = blockIdx.y, b0 = blockIdx.z * BC;
    const int nT = g.nT, NB = g.NB, nw = NW ? NW : taps.nw;   // NW > 0: taps per phase known at compile time
    const int j0 = blockIdx.x * kBinTile, j1 = min(NB, j0 + kBinTile);
    const int nlo = j0 / 3 - g.K, nhi = (j1 - 1) / 3 - g.K;
    const int zlo = nlo + taps.kmin;
    const int nz = nhi - nlo + nw, nzp = nz | 1;
    for (int i = threadIdx.x; i < 3 * nw; i += 256) tw[i] = taps.w[i / nw][i % nw];
    int2 *wl = reinterpret_cast<int2 *>(zsm + BC * nzp);   // [BC] sample window per sinogram
    for (int b = threadIdx.x; b < BC; b += 256) {
        int2 w = make_int2(0, b0 + b < B ? nT : 0);
        if (win && b0 + b < B) w = make_int2(max(0, win[b0 + b].x), min(nT, win[b0 + b].y));
        wl[b] = w;
    }
    __syncthreads();
    const float invz = 1.0f / (float)nz;
    stage_batched<16, 256>(
        threadIdx.x, BC * nz,
        [&](int e) {
            int b, q;
            divmod_small(e, nz, invz, b, q);
            const int t = zlo + q;
            const int2 w = wl[b];
            return (t >= w.x && t < w.y) ? z[((size_t)(b0 + b) * D + d) * nT + t] : 0.0f;
        },
        [&](int e, float v) {
            int b, q;
            divmod_small(e, nz, invz, b, q);
            zsm[b * nzp + q] = v;
        });
    __syncthreads();
    const int b = threadIdx.x % BC, bg = b0 + b;
    if (bg >= Bp) return;
    float *brow = bins + (size_t)d * NB * Bp + bg;
    const float *zb = zsm + b * nzp;
    const int ns = nhi - nlo + 1;
    float wr[3 * (NW ? NW : 1)];   // NW > 0: the taps in registers
    if (NW) {
#pragma unroll
        for (int i = 0; i < 3 * NW; ++i) wr[i] = tw[i];
    }
    for (int k = threadIdx.x / BC; k < ns; k += 256 / BC) {
        const float *zs = zb + k;   // z[n + kmin + q] for n = nlo + k
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
        if (NW) {
#pragma unroll
            for (int q = 0; q < (NW ? NW : 1); ++q) {
                const float v = zs[q];
                a0 = fmaf(wr[q], v, a0);
                a1 = fmaf(wr[NW + q], v, a1);
                a2 = fmaf(wr[2 * NW + q], v, a2);
            }
        } else {
            for (int q = 0; q < nw; ++q) {
                const float v = zs[q];
                a0 = fmaf(tw[q], v, a0);
                a1 = fmaf(tw[nw + q], v, a1);
                a2 = fmaf(tw[2 * nw + q], v, a2);
            }
        }
        const int j = j0 + 3 * k;
        brow[(size_t)j * Bp] = a0;
        if (j + 1 < j1) brow[(size_t)(j + 1) * Bp] = a1;
        if (j + 2 < j1) brow[(size_t)(j + 2) * Bp] = a2;
    }
}

template <int BC, int TY, bool DIRECT, int NW>
__global__ __launch_bounds__(256) void k_bins_to_y_bc(
    const float *__restrict__ hpart, float *__restrict__ y, HatTaps taps, BinGeom g, int Wb, int nbk, int lo_off,
    int WL, int Bp, int D, int B, float scale, const int2 *__restrict__ win)
{
    extern __shared__ float sm[];
    float *tw = sm;
    float *bsm = sm + 3 * MBREC_MAX_TAPS;   // [BC][njp]
    const int d = blockIdx.y, b0 = blockIdx.z * BC;
    const int nT = g.nT, nw = NW ? NW : taps.nw;   // NW > 0: taps per phase known at compile time
    const int t0 = blockIdx.x * TY, t1 = min(nT, t0 + TY);
    const int kmax = taps.kmin + nw - 1;
    const int jlo = 3 * (t0 - kmax + g.K);
    const int jhi = 3 * (t1 - 1 - taps.kmin + g.K) + 2;
    const int nj = jhi - jlo + 1, njp = nj | 1;
    for (int i = threadIdx.x; i < 3 * nw; i += 256) tw[i] = taps.w[i / nw][i % nw];
    int2 *wl = reinterpret_cast<int2 *>(bsm + BC * njp);   // [BC] sample window per sinogram
    for (int b = threadIdx.x; b < BC; b += 256)
        wl[b] = (win && b0 + b < B) ? win[b0 + b] : make_int2(0, nT);
    if (DIRECT) {   // 16-byte element e = jj * (BC / 4) + c: sinograms b0 + 4c .. b0 + 4c + 3 of bin row jlo + jj
        constexpr int C4 = BC / 4, U = 8;
        const int n4 = nj * C4;
        for (int e0 = threadIdx.x; e0 < n4; e0 += 256 * U) {
            float4 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + 256 * u, jj = e / C4, c = e % C4;
                v[u] = (e < n4 && b0 + 4 * c < Bp)
                           ? *reinterpret_cast<const float4 *>(hpart + ((size_t)d * g.NB + jlo + jj) * Bp + b0 + 4 * c)
                           : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + 256 * u, jj = e / C4, c = e % C4;
                if (e < n4) {
                    float *q = bsm + 4 * c * njp + jj;
                    q[0] = v[u].x;
                    q[njp] = v[u].y;
                    q[2 * njp] = v[u].z;
                    q[3 * njp] = v[u].w;
                }
            }
        }
    } else {
        const int b = threadIdx.x % BC;
        for (int jj = threadIdx.x / BC; jj < nj; jj += 256 / BC) {
            float acc = 0.0f;
            if (b0 + b < Bp) {
                const int rel = jlo + jj - lo_off;
                const int ka = rel >= WL - 1 ? (rel - (WL - 1) + Wb - 1) / Wb : 0;
                const int kz = min(nbk - 1, rel / Wb);
                for (int k = ka; k <= kz; ++k)
                    acc += hpart[(((size_t)d * nbk + k) * WL + (rel - k * Wb)) * Bp + b0 + b];
            }
            bsm[b * njp + jj] = acc;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float wr[3 * (NW ? NW : 1)];   // NW > 0: the taps in registers
    if (NW) {
#pragma unroll
        for (int i = 0; i < 3 * NW; ++i) wr[i] = tw[i];
    }
    for (int b = wv; b < BC; b += 4) {
        const int bg = b0 + b;
        if (bg >= B) break;
        const float *bs = bsm + b * njp;
        const int lo = wl[b].x, hi = wl[b].y;
        float *yr = y + ((size_t)bg * D + d) * nT;
        for (int r = t0 + lane; r < t1; r += 64) {
            float acc = 0.0f;
            const float *bj = bs + 3 * (r - taps.kmin + g.K) - jlo;
            if (NW) {
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int q = 0; q < (NW ? NW : 1); ++q) acc = fmaf(wr[m * NW + q], bj[m - 3 * q], acc);
            } else {
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    for (int q = 0; q < nw; ++q) acc = fmaf(tw[m * nw + q], bj[m - 3 * q], acc);
            }
            if (r < lo || r >= hi) acc = 0.0f;
            yr[r] = scale * acc;
        }
    }
}

// ---------------------------------------------------------------------------
// transpose gather: one workgroup per 64-slot pixel tile; its 4 waves split the
// detectors (unrolled by U so several travel-time loads are in flight) and are
// reduced through LDS.  BC sinograms per pass (float4 loads of interleaved bins).
// ---------------------------------------------------------------------------
template <int BC>
__device__ __forceinline__ void load_bins(const float *p, float (&v)[BC])
{
    if constexpr (BC % 4 == 0) {
#pragma unroll
        for (int q = 0; q < BC; q += 4) {
            const float4 t = *reinterpret_cast<const float4 *>(p + q);
            v[q] = t.x; v[q + 1] = t.y; v[q + 2] = t.z; v[q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < BC; ++q) v[q] = p[q];
    }
}

template <int BC, int U, bool SIR>
__global__ __launch_bounds__(256) void k_gather(
    const int32_t *__restrict__ tq, const int32_t *__restrict__ pix, const float *__restrict__ bins,
    float *__restrict__ out, const int32_t *__restrict__ sidx, const float *__restrict__ sir,
    const int32_t *__restrict__ swin, int Ls, BinGeom g, int Pd, int P, int D, int B, int Bp, float scale)
{
    __shared__ float red[4][BC][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const int b0 = blockIdx.y * BC;
    const int span = g.nT + 2 * g.K;
    const size_t rowstride = (size_t)g.NB * Bp;
    float acc[BC], cmp[BC];
#pragma unroll
    for (int b = 0; b < BC; ++b) { acc[b] = 0.0f; cmp[b] = 0.0f; }
    for (int d0 = wv; d0 < D; d0 += 4 * U) {
        int32_t q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + 4 * u;
            q[u] = d < D ? tq[(size_t)d * Pd + slot] : INT32_MIN;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + 4 * u;
            const Decoded dc = decode_tau(q[u], g);
            const float *gb = bins + (size_t)d * rowstride + b0;
            if (!SIR) {
                if ((unsigned)dc.nn < (unsigned)span) {
                    const int j = 3 * dc.nn + dc.m;
                    float g0[BC], g1[BC];
                    load_bins<BC>(gb + (size_t)j * Bp, g0);
                    load_bins<BC>(gb + (size_t)(j + 1) * Bp, g1);
#pragma unroll
                    for (int b = 0; b < BC; ++b) kahan_add(acc[b], cmp[b], fmaf(dc.u, g1[b] - g0[b], g0[b]));
                }
            } else if (d < D) {
                const int k = sidx[(size_t)d * Pd + slot];
                const int s0 = swin[2 * k], len = swin[2 * k + 1];
                const float *hk = sir + (size_t)k * Ls;
                const int shift = dc.nn - Ls / 2;
                float part[BC];
#pragma unroll
                for (int b = 0; b < BC; ++b) part[b] = 0.0f;
                // taps in runs of ST whose loads are all issued before any is used; taps
                // outside the window or the bin range read nothing and add 0 (same order)
                constexpr int ST = BC <= 1 ? 8 : (BC <= 4 ? 4 : 2);
                for (int c = 0; c < len; c += ST) {
                    float wt[ST], g0[ST][BC], g1[ST][BC];
#pragma unroll
                    for (int t = 0; t < ST; ++t) {
                        const int s = s0 + c + t, nn = shift + s;
                        const bool ok = c + t < len && (unsigned)nn < (unsigned)span;
                        const int j = ok ? 3 * nn + dc.m : 0;
                        wt[t] = ok ? hk[s] : 0.0f;
                        if (ok) {
                            load_bins<BC>(gb + (size_t)j * Bp, g0[t]);
                            load_bins<BC>(gb + (size_t)(j + 1) * Bp, g1[t]);
                        } else {
#pragma unroll
                            for (int b = 0; b < BC; ++b) g0[t][b] = g1[t][b] = 0.0f;
                        }
                    }
#pragma unroll
                    for (int t = 0; t < ST; ++t)
#pragma unroll
                        for (int b = 0; b < BC; ++b)
                            part[b] = fmaf(wt[t], fmaf(dc.u, g1[t][b] - g0[t][b], g0[t][b]), part[b]);
                }
#pragma unroll
                for (int b = 0; b < BC; ++b) kahan_add(acc[b], cmp[b], part[b]);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < BC; ++b) red[wv][b][lane] = acc[b] - cmp[b];
    __syncthreads();
    if (wv == 0) {
        const int px = pix[slot];
        if (px >= 0) {
#pragma unroll
            for (int b = 0; b < BC; ++b) {
                if (b0 + b < B) {
                    const float v = (red[0][b][lane] + red[1][b][lane]) + (red[2][b][lane] + red[3][b][lane]);
                    out[(size_t)(b0 + b) * P + px] = scale * v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// batch-lane variants for batches of >= 16 sinograms: lane l of a wave owns
// sinogram b0 + l % LB, and the wave's 64 / LB lane groups own consecutive
// records (scatter) or pixel slots (gather).  Every image / bins access is then
// a contiguous run of LB floats, the travel time and SIR taps of a pair are
// loaded once for its LB sinograms, and the LDS window atomics of one wave hit
// LB distinct rows (odd row pitch: distinct banks).
// ---------------------------------------------------------------------------
template <bool SIR, int LB>
__global__ __launch_bounds__(256) void k_scatter_bl(
    const int2 *__restrict__ rec, const int32_t *__restrict__ rec_off, const int32_t *__restrict__ rec_k,
    const float *__restrict__ imgI, float *__restrict__ hpart, const float *__restrict__ sir,
    const int32_t *__restrict__ swin, int Ls, BinGeom g, int Wb, int nbk, int lo_off, int WL, int WLp, int Bp,
    int B)
{
    extern __shared__ float L[];   // [LB][WLp]
    constexpr int RPW = 64 / LB;
    const int d = blockIdx.x, kb = blockIdx.y, b0 = blockIdx.z * LB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bl = lane % LB, rq = lane / LB;
    const int b = b0 + bl;
    const bool bvalid = b < B;
    const int nb = min(LB, B - b0);
    for (int i = threadIdx.x; i < LB * WLp; i += blockDim.x) L[i] = 0.0f;
    __syncthreads();
    const int base = kb * Wb + lo_off;
    const int r0 = rec_off[(size_t)d * (nbk + 1) + kb], r1 = rec_off[(size_t)d * (nbk + 1) + kb + 1];
    const int span = g.nT + 2 * g.K;
    float *Lb = L + bl * WLp;
    for (int r = r0 + wv * RPW + rq; r < r1; r += 4 * RPW) {
        const int2 e = rec[r];
        const Decoded dc = decode_tau(e.y, g);
        const int jl = 3 * dc.nn + dc.m - base;
        const float cv = bvalid ? imgI[(size_t)e.x * Bp + b] : 0.0f;
        const float u1 = dc.u, u0 = 1.0f - dc.u;
        if (!SIR) {
            if (!bvalid || (unsigned)jl >= (unsigned)(WL - 1)) continue;   // never for a consistent plan
            atomicAdd(&Lb[jl], cv * u0);
            atomicAdd(&Lb[jl + 1], cv * u1);
        } else {
            const int k = rec_k[r];
            const int s0 = swin[2 * k], len = swin[2 * k + 1];
            const float *hk = sir + (size_t)k * Ls;
            const int half = Ls / 2;
            for (int s = s0; s < s0 + len; ++s) {
                const int nn = dc.nn + s - half;
                if ((unsigned)nn >= (unsigned)span) continue;
                const int jj = jl + 3 * (s - half);
                if (!bvalid || (unsigned)jj >= (unsigned)(WL - 1)) continue;
                const float w = hk[s] * cv;
                atomicAdd(&Lb[jj], w * u0);
                atomicAdd(&Lb[jj + 1], w * u1);
            }
        }
    }
    __syncthreads();
    float *dst = hpart + (((size_t)d * nbk + kb) * WL) * Bp + b0;
    for (int i = threadIdx.x; i < WL * nb; i += blockDim.x) {
        const int jl = i / nb, bb = i - jl * nb;
        dst[(size_t)jl * Bp + bb] = L[bb * WLp + jl];
    }
}

template <bool SIR, int LB, int U>
__global__ __launch_bounds__(256) void k_gather_bl(
    const int32_t *__restrict__ tq, const int32_t *__restrict__ pix, const float *__restrict__ bins,
    float *__restrict__ out, const int32_t *__restrict__ sidx, const float *__restrict__ sir,
    const int32_t *__restrict__ swin, int Ls, BinGeom g, int Pd, int P, int D, int B, int Bp, float scale)
{
    constexpr int RPW = 64 / LB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bl = lane % LB, rq = lane / LB;
    const int slot = (blockIdx.x * 4 + wv) * RPW + rq;   // grid covers Pd exactly
    const int b = blockIdx.y * LB + bl;
    const bool bv = b < Bp;
    const int span = g.nT + 2 * g.K;
    const size_t rowstride = (size_t)g.NB * Bp;
    float acc = 0.0f, cmp = 0.0f;
    for (int d0 = 0; d0 < D; d0 += U) {
        int32_t q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + u;
            q[u] = d < D ? tq[(size_t)d * Pd + slot] : INT32_MIN;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int d = d0 + u;
            const Decoded dc = decode_tau(q[u], g);
            const float *gb = bins + (size_t)d * rowstride + b;
            if (!SIR) {
                if (bv && (unsigned)dc.nn < (unsigned)span) {
                    const int j = 3 * dc.nn + dc.m;
                    const float g0 = gb[(size_t)j * Bp], g1 = gb[(size_t)(j + 1) * Bp];
                    kahan_add(acc, cmp, fmaf(dc.u, g1 - g0, g0));
                }
            } else if (d < D) {
                const int k = sidx[(size_t)d * Pd + slot];
                const int s0 = swin[2 * k], len = swin[2 * k + 1];
                const float *hk = sir + (size_t)k * Ls;
                const int shift = dc.nn - Ls / 2;
                float part = 0.0f;
                // taps in runs of 8 whose loads are all issued before any is used; taps
                // outside the window or the bin range read nothing and add 0 (same order)
                for (int c = 0; c < len; c += 8) {
                    float wt[8], g0[8], g1[8];
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const int s = s0 + c + t, nn = shift + s;
                        const bool ok = c + t < len && (unsigned)nn < (unsigned)span;
                        const int j = ok ? 3 * nn + dc.m : 0;
                        wt[t] = ok ? hk[s] : 0.0f;
                        g0[t] = ok && bv ? gb[(size_t)j * Bp] : 0.0f;
                        g1[t] = ok && bv ? gb[(size_t)(j + 1) * Bp] : 0.0f;
                    }
#pragma unroll
                    for (int t = 0; t < 8; ++t) part = fmaf(wt[t], fmaf(dc.u, g1[t] - g0[t], g0[t]), part);
                }
                kahan_add(acc, cmp, part);
            }
        }
    }
    const int px = pix[slot];
    if (px >= 0 && b < B) out[(size_t)b * P + px] = scale * (acc - cmp);
}

// ---------------------------------------------------------------------------
// forward without SIR, gather form ("pull"): pairs sorted by bin (CSR), so
// bins[d][j][b] = sum_{pairs of bin j} (1 - u) c[px][b] + sum_{pairs of bin j-1} u c[px][b]
// is computed without atomics, each pair read once, in a fixed order.
// k_fwd_pull: one thread per bin (256 consecutive bins of one detector per
// workgroup), BC sinograms per pass; the u-part of bin j - 1 comes from the
// neighbouring thread through LDS (thread 0 computes it for the first bin).
// ---------------------------------------------------------------------------
template <int BC>
__device__ __forceinline__ void load_c(const float *p, float (&v)[BC])
{
    if constexpr (BC % 4 == 0) {
#pragma unroll
        for (int q = 0; q < BC; q += 4) {
            const float4 t = *reinterpret_cast<const float4 *>(p + q);
            v[q] = t.x; v[q + 1] = t.y; v[q + 2] = t.z; v[q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < BC; ++q) v[q] = p[q];
    }
}

template <int BC>
__global__ __launch_bounds__(256) void k_fwd_pull(const int2 *__restrict__ rec, const int32_t *__restrict__ off,
                                                  const float *__restrict__ imgI, float *__restrict__ bins, int NB,
                                                  int Bp)
{
    __shared__ float s1s[BC][257];
    const int d = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int b0 = blockIdx.z * BC;
    const int32_t *o = off + (size_t)d * (NB + 1);
    float s0[BC], s1[BC], c[BC];
#pragma unroll
    for (int q = 0; q < BC; ++q) { s0[q] = 0.0f; s1[q] = 0.0f; }
    (void)c;
    if (j < NB) {
        const int re = o[j + 1];
        constexpr int U = BC <= 4 ? 4 : (BC <= 8 ? 2 : 1);   // pairs in flight per thread
        for (int r = o[j]; r < re; r += U) {
            int2 e[U];
            float cu[U][BC];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = r + u < re ? rec[r + u] : make_int2(0, 0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (r + u < re) {
                    load_c<BC>(imgI + (size_t)e[u].x * Bp + b0, cu[u]);
                } else {
#pragma unroll
                    for (int q = 0; q < BC; ++q) cu[u][q] = 0.0f;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float uu = r + u < re ? __int_as_float(e[u].y) : 0.0f;
#pragma unroll
                for (int q = 0; q < BC; ++q) {
                    s0[q] = fmaf(1.0f - uu, cu[u][q], s0[q]);
                    s1[q] = fmaf(uu, cu[u][q], s1[q]);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < BC; ++q) s1s[q][threadIdx.x + 1] = s1[q];
    if (threadIdx.x == 0) {
        float p1[BC];
#pragma unroll
        for (int q = 0; q < BC; ++q) p1[q] = 0.0f;
        if (j > 0 && j - 1 < NB) {
            const int re = o[j];
            for (int r = o[j - 1]; r < re; ++r) {
                const int2 e = rec[r];
                const float u = __int_as_float(e.y);
                load_c<BC>(imgI + (size_t)e.x * Bp + b0, c);
#pragma unroll
                for (int q = 0; q < BC; ++q) p1[q] = fmaf(u, c[q], p1[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < BC; ++q) s1s[q][0] = p1[q];
    }
    __syncthreads();
    if (j < NB) {
        float *dst = bins + ((size_t)d * NB + j) * Bp + b0;
#pragma unroll
        for (int q = 0; q < BC; ++q) dst[q] = s0[q] + s1s[q][threadIdx.x];
    }
}

// batch-lane form: lane l owns sinogram b0 + l % LB; each lane group walks JR
// consecutive bins of one detector, carrying the u-part into the next bin.  The
// pairs are consumed in runs of U whose records and image rows are all loaded
// before any is accumulated, so U loads are in flight per lane.
template <int LB, int U>
__global__ __launch_bounds__(256) void k_fwd_pull_bl(const int2 *__restrict__ rec, const int32_t *__restrict__ off,
                                                     const float *__restrict__ imgI, float *__restrict__ bins, int NB,
                                                     int Bp, int JR, int xcd_map /* D / 8, or 0 */)
{
    constexpr int RPW = 64 / LB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bl = lane % LB, rq = lane / LB;
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so XCD x
    // (= blockIdx.x % 8) is given the detectors [x * D/8, (x + 1) * D/8), walked
    // detector-fastest: its concurrent workgroups read the same arc of image rows
    // (neighbouring detectors, same bins) out of that XCD's L2.
    int d, jblk;
    if (xcd_map) {
        const int dpx = xcd_map;                   // detectors per XCD
        const int x = blockIdx.x % 8, sidx = blockIdx.x / 8;
        d = x * dpx + sidx % dpx;
        jblk = sidx / dpx;
    } else {
        d = blockIdx.y;
        jblk = blockIdx.x;
    }
    const int ja = ((jblk * 4 + wv) * RPW + rq) * JR;
    if (ja >= NB) return;   // no barriers in this kernel
    const int jb = min(NB, ja + JR);
    const int b = blockIdx.z * LB + bl;
    const bool bv = b < Bp;
    const int32_t *o = off + (size_t)d * (NB + 1);
    const float *cb = imgI + b;
    float *dst = bins + (size_t)d * NB * Bp + b;
    float prev = 0.0f;
    if (ja > 0) {
        const int re = o[ja];
        for (int r = o[ja - 1]; r < re; ++r) {
            const int2 e = rec[r];
            const float cv = bv ? cb[(size_t)e.x * Bp] : 0.0f;
            prev = fmaf(__int_as_float(e.y), cv, prev);
        }
    }
    int r = o[ja];
    const int rend = o[jb];
    int j = ja;
    int jend = o[ja + 1];   // first record past bin j
    float s0 = 0.0f, s1 = 0.0f;
    while (r < rend) {
        int2 e[U];
        float cv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = r + u < rend ? rec[r + u] : make_int2(0, 0);
#pragma unroll
        for (int u = 0; u < U; ++u) cv[u] = (bv && r + u < rend) ? cb[(size_t)e[u].x * Bp] : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (r + u >= rend) break;
            while (r + u >= jend) {   // bin j is complete
                if (bv) dst[(size_t)j * Bp] = s0 + prev;
                prev = s1;
                s0 = s1 = 0.0f;
                ++j;
                jend = o[j + 1];
            }
            const float uu = __int_as_float(e[u].y);
            s0 = fmaf(1.0f - uu, cv[u], s0);
            s1 = fmaf(uu, cv[u], s1);
        }
        r += U;
    }
    for (; j < jb; ++j) {   // the last bin with pairs and any empty bins after it
        if (bv) dst[(size_t)j * Bp] = s0 + prev;
        prev = s1;
        s0 = s1 = 0.0f;
    }
}

// batch-lane form with lane groups of 64 sinograms: each wave owns JR bins of
// one detector, so bin offsets and records are uniform over the wave and live
// in scalar registers.  Per bin, its records are consumed in runs of U whose
// image rows are all loaded before any is accumulated (records past the bin's
// end read the zero padding after the table with weight 0).
template <int U>
__global__ __launch_bounds__(256) void k_fwd_pull_wu(const int2 *__restrict__ rec, const int32_t *__restrict__ off,
                                                     const float *__restrict__ imgI, float *__restrict__ bins, int NB,
                                                     int Bp, int JR, int xcd_map /* D / 8, or 0 */)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int d, jblk;
    if (xcd_map) {   // XCD x (= blockIdx.x % 8) takes detectors [x * D/8, (x + 1) * D/8)
        const int x = blockIdx.x % 8, sidx = blockIdx.x / 8;
        d = x * xcd_map + sidx % xcd_map;
        jblk = sidx / xcd_map;
    } else {
        d = blockIdx.y;
        jblk = blockIdx.x;
    }
    const int ja = __builtin_amdgcn_readfirstlane((jblk * 4 + wv) * JR);
    if (ja >= NB) return;   // no barriers in this kernel
    const int jb = min(NB, ja + JR);
    const int b = blockIdx.z * 64 + lane;
    const int bo = min(b, Bp - 1);   // lanes past Bp compute a valid column and store nothing
    const int32_t *o = off + (size_t)d * (NB + 1);
    const float *cb = imgI + bo;
    float *dst = bins + (size_t)d * NB * Bp + b;
    float prev = 0.0f;
    if (ja > 0)   // u-part of bin ja - 1
        for (int r = o[ja - 1], re = o[ja]; r < re; ++r) {
            const int2 e = rec[r];
            prev = fmaf(__int_as_float(e.y), cb[(size_t)e.x * Bp], prev);
        }
    int r0 = o[ja];
    for (int j = ja; j < jb; ++j) {
        const int r1 = o[j + 1];
        float s0 = 0.0f, s1 = 0.0f;
        for (int r = r0; r < r1; r += U) {
            int2 e[U];
            float cv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = rec[r + u];
#pragma unroll
            for (int u = 0; u < U; ++u) cv[u] = cb[(size_t)e[u].x * Bp];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool in = r + u < r1;
                const float uu = in ? __int_as_float(e[u].y) : 0.0f;
                s0 = fmaf(in ? 1.0f - uu : 0.0f, cv[u], s0);
                s1 = fmaf(uu, cv[u], s1);
            }
        }
        if (b < Bp) dst[(size_t)j * Bp] = s0 + prev;
        prev = s1;
        r0 = r1;
    }
}

// Wave-uniform pull with raw-buffer image-row loads: records are read with
// scalar loads, each image row at a scalar byte offset (no per-pair address
// arithmetic in vector registers); full runs of U records carry no tests and
// the tail of a bin runs once more with weights zeroed past its end.  The image
// is read as a raw buffer with 32-bit offsets: P * Bp floats below 2^31 bytes
// (checked by the launcher).
template <int U>
__global__ __launch_bounds__(256) void k_fwd_pull_wt(const int2 *__restrict__ rec, const int32_t *__restrict__ off,
                                                     const float *__restrict__ imgI, float *__restrict__ bins, int NB,
                                                     int Bp, int JR, int xcd_map /* D / 8, or 0 */, int img_bytes)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int d, jblk;
    if (xcd_map) {   // XCD x (= blockIdx.x % 8) takes detectors [x * D/8, (x + 1) * D/8)
        const int x = blockIdx.x % 8, sidx = blockIdx.x / 8;
        d = x * xcd_map + sidx % xcd_map;
        jblk = sidx / xcd_map;
    } else {
        d = blockIdx.y;
        jblk = blockIdx.x;
    }
    const int ja = __builtin_amdgcn_readfirstlane((jblk * 4 + wv) * JR);
    if (ja >= NB) return;   // no barriers in this kernel
    const int jb = min(NB, ja + JR);
    const int b = blockIdx.z * 64 + lane;
    const int vo = 4 * min(b, Bp - 1);   // lanes past Bp compute a valid column and store nothing
    const int rowb = 4 * Bp;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(imgI), (short)0, img_bytes, kBufferWord3);
    const int32_t *o = off + (size_t)d * (NB + 1);
    float *dst = bins + (size_t)d * NB * Bp + b;
    auto ld = [&](int px) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, px * rowb, 0));
    };
    float prev = 0.0f;
    if (ja > 0)   // u-part of bin ja - 1
        for (int r = o[ja - 1], re = o[ja]; r < re; ++r) {
            const int2 e = rec[r];
            prev = fmaf(__int_as_float(e.y), ld(e.x), prev);
        }
    int r0 = o[ja];
    for (int j = ja; j < jb; ++j) {
        const int r1 = o[j + 1];
        float s0 = 0.0f, s1 = 0.0f;
        int r = r0;
        for (; r + U <= r1; r += U) {
            int2 e[U];
            float cv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = rec[r + u];
#pragma unroll
            for (int u = 0; u < U; ++u) cv[u] = ld(e[u].x);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float uu = __int_as_float(e[u].y);
                s0 = fmaf(1.0f - uu, cv[u], s0);
                s1 = fmaf(uu, cv[u], s1);
            }
        }
        if (r < r1) {   // tail: records past r1 (next bin's, or the zero pad) get weight 0
            int2 e[U];
            float cv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = rec[r + u];
#pragma unroll
            for (int u = 0; u < U; ++u) cv[u] = ld(e[u].x);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool in = r + u < r1;
                const float uu = in ? __int_as_float(e[u].y) : 0.0f;
                s0 = fmaf(in ? 1.0f - uu : 0.0f, cv[u], s0);
                s1 = fmaf(uu, cv[u], s1);
            }
        }
        if (b < Bp) dst[(size_t)j * Bp] = s0 + prev;
        prev = s1;
        r0 = r1;
    }
}

// forward pull for a single sinogram (Bp == 1), record-streaming form: a wave
// owns 64 consecutive bins of one detector (lane i: bin ja + i) and streams the
// records of bins ja - 1 .. ja + 63 (contiguous in the CSR table) in coalesced
// chunks of 64, lane l taking record r + l.  The per-record terms (1 - u) c and
// u c are prefix-summed over the wave; lane i then takes the sum of its
// bin's run of records (the (1 - u) parts) and of the previous bin's run (the u
// parts) as differences of the prefix sums at the run ends.
// v plus its DPP-moved copy; lanes without a source (or rows outside ROWS) add 0
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS,
                                                                      0xf, false));
}

// inclusive prefix sum over the 64 lanes: shifts by 1, 2, 4, 8 within each row of
// 16 lanes, then rows 1 and 3 add the last lane of the row before (row_bcast:15),
// rows 2 and 3 add lane 31 (row_bcast:31)
__device__ __forceinline__ float wave_scan_add(float v)
{
    v = dpp_add<0x111, 0xf>(v);   // row_shr:1
    v = dpp_add<0x112, 0xf>(v);   // row_shr:2
    v = dpp_add<0x114, 0xf>(v);   // row_shr:4
    v = dpp_add<0x118, 0xf>(v);   // row_shr:8
    v = dpp_add<0x142, 0xa>(v);   // row_bcast:15 into rows 1, 3
    v = dpp_add<0x143, 0xc>(v);   // row_bcast:31 into rows 2, 3
    return v;
}

// inclusive prefix s (over the wave) summed over lanes [lo, hi) of the chunk
__device__ __forceinline__ float run_sum(float s, int lo, int hi)
{
    const float a = __shfl(s, max(hi - 1, 0), 64), b = __shfl(s, max(lo - 1, 0), 64);
    return hi > lo ? a - (lo > 0 ? b : 0.0f) : 0.0f;
}

template <int NG>
__global__ __launch_bounds__(256) void k_fwd_pull_b1(const int2 *__restrict__ rec, const int32_t *__restrict__ off,
                                                     const float *__restrict__ img, float *__restrict__ bins, int NB,
                                                     int xcd_map /* D / 8, or 0 */)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int d, jblk;
    if (xcd_map) {   // XCD x (= blockIdx.x % 8) takes detectors [x * D/8, (x + 1) * D/8)
        const int x = blockIdx.x % 8, sidx = blockIdx.x / 8;
        d = x * xcd_map + sidx % xcd_map;
        jblk = sidx / xcd_map;
    } else {
        d = blockIdx.y;
        jblk = blockIdx.x;
    }
    // the wave owns NG groups of 64 bins, [ja, ja + 64 NG); lane i of group g: bin ja + 64 g + i
    const int ja = (jblk * 4 + wv) * 64 * NG;
    if (ja >= NB) return;   // no barriers in this kernel
    const int32_t *o = off + (size_t)d * (NB + 1);
    // record runs: bin J is [o[J], o[J + 1]); bin J - 1 is [o[J - 1], o[J])
    int oJ[NG], oJ1[NG], oJm[NG], gs[NG], ge[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int J = ja + 64 * g + lane;
        oJ[g] = o[min(J, NB)];
        oJ1[g] = o[min(J + 1, NB)];
        oJm[g] = J > 0 ? o[min(J - 1, NB)] : oJ[g];
        const int j0 = ja + 64 * g;   // the group's records: bins j0 - 1 .. j0 + 63
        gs[g] = __builtin_amdgcn_readfirstlane(o[min(max(j0 - 1, 0), NB)]);
        ge[g] = __builtin_amdgcn_readfirstlane(o[min(j0 + 64, NB)]);
    }
    const int rs = gs[0], re = ge[NG - 1];
    float a0[NG], a1[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) a0[g] = a1[g] = 0.0f;
    // NC chunks per step, software-pipelined: the records of step k + 2 and the
    // pixel values of step k + 1 are in flight while step k is summed (records past
    // re, the next bins' or the zero pad, carry weight 0)
    constexpr int NC = 8;
    int2 e[NC], en[NC];
    float c[NC];
    auto ld_e = [&](int r0, int2(&q)[NC]) {
#pragma unroll
        for (int k = 0; k < NC; ++k) q[k] = rec[r0 + 64 * k + lane];
    };
    auto ld_c = [&](const int2(&q)[NC], float(&v)[NC]) {
#pragma unroll
        for (int k = 0; k < NC; ++k) v[k] = img[q[k].x];
    };
    int r0 = rs;
    if (r0 < re) {
        ld_e(r0, e);
        ld_c(e, c);
        if (r0 + 64 * NC < re) ld_e(r0 + 64 * NC, en);
    }
    for (; r0 < re; r0 += 64 * NC) {
        float cn[NC];
        int2 enn[NC];
        if (r0 + 64 * NC < re) ld_c(en, cn);
        if (r0 + 128 * NC < re) ld_e(r0 + 128 * NC, enn);
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const int r = r0 + 64 * q;
            if (r >= re) break;
            const bool in = r + lane < re;
            const float uq = __int_as_float(e[q].y);
            const float w0 = in ? (1.0f - uq) * c[q] : 0.0f, w1 = in ? uq * c[q] : 0.0f;
            const float s0 = wave_scan_add(w0), s1 = wave_scan_add(w1);
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                if (r >= ge[g] || r + 64 <= gs[g]) continue;   // the chunk misses the group's records
                a0[g] += run_sum(s0, min(max(oJ[g] - r, 0), 64), min(max(oJ1[g] - r, 0), 64));
                a1[g] += run_sum(s1, min(max(oJm[g] - r, 0), 64), min(max(oJ[g] - r, 0), 64));
            }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            e[q] = en[q];
            c[q] = cn[q];
            en[q] = enn[q];
        }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int J = ja + 64 * g + lane;
        if (J < NB) bins[(size_t)d * NB + J] = a0[g] + a1[g];
    }
}

// SIR forward for a single sinogram, pull form (records of build_sir_pull): a wave owns
// samples a - 1 .. a + 14 of one detector in four lane groups of 16 (lane i of a group:
// sample a - 1 + i).  Group g takes every fourth record of the records whose taps can
// reach those samples; per record each lane adds the tap landing on its sample (if any)
// to its phase accumulators (u-part of phase 2 into a carry for the next sample's phase
// 0).  No atomics: the four groups are added at the end, and lane 0 (sample a - 1) only
// hands its carry on.  Same terms as k_scatter_bk (c (h (1 - u)), c (h u)), summed in
// record order instead of atomic order.
__global__ __launch_bounds__(256) void k_fwd_sir_pull1(const int4 *__restrict__ rec, const int32_t *__restrict__ off,
                                                       const float *__restrict__ img, const float *__restrict__ sir,
                                                       float *__restrict__ bins, int NB, int span, int lo, int n_nf,
                                                       int maxlen)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, i = lane & 15;
    const int d = blockIdx.y;
    const int a = (blockIdx.x * 4 + wv) * 15;   // first output sample of the wave
    if (a > span) return;                       // samples 0 .. span (span: only bin 3 span = NB - 1); no barriers
    const int n = a - 1 + i;
    const bool live = n >= 0 && n < span;       // taps land on samples in [0, span) only
    // records with nf in [a - maxlen, a + 14] can reach samples a - 1 .. a + 14
    const int ia = min(max(a - maxlen - lo, 0), n_nf), ib = min(max(a + 15 - lo, 0), n_nf);
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, carry = 0.0f;
    constexpr int U = 4;   // records per group whose loads are in flight together
    for (int mp = 0; mp < 3; ++mp) {   // the record lists of the three phases
        const int32_t *o = off + ((size_t)d * 3 + mp) * (n_nf + 1);
        const int rs = o[ia], re = o[ib];
        for (int r0 = rs + g; r0 < re; r0 += 4 * U) {
            int4 e[U];
            float c[U], h[U];
#pragma unroll
            for (int q = 0; q < U; ++q) e[q] = rec[min(r0 + 4 * q, re - 1)];   // clamped: duplicates are masked below
#pragma unroll
            for (int q = 0; q < U; ++q) {
                c[q] = img[e[q].x & 0x0fffffff];
                const int t = n - ((e[q].y << 16) >> 16), len = (int)((unsigned)e[q].y >> 16);
                const bool ok = live && r0 + 4 * q < re && t >= 0 && t < len;
                h[q] = ok ? sir[e[q].w + t] : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const int m = (int)((unsigned)e[q].x >> 28);
                const float u = __int_as_float(e[q].z);
                const float w0 = c[q] * (h[q] * (1.0f - u)), w1 = c[q] * (h[q] * u);
                acc0 += m == 0 ? w0 : 0.0f;
                acc1 += m == 0 ? w1 : (m == 1 ? w0 : 0.0f);
                acc2 += m == 1 ? w1 : (m == 2 ? w0 : 0.0f);
                carry += m == 2 ? w1 : 0.0f;
            }
        }
    }
    // the four groups' sums (lanes i, i + 16, i + 32, i + 48), then the carry of sample n - 1
    acc0 += __shfl_xor(acc0, 32, 64);
    acc1 += __shfl_xor(acc1, 32, 64);
    acc2 += __shfl_xor(acc2, 32, 64);
    carry += __shfl_xor(carry, 32, 64);
    acc0 += __shfl_xor(acc0, 16, 64);
    acc1 += __shfl_xor(acc1, 16, 64);
    acc2 += __shfl_xor(acc2, 16, 64);
    carry += __shfl_xor(carry, 16, 64);
    const float cin = __shfl_up(carry, 1, 64);
    if (g != 0 || i == 0 || n > span) return;
    float *row = bins + (size_t)d * NB;
    if (3 * n < NB) row[3 * n] = acc0 + cin;
    if (n < span) {
        row[3 * n + 1] = acc1;
        row[3 * n + 2] = acc2;
    }
}

// SIR forward for batches (lane = sinogram), pull form: a wave owns TS samples
// a .. a + TS - 1 of one detector for 64 sinograms and keeps their 3 TS + 1 bins
// (the last one: phase 0 of sample a + TS, never stored) in registers.  Records whose
// taps reach samples a - 1 .. a + TS - 1 are read as wave-uniform scalars, one phase
// at a time, with their TS + 1 weights at those samples: the weights table holds every
// window between kSirWinPad zeros on each side, so a record needs no per-tap test.
// Taps on samples outside [0, span) are dropped where they would reach a stored bin:
// the phase-2 u-part from sample -1 (bin 0) and the phase-0 part at sample span (bin
// 3 span, which only takes the u-part of sample span - 1).
template <int TS, int U>
__device__ __forceinline__ void sir_pull_phase(float (&acc)[3 * TS + 1], const int4 *__restrict__ rec, int rs, int re,
                                               const float *__restrict__ imgI, const float *__restrict__ sirw,
                                               int a, int maxlen, int Bp, int bo, int m, float f0)
{
    for (int r0 = rs; r0 < re; r0 += U) {
        int4 e[U];
        float c[U];
#pragma unroll
        for (int q = 0; q < U; ++q) e[q] = rec[r0 + q];   // the table is padded past its end
#pragma unroll
        for (int q = 0; q < U; ++q) c[q] = imgI[(size_t)(e[q].x & 0x0fffffff) * Bp + bo];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            if (r0 + q >= re) break;
            // weight j is tap a - 1 + j - nf of the window: offset a - 1 - nf in [-TS, maxlen - 1]
            const int sh = min(max(a - 1 - ((e[q].y << 16) >> 16), -TS), maxlen - 1);
            const float *w = sirw + (size_t)e[q].w + sh;
            float h[TS + 1];
#pragma unroll
            for (int j = 0; j <= TS; ++j) h[j] = w[j];
            const float u = __int_as_float(e[q].z);
            const float c0 = c[q] * (1.0f - u), c1 = c[q] * u;
            if (m == 0) {
#pragma unroll
                for (int t = 0; t < TS; ++t) {
                    acc[3 * t] += c0 * h[t + 1];
                    acc[3 * t + 1] += c1 * h[t + 1];
                }
            } else if (m == 1) {
#pragma unroll
                for (int t = 0; t < TS; ++t) {
                    acc[3 * t + 1] += c0 * h[t + 1];
                    acc[3 * t + 2] += c1 * h[t + 1];
                }
            } else {
                acc[0] += (c1 * f0) * h[0];
#pragma unroll
                for (int t = 0; t < TS; ++t) {
                    acc[3 * t + 2] += c0 * h[t + 1];
                    acc[3 * t + 3] += c1 * h[t + 1];
                }
            }
        }
    }
}

template <int TS, int U>
__global__ __launch_bounds__(256) void k_fwd_sir_pull_bl(const int4 *__restrict__ rec,
                                                         const int32_t *__restrict__ off,
                                                         const float *__restrict__ imgI, const float *__restrict__ sirw,
                                                         float *__restrict__ bins, int NB, int span, int lo, int n_nf,
                                                         int maxlen, int Bp)
{
    static_assert(TS <= kSirWinPad, "window margins must cover the samples of a wave");
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int d = blockIdx.y;
    const int a = __builtin_amdgcn_readfirstlane((blockIdx.x * 4 + wv) * TS);
    if (a > span) return;   // samples 0 .. span; no barriers in this kernel
    const int b = blockIdx.z * 64 + lane;
    const int bo = min(b, Bp - 1);
    // records with nf in [a - maxlen, a + TS - 1] can reach samples a - 1 .. a + TS - 1
    const int ia = min(max(a - maxlen - lo, 0), n_nf), ib = min(max(a + TS - lo, 0), n_nf);
    float acc[3 * TS + 1];
#pragma unroll
    for (int k = 0; k < 3 * TS + 1; ++k) acc[k] = 0.0f;
    const int32_t *o = off + (size_t)d * 3 * (n_nf + 1);
    sir_pull_phase<TS, U>(acc, rec, o[ia], o[ib], imgI, sirw, a, maxlen, Bp, bo, 0, 1.0f);
    // bin 3 span keeps only the u-part of sample span - 1 (added below)
#pragma unroll
    for (int t = 0; t < TS; ++t)
        if (a + t == span) acc[3 * t] = 0.0f;
    o += n_nf + 1;
    sir_pull_phase<TS, U>(acc, rec, o[ia], o[ib], imgI, sirw, a, maxlen, Bp, bo, 1, 1.0f);
    o += n_nf + 1;
    sir_pull_phase<TS, U>(acc, rec, o[ia], o[ib], imgI, sirw, a, maxlen, Bp, bo, 2, a > 0 ? 1.0f : 0.0f);
    if (b >= Bp) return;
    float *dst = bins + (size_t)d * NB * Bp + b;
#pragma unroll
    for (int t = 0; t < TS; ++t) {
        const int n = a + t;
        if (n > span) break;
        if (3 * n < NB) dst[(size_t)(3 * n) * Bp] = acc[3 * t];
        if (n < span) {
            dst[(size_t)(3 * n + 1) * Bp] = acc[3 * t + 1];
            dst[(size_t)(3 * n + 2) * Bp] = acc[3 * t + 2];
        }
    }
}

// SIR forward for batches, band form, pass 1 (lane = sinogram): wave w takes the records
// whose first affected sample nf lies in its band [a, a + nfb), a = lo + w nfb, so every
// record is visited once; its taps land on samples a .. a + TS - 1 (TS >= nfb - 1 + maxlen)
// and the wave keeps their 3 TS + 1 bins in registers.  A record's TS weights from sample
// a on come from the zero-margin weights table (offset nf - a before its first tap).  The
// bin of sample -1's phase-2 u-part and the phase-0 part at sample span are dropped
// where they fall in the wave's range (the reference's range check).  The wave's rows go
// to its slot of the band scratch; pass 2 adds the overlapping slots.
template <int TS>
__device__ __forceinline__ void sir_band_phase(float (&acc)[3 * TS + 1], const int4 *__restrict__ rec, int rs,
                                               int re, const float *__restrict__ imgI,
                                               const float *__restrict__ sirw, int a, int Bp, int bo, int m)
{
    for (int r = rs; r < re; ++r) {
        const int4 e = rec[r];
        const float c = imgI[(size_t)(e.x & 0x0fffffff) * Bp + bo];
        const int t0 = ((e.y << 16) >> 16) - a;   // in [0, nfb)
        const float *w = sirw + (size_t)e.w - t0;   // w[t]: weight of the tap on sample a + t
        float h[TS];
#pragma unroll
        for (int t = 0; t < TS; ++t) h[t] = w[t];
        const float u = __int_as_float(e.z);
        const float c0 = c * (1.0f - u), c1 = c * u;
#pragma unroll
        for (int t = 0; t < TS; ++t) {
            acc[3 * t + m] += c0 * h[t];
            acc[3 * t + m + 1] += c1 * h[t];
        }
    }
}

template <int TS>
__global__ __launch_bounds__(256) void k_fwd_sir_band(const int4 *__restrict__ rec, const int32_t *__restrict__ off,
                                                      const float *__restrict__ imgI, const float *__restrict__ sirw,
                                                      float *__restrict__ scratch, int span, int lo, int n_nf,
                                                      int nfb, int nwave, int Bp)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= nwave) return;   // no barriers in this kernel
    const int d = blockIdx.y;
    const int b = blockIdx.z * 64 + lane;
    const int bo = min(b, Bp - 1);
    const int a = lo + w * nfb;
    const int ia = w * nfb, ib = min(ia + nfb, n_nf);
    float acc[3 * TS + 1];
#pragma unroll
    for (int k = 0; k < 3 * TS + 1; ++k) acc[k] = 0.0f;
    const int32_t *o = off + (size_t)d * 3 * (n_nf + 1);
    // phase 2 first: the u-part of sample -1 (row 0) is dropped before other phases add there
    sir_band_phase<TS>(acc, rec, o[2 * (n_nf + 1) + ia], o[2 * (n_nf + 1) + ib], imgI, sirw, a, Bp, bo, 2);
#pragma unroll
    for (int t = 0; t < TS; ++t)
        if (a + t == -1) acc[3 * t + 3] = 0.0f;
    // phase 0: the part at sample span (row 3 span) is dropped, the u-part of span - 1 kept
    float keep = 0.0f;
#pragma unroll
    for (int t = 0; t < TS; ++t)
        if (a + t == span) keep = acc[3 * t];
    sir_band_phase<TS>(acc, rec, o[ia], o[ib], imgI, sirw, a, Bp, bo, 0);
#pragma unroll
    for (int t = 0; t < TS; ++t)
        if (a + t == span) acc[3 * t] = keep;
    sir_band_phase<TS>(acc, rec, o[(n_nf + 1) + ia], o[(n_nf + 1) + ib], imgI, sirw, a, Bp, bo, 1);
    if (b >= Bp) return;
    float *dst = scratch + (((size_t)d * nwave + w) * (3 * TS + 1)) * Bp + b;
#pragma unroll
    for (int k = 0; k < 3 * TS + 1; ++k) dst[(size_t)k * Bp] = acc[k];
}

// pass 2: bins[d][j][b] = sum over the waves whose rows [3 a_w, 3 (a_w + TS)] hold j, in
// wave order (rows of samples outside [0, span] are never asked for)
template <int TS>
__global__ __launch_bounds__(256) void k_fwd_sir_band_sum(const float *__restrict__ scratch,
                                                          float *__restrict__ bins, int NB, int lo, int nfb,
                                                          int nwave, int Bp)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;   // (j, b) of detector blockIdx.y
    if (i >= (size_t)NB * Bp) return;
    const int d = blockIdx.y;
    const int j = (int)(i / Bp), b = (int)(i - (size_t)j * Bp);
    // wave w covers rows 3 (lo + w nfb) .. 3 (lo + w nfb + TS)
    const int wlo = max(0, (j - 3 * (lo + TS) + 3 * nfb - 1) / (3 * nfb)), whi = min(nwave - 1, (j - 3 * lo) / (3 * nfb));
    // up to 8 rows in flight per step, added in wave order (absent rows add 0)
    float acc = 0.0f;
    for (int w0 = wlo; w0 <= whi; w0 += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int w = w0 + k, r = j - 3 * (lo + w * nfb);
            v[k] = (w <= whi && r >= 0 && r <= 3 * TS)
                       ? scratch[(((size_t)d * nwave + w) * (3 * TS + 1) + r) * Bp + b]
                       : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    bins[((size_t)d * NB + j) * Bp + b] = acc;
}

// ---------------------------------------------------------------------------
// tile gather (batch-lane, Bp >= 16): one workgroup per 16 x 16 pixel tile and
// chunk of LB sinograms.  For every detector the tile's bin window (all bins its
// pairs read, at most a few hundred) is staged in LDS with 16-byte loads, the
// tile's travel times next to it; each wave then interpolates 64 pixels from LDS
// (lane l: sinogram b0 + l % LB, its lane group a run of LB pixels), keeping
// Kahan-compensated sums in registers across all detectors.  Every bin row is
// read once per (tile, detector) instead of twice per pair.
// ---------------------------------------------------------------------------
template <int LB>
__global__ __launch_bounds__(256) void k_gather_tile(
    const int32_t *__restrict__ tq, const int2 *__restrict__ twin, const int32_t *__restrict__ pix,
    const float *__restrict__ bins, float *__restrict__ out, BinGeom g, int Pd, int P, int D, int B, int Bp,
    int WMAX, float scale)
{
    extern __shared__ float sm[];                 // [WMAX][LB] window, then [256] travel times
    float *W = sm;
    int32_t *T = reinterpret_cast<int32_t *>(sm + (size_t)WMAX * LB);
    constexpr int NG = 64 / LB;                   // lane groups per wave
    constexpr int PPG = 64 / NG;                  // pixels per lane group
    const int tile = blockIdx.x, b0 = blockIdx.y * LB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bl = lane % LB, gi = lane / LB;
    const int pbase = wv * 64 + gi * PPG;
    const int span = g.nT + 2 * g.K;
    const size_t rowstride = (size_t)g.NB * Bp;
    const int c4n = LB / 4;                       // float4 per window row
    float acc[PPG], cmp[PPG];
#pragma unroll
    for (int k = 0; k < PPG; ++k) { acc[k] = 0.0f; cmp[k] = 0.0f; }
    for (int d = 0; d < D; ++d) {
        const int2 wd = twin[(size_t)tile * D + d];
        __syncthreads();                          // previous detector's window no longer read
        T[threadIdx.x] = tq[(size_t)d * Pd + (size_t)tile * 256 + threadIdx.x];
        const float *src = bins + (size_t)d * rowstride + (size_t)wd.x * Bp + b0;
        for (int i = threadIdx.x; i < wd.y * c4n; i += blockDim.x) {
            const int row = i / c4n, c4 = i - row * c4n;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (b0 + 4 * c4 < Bp) v = *reinterpret_cast<const float4 *>(src + (size_t)row * Bp + 4 * c4);
            *reinterpret_cast<float4 *>(W + row * LB + 4 * c4) = v;
        }
        __syncthreads();
        if (wd.y == 0) continue;
#pragma unroll
        for (int k = 0; k < PPG; ++k) {
            const Decoded dc = decode_tau(T[pbase + k], g);
            if ((unsigned)dc.nn < (unsigned)span) {
                const int jl = 3 * dc.nn + dc.m - wd.x;
                const float g0 = W[jl * LB + bl], g1 = W[(jl + 1) * LB + bl];
                kahan_add(acc[k], cmp[k], fmaf(dc.u, g1 - g0, g0));
            }
        }
    }
    const int b = b0 + bl;
    if (b >= B) return;
#pragma unroll
    for (int k = 0; k < PPG; ++k) {
        const int px = pix[(size_t)tile * 256 + pbase + k];
        if (px >= 0) out[(size_t)b * P + px] = scale * (acc[k] - cmp[k]);
    }
}

// Pipelined tile gather: while the waves interpolate detector d from LDS, every
// thread already holds its share of detector d+1's bin window (up to PF 16-byte
// loads) and travel time in registers; each pixel's (bin, weight) is decoded
// once per detector by one thread and shared through LDS.
template <int LB, int PF>
__global__ __launch_bounds__(256) void k_gather_tile2(
    const int32_t *__restrict__ tq, const int2 *__restrict__ twin, const int32_t *__restrict__ pix,
    const float *__restrict__ bins, float *__restrict__ out, BinGeom g, int Pd, int P, int D, int B, int Bp,
    int WMAX, float scale)
{
    extern __shared__ float sm[];                 // [WMAX][LB] window, [256] bin, [256] weight
    float *W = sm;
    int32_t *J = reinterpret_cast<int32_t *>(sm + (size_t)WMAX * LB);
    float *Uw = sm + (size_t)WMAX * LB + 256;
    constexpr int NG = 64 / LB, PPG = 64 / NG, C4N = LB / 4;
    const int tile = blockIdx.x, b0 = blockIdx.y * LB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bl = lane % LB, gi = lane / LB;
    const int pbase = wv * 64 + gi * PPG;
    const int span = g.nT + 2 * g.K;
    const size_t rowstride = (size_t)g.NB * Bp;
    float acc[PPG], cmp[PPG];
#pragma unroll
    for (int k = 0; k < PPG; ++k) { acc[k] = 0.0f; cmp[k] = 0.0f; }
    float4 pf[PF];
    int32_t qn;
    int2 wn;
    auto prefetch = [&](int d) {
        wn = twin[(size_t)tile * D + d];
        qn = tq[(size_t)d * Pd + (size_t)tile * 256 + threadIdx.x];
        const float *src = bins + (size_t)d * rowstride + (size_t)wn.x * Bp + b0;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int i = threadIdx.x + k * 256;
            const int row = i / C4N, c4 = i - row * C4N;
            pf[k] = (row < wn.y && b0 + 4 * c4 < Bp)
                        ? *reinterpret_cast<const float4 *>(src + (size_t)row * Bp + 4 * c4)
                        : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    prefetch(0);
    for (int d = 0; d < D; ++d) {
        const int2 wd = wn;
        __syncthreads();                          // the previous detector is no longer read
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int i = threadIdx.x + k * 256;
            if (i < wd.y * C4N) *reinterpret_cast<float4 *>(W + 4 * i) = pf[k];
        }
        {
            const Decoded dc = decode_tau(qn, g);
            const bool ok = (unsigned)dc.nn < (unsigned)span;
            J[threadIdx.x] = ok ? 3 * dc.nn + dc.m - wd.x : -1;
            Uw[threadIdx.x] = dc.u;
        }
        __syncthreads();
        if (d + 1 < D) prefetch(d + 1);
#pragma unroll
        for (int k = 0; k < PPG; ++k) {
            const int jl = J[pbase + k];
            if (jl >= 0) {
                const float u = Uw[pbase + k];
                const float g0 = W[jl * LB + bl], g1 = W[(jl + 1) * LB + bl];
                kahan_add(acc[k], cmp[k], fmaf(u, g1 - g0, g0));
            }
        }
    }
    const int b = b0 + bl;
    if (b >= B) return;
#pragma unroll
    for (int k = 0; k < PPG; ++k) {
        const int px = pix[(size_t)tile * 256 + pbase + k];
        if (px >= 0) out[(size_t)b * P + px] = scale * (acc[k] - cmp[k]);
    }
}

// XCD-grouped batch-lane gather: workgroups are dealt round-robin to the 8 XCDs;
// XCD g (= blockIdx.x % 8) takes detectors [g*DG, (g+1)*DG), DG = ceil(D/8), for every pixel slot
// and writes a partial image, so the bin rows its concurrent workgroups read
// (DG detectors x one small pixel region) stay in that XCD's L2.  The 8
// partials are summed in a fixed order by k_sum_parts.
template <int LB, int U>
__global__ __launch_bounds__(256) void k_gather_bl_xcd(
    const int32_t *__restrict__ tq, const int32_t *__restrict__ pix, const float *__restrict__ bins,
    float *__restrict__ part, BinGeom g, int Pd, int P, int D, int B, int Bp, int nsb)
{
    constexpr int RPW = 64 / LB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bl = lane % LB, rq = lane / LB;
    const int grp = blockIdx.x % 8, s = blockIdx.x / 8;
    const int sb = s % nsb, bc = s / nsb;
    const int slot = (sb * 4 + wv) * RPW + rq;
    const int b = bc * LB + bl;
    const bool bv = b < Bp;
    const int DG = (D + 7) / 8, d0 = grp * DG, dn = min(DG, D - d0);   // dn <= 0: empty group
    const int span = g.nT + 2 * g.K;
    const size_t rowstride = (size_t)g.NB * Bp;
    float acc = 0.0f, cmp = 0.0f;
    for (int dd = 0; dd < dn; dd += U) {
        int32_t q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) q[u] = dd + u < dn ? tq[(size_t)(d0 + dd + u) * Pd + slot] : INT32_MIN;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Decoded dc = decode_tau(q[u], g);
            if (bv && (unsigned)dc.nn < (unsigned)span) {
                const float *gb = bins + (size_t)(d0 + dd + u) * rowstride + b;
                const int j = 3 * dc.nn + dc.m;
                const float g0 = gb[(size_t)j * Bp], g1 = gb[(size_t)(j + 1) * Bp];
                kahan_add(acc, cmp, fmaf(dc.u, g1 - g0, g0));
            }
        }
    }
    const int px = pix[slot];
    if (px >= 0 && b < B) part[((size_t)grp * B + b) * P + px] = acc - cmp;
}

// Batch-lane gather with lane groups of 64 sinograms: each wave owns one pixel
// slot, so the travel time of every (slot, detector) pair is uniform over the
// wave.  Lane i decodes the pair of detector d0 + i (one strided load of 64
// travel times and one decode for 64 pairs); the wave then walks the 64 pairs
// with the bin offset and hat weight of pair k read out of lane k into scalar
// registers, issuing the two bin-row loads of U pairs before it accumulates
// them (plain sums of U pairs, compensated across groups).  XG: workgroups are
// dealt round-robin to the 8 XCDs and XCD x sums detectors
// [x*ceil(D/8), (x+1)*ceil(D/8)) into partial image x (k_sum_parts adds them).
// Bin offsets are 32-bit: D * NB * Bp < 2^31 (checked by the launcher).
template <int U, bool XG>
__global__ __launch_bounds__(256) void k_gather_wu(
    const int32_t *__restrict__ tq, const int32_t *__restrict__ pix, const float *__restrict__ bins,
    float *__restrict__ out, BinGeom g, int Pd, int P, int D, int B, int Bp, int nsb, float scale)
{
    const int lane = threadIdx.x & 63;
    int grp = 0, sb, bc;
    if (XG) {
        grp = blockIdx.x % 8;
        const int s = blockIdx.x / 8;
        sb = s % nsb;
        bc = s / nsb;
    } else {
        sb = blockIdx.x;
        bc = blockIdx.y;
    }
    const int slot = __builtin_amdgcn_readfirstlane(sb * 4 + (threadIdx.x >> 6));
    const int b = bc * 64 + lane;
    const int bo = min(b, Bp - 1);   // lanes past Bp read a valid column; their sums are dropped
    const int bo1 = bo + Bp;
    const int DG = XG ? (D + 7) / 8 : D;
    const int dbeg = grp * DG, dend = min(D, dbeg + DG);
    const int span = g.nT + 2 * g.K;
    float acc = 0.0f, cmp = 0.0f;
    for (int d0 = dbeg; d0 < dend; d0 += 64) {
        const int d = d0 + lane;
        const int32_t q = d < dend ? tq[(size_t)d * Pd + slot] : INT32_MIN;
        const Decoded dc = decode_tau(q, g);
        const bool ok = (unsigned)dc.nn < (unsigned)span;
        const int off = ok ? (d * g.NB + 3 * dc.nn + dc.m) * Bp : 0;
        const float wu = ok ? dc.u : 0.0f;
        const float wv = ok ? 1.0f : 0.0f;
        const int n = min(64, dend - d0);
        for (int k = 0; k < n; k += U) {
            float g0[U], g1[U], uu[U], vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int o = __builtin_amdgcn_readlane(off, k + u);
                uu[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, wu), k + u));
                vv[u] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, wv), k + u));
                const float *row = bins + o;
                g0[u] = row[bo];
                g1[u] = row[bo1];
            }
            float part = 0.0f;
#pragma unroll
            for (int u = 0; u < U; ++u) part += vv[u] * fmaf(uu[u], g1[u] - g0[u], g0[u]);
            kahan_add(acc, cmp, part);
        }
    }
    const int px = pix[slot];
    if (px >= 0 && b < B) {
        if (XG)
            out[((size_t)grp * B + b) * P + px] = acc - cmp;
        else
            out[(size_t)b * P + px] = scale * (acc - cmp);
    }
}

// Batch-lane gather from the decoded pair table (lane groups of 64 sinograms,
// one pixel slot per wave): the (bin row, u) of every pair of the wave's slot
// is read with scalar loads, so a pair costs two bin-row loads at a scalar base
// address and three vector operations; pairs outside the bin range read the
// zero rows.  XG: detector groups per XCD and partial images as k_gather_wu.
// Bin rows are read as a raw buffer with 32-bit byte offsets: the bins buffer,
// (D * NB + 2) * Bp floats, stays below 2^31 bytes (checked by the launcher).
template <int U, bool XG>
__global__ __launch_bounds__(256) void k_gather_wt(
    const int2 *__restrict__ ptab, const int32_t *__restrict__ pix, const float *__restrict__ bins,
    float *__restrict__ out, int Pd, int P, int D, int B, int Bp, int nsb, float scale, int bins_bytes)
{
    const int lane = threadIdx.x & 63;
    int grp = 0, sb, bc;
    if (XG) {
        grp = blockIdx.x % 8;
        const int s = blockIdx.x / 8;
        sb = s % nsb;
        bc = s / nsb;
    } else {
        sb = blockIdx.x;
        bc = blockIdx.y;
    }
    const int slot = __builtin_amdgcn_readfirstlane(sb * 4 + (threadIdx.x >> 6));
    const int b = bc * 64 + lane;
    // byte offsets of the lane's column in a bin row and the row after it (lanes
    // past Bp read a valid column; their sums are dropped)
    const int vo0 = 4 * min(b, Bp - 1), vo1 = vo0 + 4 * Bp;
    const int rowb = 4 * Bp;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(bins), (short)0, bins_bytes, kBufferWord3);
    const int DG = XG ? (D + 7) / 8 : D;
    const int dbeg = grp * DG, dend = min(D, dbeg + DG);
    const int2 *tr = ptab + (size_t)slot * D;
    float acc = 0.0f, cmp = 0.0f;
    int d0 = dbeg;
    int2 e[U];   // table entries of the current run; the next run's are read while its rows load
#pragma unroll
    for (int u = 0; u < U; ++u) e[u] = tr[d0 + u];   // the table is padded: reads past a row stay inside
    for (; d0 + U <= dend; d0 += U) {
        float g0[U], g1[U], uu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int so = e[u].x * rowb;
            g0[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, vo0, so, 0));
            g1[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, vo1, so, 0));
            uu[u] = __int_as_float(e[u].y);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = tr[d0 + U + u];
        float part = 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) part += fmaf(uu[u], g1[u] - g0[u], g0[u]);
        kahan_add(acc, cmp, part);
    }
    for (; d0 < dend; ++d0) {   // detectors past the last full run
        const int2 e = tr[d0];
        const int so = e.x * rowb;
        const float g0 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, vo0, so, 0));
        const float g1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, vo1, so, 0));
        kahan_add(acc, cmp, fmaf(__int_as_float(e.y), g1 - g0, g0));
    }
    const int px = pix[slot];
    if (px >= 0 && b < B) {
        if (XG)
            out[((size_t)grp * B + b) * P + px] = acc - cmp;
        else
            out[(size_t)b * P + px] = scale * (acc - cmp);
    }
}

// Tile gather from LDS, table-driven: one workgroup of 16 waves per 256-slot
// tile, chunk of 64 sinograms (lane = sinogram) and detector group.  For each
// detector the tile's bin window (rows [lo, lo + len) of 64 sinograms) is
// staged in one of two LDS buffers while the other is read: every thread holds
// its share of the next detector's window in registers across the current
// detector's work, so there is one barrier per detector.  Wave w interpolates
// the tile's pixels 16w .. 16w + 15: its (row, u) pairs come from the tile
// table with scalar loads; out-of-range pairs read two zero rows at index WMAX.
// Sums: over runs of 8 detectors, then of the runs (pairwise-like, two levels).  G > 1: the
// detector groups write partial images [G][B][P], summed by k_sum_parts.
template <int G>
__global__ __launch_bounds__(1024) void k_gather_tt(
    const int2 *__restrict__ ttab, const int2 *__restrict__ twin, const int32_t *__restrict__ pix,
    const float *__restrict__ bins, float *__restrict__ out, int D, int NB, int P, int B, int Bp, float scale,
    int bins_bytes, int ZR)
{
    extern __shared__ float sm[];   // 2 buffers x (ZR + 2) rows x 64 floats
    const int RS = ZR + 2;
    const int tile = blockIdx.x, grp = blockIdx.y % G, b0 = (blockIdx.y / G) * 64;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int DG = (D + G - 1) / G, dbeg = grp * DG, dend = min(D, dbeg + DG);
    for (int i = threadIdx.x; i < 2 * 2 * 64; i += 1024)   // the zero rows of both buffers
        sm[(i / 128) * RS * 64 + (ZR + (i / 64) % 2) * 64 + i % 64] = 0.0f;
    // window staging: thread t moves 16-byte pieces f = t + 1024 k (row f / 16, column
    // 4 (f % 16)) of the window's rows [lo, lo + len); columns past Bp carry the next
    // row's values and are never used (raw-buffer reads past the end return 0)
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(bins), (short)0, bins_bytes, kBufferWord3);
    const int rowb = 4 * Bp;
    int voff[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int f = threadIdx.x + 1024 * k;
        voff[k] = (f >> 4) * rowb + 4 * (b0 + 4 * (f & 15));
    }
    float4 pf[4];
    int plen = 0;   // rows of the window held in pf
    auto fetch = [&](int d) {
        const int2 w = twin[(size_t)tile * D + d];
        const int so = (d * NB + w.x) * rowb;
        plen = w.y;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (((threadIdx.x + 1024 * k) >> 4) < plen)
                pf[k] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff[k], so, 0));
    };
    auto store = [&](int buf) {
        float *dstb = sm + buf * RS * 64;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int f = threadIdx.x + 1024 * k;
            if ((f >> 4) < plen) *reinterpret_cast<float4 *>(dstb + (f >> 4) * 64 + 4 * (f & 15)) = pf[k];
        }
    };
    float acc[16], part[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = part[k] = 0.0f;
    const int2 *tt = ttab + (size_t)tile * D * 256 + 16 * wv;
    if (dbeg < dend) {
        fetch(dbeg);
        store(0);
        __syncthreads();
        if (dbeg + 1 < dend) fetch(dbeg + 1);
        int2 ec[16];   // this wave's table entries for the current detector (scalar registers)
#pragma unroll
        for (int k = 0; k < 16; ++k) ec[k] = tt[(size_t)dbeg * 256 + k];
        for (int d = dbeg; d < dend; ++d) {
            const int cur = (d - dbeg) & 1;
            const float *wb = sm + cur * RS * 64 + lane;
            int2 en[16];   // next detector's entries, read while this one is interpolated
            const int dn = min(d + 1, dend - 1);
#pragma unroll
            for (int k = 0; k < 16; ++k) en[k] = tt[(size_t)dn * 256 + k];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const float g0 = wb[ec[k].x * 64], g1 = wb[ec[k].x * 64 + 64];
                part[k] += fmaf(__int_as_float(ec[k].y), g1 - g0, g0);
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) ec[k] = en[k];
            if (((d - dbeg) & 7) == 7) {
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    acc[k] += part[k];
                    part[k] = 0.0f;
                }
            }
            if (d + 1 < dend) store(cur ^ 1);
            __syncthreads();
            if (d + 2 < dend) fetch(d + 2);
        }
    }
    const int b = b0 + lane;
    if (b >= B) return;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float v = acc[k] + part[k];
        const int px = pix[(size_t)tile * 256 + 16 * wv + k];
        if (px < 0) continue;
        if (G > 1)
            out[((size_t)grp * B + b) * P + px] = v;
        else
            out[(size_t)b * P + px] = scale * v;
    }
}

// Tile gather with asynchronous staging: as k_gather_tt, but each detector's
// bin window and the tile's pair-table entries go from memory straight into a
// ring of three LDS buffers (buffer loads with the LDS destination), two
// detectors ahead of the one being interpolated, without passing through
// registers.  Each wave issues a fixed number of loads per detector (the window
// rounded up to NCH rounds of 64 rows, ZR = 64 NCH; waves 0 and 1 also move the
// 256 table entries), so a wait for all but the latest detector's loads covers
// the detector about to be read.  The three buffers are distinct arrays and the
// detector loop is unrolled by three, so the compiler can tell the buffer being
// read from the one being filled (and does not wait for the fill before reading).
constexpr int kTaMaxD = 512;   // detectors whose window origins the async tile gather keeps in LDS

template <int NCH>
struct TaLayout {
    static constexpr int ZR = 64 * NCH, RS = ZR + 2;    // window rows, then two zero rows
    static constexpr int BUF = RS * 64 + 2 * 256;       // floats per buffer: rows, then 256 table entries
};

template <int G, int NCH>
__global__ __launch_bounds__(1024) void k_gather_ta(
    const int2 *__restrict__ ttab, const int2 *__restrict__ twin, const int32_t *__restrict__ pix,
    const float *__restrict__ bins, float *__restrict__ out, int D, int NB, int P, int B, int Bp, float scale,
    int bins_bytes, int ttab_bytes, const int32_t *__restrict__ tord)
{
    using L = TaLayout<NCH>;
    __shared__ __attribute__((aligned(16))) float s0[L::BUF], s1[L::BUF], s2[L::BUF];
    __shared__ int2 tw[kTaMaxD];   // the tile's window origins
    // tord: the tile of each workgroup (XCD-aware order; -1: a padding workgroup, which
    // exits before any barrier), or null: workgroup x = tile x
    const int tile = tord ? tord[blockIdx.x] : (int)blockIdx.x;
    if (tile < 0) return;
    const int grp = blockIdx.y % G, b0 = (blockIdx.y / G) * 64;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int DG = (D + G - 1) / G, dbeg = grp * DG, dend = min(D, dbeg + DG);
    for (int i = threadIdx.x; i < 2 * 64; i += 1024) {   // the zero rows of the three buffers
        s0[L::ZR * 64 + i] = 0.0f;
        s1[L::ZR * 64 + i] = 0.0f;
        s2[L::ZR * 64 + i] = 0.0f;
    }
    for (int i = threadIdx.x; i < D; i += 1024) tw[i] = twin[(size_t)tile * D + i];
    __syncthreads();
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(bins), (short)0, bins_bytes, kBufferWord3);
    const __amdgpu_buffer_rsrc_t rt =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int2 *>(ttab), (short)0, ttab_bytes, kBufferWord3);
    const int rowb = 4 * Bp;
    // load c of wave w moves rows 4 (16 c + w) .. + 3: lane l takes row + l / 16, columns b0 + 4 (l % 16) ..
    const int vo = (lane >> 4) * rowb + 4 * (b0 + 4 * (lane & 15));
    auto stage = [&](int d, float *bb) __attribute__((always_inline)) {
        const int so = (d * NB + __builtin_amdgcn_readfirstlane(tw[d].x)) * rowb;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int r0 = 4 * (16 * c + wv);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)(bb + r0 * 64),
                                                     16, vo, so + r0 * rowb, 0, 0);
        }
        if (wv < 2)   // table entries 128 w .. 128 w + 127 of (tile, d): 16 bytes (2 entries) per lane
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rt, (__attribute__((address_space(3))) void *)(bb + L::RS * 64 + 256 * wv), 16, 16 * lane,
                (int)((((size_t)tile * D + d) * 256 + 128 * wv) * sizeof(int2)), 0, 0);
    };
    float acc[16], part[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = part[k] = 0.0f;
    // one detector: wait for its loads, barrier, start the fill two ahead, interpolate
    auto step = [&](int d, const float *rd, float *wr) __attribute__((always_inline)) {
        if (d + 1 < dend) {   // this wave's loads for d are done (those for d + 1 may be in flight)
            if (wv < 2)
                __builtin_amdgcn_s_waitcnt(kWaitVm(NCH + 1));
            else
                __builtin_amdgcn_s_waitcnt(kWaitVm(NCH));
        } else {
            __builtin_amdgcn_s_waitcnt(kWaitVm(0));
        }
        // everyone's loads for d done; the buffer of d - 1 no longer read (a bare
        // barrier: a full fence here would also wait for the loads of d + 1)
        __builtin_amdgcn_s_barrier();
        if (d + 2 < dend) stage(d + 2, wr);
        const int4 *te = reinterpret_cast<const int4 *>(rd + L::RS * 64) + 8 * wv;   // this wave's 16 entries
        const float *wb = rd + lane;
        float g0[16], g1[16], uu[16];
#pragma unroll
        for (int k2 = 0; k2 < 8; ++k2) {
            const int4 e = te[k2];   // entries 2 k2, 2 k2 + 1 (same address in every lane)
            g0[2 * k2] = wb[e.x * 64];
            g1[2 * k2] = wb[e.x * 64 + 64];
            uu[2 * k2] = __int_as_float(e.y);
            g0[2 * k2 + 1] = wb[e.z * 64];
            g1[2 * k2 + 1] = wb[e.z * 64 + 64];
            uu[2 * k2 + 1] = __int_as_float(e.w);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) part[k] += fmaf(uu[k], g1[k] - g0[k], g0[k]);
        if (((d - dbeg) & 7) == 7) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                acc[k] += part[k];
                part[k] = 0.0f;
            }
        }
    };
    if (dbeg < dend) {
        stage(dbeg, s0);
        if (dbeg + 1 < dend) stage(dbeg + 1, s1);
        // unrolled by six: the compiler's wait before the first step of an iteration
        // (where it loses track of which buffer the pending fills target) is paid
        // once per six detectors
        for (int d = dbeg; d < dend; d += 6) {
            step(d, s0, s2);
            if (d + 1 < dend) step(d + 1, s1, s0);
            if (d + 2 < dend) step(d + 2, s2, s1);
            if (d + 3 < dend) step(d + 3, s0, s2);
            if (d + 4 < dend) step(d + 4, s1, s0);
            if (d + 5 < dend) step(d + 5, s2, s1);
        }
    }
    const int b = b0 + lane;
    if (b >= B) return;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float v = acc[k] + part[k];
        const int px = pix[(size_t)tile * 256 + 16 * wv + k];
        if (px < 0) continue;
        if (G > 1)
            out[((size_t)grp * B + b) * P + px] = v;
        else
            out[(size_t)b * P + px] = scale * v;
    }
}

// SIR transpose for batches, LDS tile form: WG = 16 waves, one 256-slot tile x 64
// sinograms x one of G detector groups.  Per detector the tile's window of bin rows
// (every in-range tap of its pairs) and its 256 pair entries are copied
// memory -> LDS by buffer loads with the LDS destination, one detector ahead into the
// other of two buffers.  Wave w takes pairs 16 w .. 16 w + 15: their weights (32 taps per
// pair, zero past the pair's in-range count) are loaded into 8 registers, two pairs per
// register, and each tap's weight is taken from its lane with readlane.  Per pair
// A = sum w g0 and C = sum w g1 over its taps (g0, g1: the tap's two bin rows), and the
// pair adds A + u (C - A) = sum w ((1 - u) g0 + u g1).
template <int G, int NCH>
__global__ __launch_bounds__(1024) void k_gather_tas(const int4 *__restrict__ stab, const int32_t *__restrict__ stwin,
                                                     const int32_t *__restrict__ pix, const float *__restrict__ bins,
                                                     const float *__restrict__ sirw, float *__restrict__ out, int D,
                                                     int NB, int P, int B, int Bp, float scale, int bins_bytes,
                                                     int stab_bytes)
{
    constexpr int ZR = 64 * NCH;
    constexpr int BUF = ZR * 64 + 4 * 256;   // floats per buffer: rows, then 256 int4 entries
    __shared__ __attribute__((aligned(16))) float s0[BUF], s1[BUF];
    __shared__ int32_t tw[kTaMaxD];   // the tile's window origins
    const int tile = blockIdx.x, grp = blockIdx.y % G, b0 = (blockIdx.y / G) * 64;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int DG = (D + G - 1) / G, dbeg = grp * DG, dend = min(D, dbeg + DG);
    for (int i = threadIdx.x; i < D; i += 1024) tw[i] = stwin[(size_t)tile * D + i];
    __syncthreads();
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(bins), (short)0, bins_bytes, kBufferWord3);
    // the tile's own table rows: D x 256 entries from its base (32-bit offsets whatever
    // the table's size)
    const __amdgpu_buffer_rsrc_t rt = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int4 *>(stab + (size_t)tile * D * 256), (short)0, stab_bytes, kBufferWord3);
    const int rowb = 4 * Bp;
    // load c of wave w moves rows 4 (16 c + w) .. + 3: lane l takes row + l / 16, columns b0 + 4 (l % 16) ..
    const int vo = (lane >> 4) * rowb + 4 * (b0 + 4 * (lane & 15));
    auto stage = [&](int d, float *bb) __attribute__((always_inline)) {
        const int so = (d * NB + __builtin_amdgcn_readfirstlane(tw[d])) * rowb;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int r0 = 4 * (16 * c + wv);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)(bb + r0 * 64),
                                                     16, vo, so + r0 * rowb, 0, 0);
        }
        // entries 16 w .. 16 w + 15 of (tile, d), 4 bytes per lane: every wave issues the same
        // number of loads, so one wait count fits all
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rt, (__attribute__((address_space(3))) void *)(bb + ZR * 64 + 64 * wv), 4, 4 * lane,
            __builtin_amdgcn_readfirstlane((d * 256 + 16 * wv) * (int)sizeof(int4)), 0, 0);
    };
    float acc[16], part[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = part[k] = 0.0f;
    auto step = [&](int d, const float *rd, float *wr) __attribute__((always_inline)) {
        // this wave's loads for d are done, then everyone's (and the buffer of d - 1 is free)
        __builtin_amdgcn_s_waitcnt(kWaitVm(0));
        __builtin_amdgcn_s_barrier();
        const int4 *te = reinterpret_cast<const int4 *>(rd + ZR * 64) + 16 * wv;
        // the weights are loaded with explicit instructions and waited for by hand: the
        // compiler's own wait for them would also wait for the staging loads issued after
        // them (its count cannot tell the two kinds of load apart); sirw[0] is a zero margin
        float W[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int4 e = te[2 * r + (lane >> 5)];
            const int t = lane & 31;
            const float *wp = sirw + (t < e.w ? e.z + t : 0);
            asm volatile("global_load_dword %0, %1, off" : "=v"(W[r]) : "v"(wp) : "memory");
        }
        if (d + 1 < dend) {
            stage(d + 1, wr);
            __builtin_amdgcn_s_waitcnt(kWaitVm(NCH + 1));   // the weights (older than the staging loads)
        } else {
            __builtin_amdgcn_s_waitcnt(kWaitVm(0));
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) asm volatile("" : "+v"(W[r]));   // uses stay after the wait
        const int4 ev = te[lane & 15];   // lane k < 16: entry k (taken with readlane below)
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int row0 = __builtin_amdgcn_readlane(ev.x, k), cnt = __builtin_amdgcn_readlane(ev.w, k);
            const float u = __int_as_float(__builtin_amdgcn_readlane(ev.y, k));
            const float *g = rd + row0 * 64 + lane;
            const int wk = __float_as_int(W[k >> 1]);
            float A = 0.0f, C = 0.0f;
            int c = 0;
            for (; c + 8 <= cnt; c += 8) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float w = __int_as_float(__builtin_amdgcn_readlane(wk, (k & 1) * 32 + c + i));
                    A = fmaf(w, g[(3 * (c + i)) * 64], A);
                    C = fmaf(w, g[(3 * (c + i) + 1) * 64], C);
                }
            }
#pragma unroll
            for (int i = 0; i < 7; ++i) {   // the last cnt % 8 taps (no reads past the pair's rows)
                if (c + i >= cnt) break;
                const float w = __int_as_float(__builtin_amdgcn_readlane(wk, (k & 1) * 32 + c + i));
                A = fmaf(w, g[(3 * (c + i)) * 64], A);
                C = fmaf(w, g[(3 * (c + i) + 1) * 64], C);
            }
            part[k] += fmaf(u, C - A, A);
        }
        if (((d - dbeg) & 7) == 7) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                acc[k] += part[k];
                part[k] = 0.0f;
            }
        }
    };
    if (dbeg < dend) {
        stage(dbeg, s0);
        for (int d = dbeg; d < dend; d += 2) {
            step(d, s0, s1);
            if (d + 1 < dend) step(d + 1, s1, s0);
        }
    }
    const int b = b0 + lane;
    if (b >= B) return;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float v = acc[k] + part[k];
        const int px = pix[(size_t)tile * 256 + 16 * wv + k];
        if (px < 0) continue;
        if (G > 1)
            out[((size_t)grp * B + b) * P + px] = v;
        else
            out[(size_t)b * P + px] = scale * v;
    }
}

// Transpose without SIR for one sinogram, LDS tile form: WG = one 256-slot tile (thread k:
// slot k) x one of G detector groups.  The group's window origins go to LDS first; per
// detector the tile's bin window (at most tile_zr <= 256 rows; one row per thread) is
// staged in LDS, double buffered, with the window rows and travel times of the next PF
// detectors in flight (a register ring); each thread decodes its travel time and
// interpolates its two bins from LDS (Kahan sum over detectors, as the record-lane gather).
// DPB detectors share one barrier (their windows staged together; PF a multiple of DPB).
template <int G, int PF, int DPB = 1>
__global__ __launch_bounds__(256) void k_gather_ta1(const int32_t *__restrict__ tq, const int2 *__restrict__ twin,
                                                    const int32_t *__restrict__ pix, const float *__restrict__ bins,
                                                    float *__restrict__ out, BinGeom g, int D, int Pd, int P,
                                                    float scale, int ZR)
{
    static_assert(PF % DPB == 0, "the ring holds whole barrier steps");
    __shared__ float win[2][DPB * 256];
    __shared__ int los[256];
    const int tile = blockIdx.x, grp = blockIdx.y, k = threadIdx.x;
    const int DG = (D + G - 1) / G, dbeg = grp * DG, dend = min(D, dbeg + DG);
    const int span = g.nT + 2 * g.K, NB = g.NB;
    const size_t nrows = (size_t)D * NB + 2;   // bins rows incl. the two zero rows
    const size_t slot = (size_t)tile * 256 + k;
    if (k < dend - dbeg) los[k] = twin[(size_t)tile * D + dbeg + k].x;   // DG <= 256
    __syncthreads();
    auto row_of = [&](int d) -> float {
        const size_t r = (size_t)d * NB + los[d - dbeg] + k;
        return (k < ZR && r < nrows) ? bins[r] : 0.0f;
    };
    auto tq_of = [&](int d) -> int32_t { return tq[(size_t)d * Pd + slot]; };
    float acc = 0.0f, cmp = 0.0f;
    // register ring: slot i holds detector d0 + i's travel time and window row until its
    // step, then is refilled with detector d0 + i + PF
    int32_t qa[PF];
    float wa[PF];
#pragma unroll
    for (int i = 0; i < PF; ++i) {
        const int d = dbeg + i;
        qa[i] = d < dend ? tq_of(d) : 0;
        wa[i] = d < dend ? row_of(d) : 0.0f;
    }
    for (int d0 = dbeg; d0 < dend; d0 += PF) {
#pragma unroll
        for (int i = 0; i < PF; i += DPB) {
            const int d = d0 + i;
            if (d >= dend) break;   // uniform over the workgroup
            float *wb = win[((d - dbeg) / DPB) & 1];
#pragma unroll
            for (int h = 0; h < DPB; ++h) wb[h * 256 + k] = wa[i + h];
            // the buffer written here was last read two steps ago: every thread has
            // passed the previous step's barrier since
            __syncthreads();
#pragma unroll
            for (int h = 0; h < DPB; ++h) {
                const int dd = d + h;
                const int32_t qd = qa[i + h];
                if (dd + PF < dend) {
                    qa[i + h] = tq_of(dd + PF);
                    wa[i + h] = row_of(dd + PF);
                }
                if (dd < dend) {
                    const Decoded dc = decode_tau(qd, g);
                    if ((unsigned)dc.nn < (unsigned)span) {
                        const int r = 3 * dc.nn + dc.m - los[dd - dbeg];
                        const float g0 = wb[h * 256 + r], g1 = wb[h * 256 + r + 1];
                        kahan_add(acc, cmp, fmaf(dc.u, g1 - g0, g0));
                    }
                }
            }
        }
    }
    const int px = pix[slot];
    if (px < 0) return;
    if (G > 1)
        out[(size_t)grp * P + px] = acc - cmp;
    else
        out[px] = scale * (acc - cmp);
}

// As k_gather_ta1, reading the packed tile-major pair table (row inside the tile's window |
// u in 23 bits): each workgroup's stream is one contiguous run per detector group, and the
// pair needs no decode; rows outside the bin range point at the zero rows tile_zr, +1.
template <int G>
__global__ __launch_bounds__(256) void k_gather_tp1(const uint32_t *__restrict__ tq1, const int2 *__restrict__ twin,
                                                    const int32_t *__restrict__ pix, const float *__restrict__ bins,
                                                    float *__restrict__ out, int D, int NB, int P, float scale, int ZR)
{
    __shared__ float win[2][258];
    __shared__ int los[256];
    const int tile = blockIdx.x, grp = blockIdx.y, k = threadIdx.x;
    const int DG = (D + G - 1) / G, dbeg = grp * DG, dend = min(D, dbeg + DG);
    const size_t nrows = (size_t)D * NB + 2;   // bins rows incl. the two zero rows
    const size_t slot = (size_t)tile * 256 + k;
    if (k < dend - dbeg) los[k] = twin[(size_t)tile * D + dbeg + k].x;   // DG <= 256
    if (k < 4) win[k >> 1][256 + (k & 1)] = 0.0f;
    __syncthreads();
    auto row_of = [&](int d) -> float {
        const size_t r = (size_t)d * NB + los[d - dbeg] + k;
        return (k < ZR && r < nrows) ? bins[r] : 0.0f;
    };
    const uint32_t *tq = tq1 + (size_t)tile * D * 256 + k;
    const float su = 1.0f / 8388608.0f;
    float acc = 0.0f, cmp = 0.0f;
    if (dbeg < dend) {
        uint32_t q0 = tq[(size_t)dbeg * 256], q1 = 0;
        float w0 = row_of(dbeg), w1 = 0.0f;
        if (dbeg + 1 < dend) {
            q1 = tq[(size_t)(dbeg + 1) * 256];
            w1 = row_of(dbeg + 1);
        }
        for (int d = dbeg; d < dend; ++d) {
            float *wb = win[(d - dbeg) & 1];
            wb[k] = w0;
            // the buffer written here was last read two detectors ago: every thread has
            // passed the previous detector's barrier since
            __syncthreads();
            const uint32_t qd = q0;
            q0 = q1;
            w0 = w1;
            if (d + 2 < dend) {
                q1 = tq[(size_t)(d + 2) * 256];
                w1 = row_of(d + 2);
            }
            const int r = (int)(qd & 511u);
            const float u = (float)(qd >> 9) * su, g0 = wb[r], g1 = wb[r + 1];
            kahan_add(acc, cmp, fmaf(u, g1 - g0, g0));
        }
    }
    const int px = pix[slot];
    if (px < 0) return;
    if (G > 1)
        out[(size_t)grp * P + px] = acc - cmp;
    else
        out[px] = scale * (acc - cmp);
}

// SIR transpose for one sinogram, LDS tile form: WG = one 256-slot tile (thread k: slot k)
// x one of G detector groups.  Per detector the tile's window of bin rows (the SIR tile
// table's, at most 256 rows) is copied to LDS (one row per thread, double buffered); the
// weights of a wave's 64 pairs are loaded cooperatively (each load instruction reads two
// pairs' contiguous windows, 32 taps each, instead of 64 scattered words) one detector
// ahead and passed through a per-wave LDS area; each thread then sums its pair's in-range
// taps w (g0 + u (g1 - g0)) from LDS.
constexpr int kTas1WS = 65;   // LDS floats per pair couple: pair 2i at 65 i, pair 2i + 1 at 65 i + 32

template <int G>
__global__ __launch_bounds__(256) void k_gather_tas1(const int4 *__restrict__ stab, const int32_t *__restrict__ stwin,
                                                     const int32_t *__restrict__ pix, const float *__restrict__ bins,
                                                     const float *__restrict__ sirw, float *__restrict__ out, int D,
                                                     int NB, int P, float scale, int ZR)
{
    __shared__ float win[2][256];
    __shared__ float wts[4][32 * kTas1WS];
    const int tile = blockIdx.x, grp = blockIdx.y, k = threadIdx.x, lane = k & 63;
    const int DG = (D + G - 1) / G, dbeg = grp * DG, dend = min(D, dbeg + DG);
    const size_t nrows = (size_t)D * NB + 2;   // bins rows incl. the two zero rows
    float *ws = wts[k >> 6];
    auto row_of = [&](int d) -> float {
        const size_t r = (size_t)d * NB + stwin[(size_t)tile * D + d] + k;
        return (k < ZR && r < nrows) ? bins[r] : 0.0f;
    };
    // instruction i: lanes 0..31 load taps 0..31 of the wave's pair 2i, lanes 32..63 of pair
    // 2i + 1 (zero past the pair's count: sirw[0] is a zero margin)
    auto wload = [&](const int4 &ee, float (&wr)[32]) {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            const int q = 2 * i + (lane >> 5), t = lane & 31;
            const int cq = __shfl(ee.w, q, 64), zq = __shfl(ee.z, q, 64);
            wr[i] = sirw[t < cq ? zq + t : 0];
        }
    };
    float wr[32];
    float acc = 0.0f, part = 0.0f;
    if (dbeg < dend) {
        int4 e = stab[((size_t)tile * D + dbeg) * 256 + k];
        float wv = row_of(dbeg);
        wload(e, wr);
#pragma unroll
        for (int i = 0; i < 32; ++i) ws[kTas1WS * i + lane] = wr[i];
        for (int d = dbeg; d < dend; ++d) {
            float *wb = win[(d - dbeg) & 1];
            wb[k] = wv;
            // the window buffer written here was last read two detectors ago: every thread
            // has passed the previous detector's barrier since
            __syncthreads();
            int4 en = e;
            if (d + 1 < dend) {
                en = stab[((size_t)tile * D + d + 1) * 256 + k];
                wv = row_of(d + 1);
                wload(en, wr);
            }
            const int cnt = e.w;
            const float *wp = ws + kTas1WS * (lane >> 1) + 32 * (lane & 1);
            const float *g = wb + e.x;
            float A = 0.0f, C = 0.0f;
            int t = 0;
            for (; t + 4 <= cnt; t += 4) {
                A = fmaf(wp[t], g[3 * t], A);
                C = fmaf(wp[t], g[3 * t + 1], C);
                A = fmaf(wp[t + 1], g[3 * t + 3], A);
                C = fmaf(wp[t + 1], g[3 * t + 4], C);
                A = fmaf(wp[t + 2], g[3 * t + 6], A);
                C = fmaf(wp[t + 2], g[3 * t + 7], C);
                A = fmaf(wp[t + 3], g[3 * t + 9], A);
                C = fmaf(wp[t + 3], g[3 * t + 10], C);
            }
            for (; t < cnt; ++t) {
                A = fmaf(wp[t], g[3 * t], A);
                C = fmaf(wp[t], g[3 * t + 1], C);
            }
            part += fmaf(__int_as_float(e.y), C - A, A);
            if (((d - dbeg) & 7) == 7) {
                acc += part;
                part = 0.0f;
            }
            if (d + 1 < dend) {   // the wave's own area: its reads above come first (LDS order)
#pragma unroll
                for (int i = 0; i < 32; ++i) ws[kTas1WS * i + lane] = wr[i];
            }
            e = en;
        }
    }
    const int px = pix[(size_t)tile * 256 + k];
    if (px < 0) return;
    if (G > 1)
        out[(size_t)grp * P + px] = acc + part;
    else
        out[px] = scale * (acc + part);
}

// out[b][x] = scale * sum of the NP partial images (fixed order)
template <int NP>
__global__ __launch_bounds__(256) void k_sum_parts(const float *__restrict__ part, float *__restrict__ out,
                                                   size_t n, float scale)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < NP; ++k) s += part[(size_t)k * n + i];
        out[i] = scale * s;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }
constexpr size_t kLdsBudget = 160 * 1024;   // LDS per workgroup on gfx950
constexpr size_t kHatLds = 32 * 1024;       // LDS per workgroup of the hat-filter kernels

static int scatter_chunk_max(const Plan *p, int Bp)
{
    const int WLp = p->WL | 1;
    if (Bp == 1) return 1;
    int bc = (int)(kLdsBudget / (sizeof(float) * (size_t)WLp));
    bc = std::min(bc, Bp) / 4 * 4;
    return std::max(bc, 4);
}

// batch-lane scatter lane-group width for this plan and batch stride (0 = not usable)
static int scatter_lb_max(const Plan *p, int Bp)
{
    if (Bp < 16) return 0;
    const size_t WLp = (size_t)(p->WL | 1);
    for (int lb = 64; lb >= 16; lb /= 2)
        if (lb <= std::max(16, Bp) && lb * WLp * sizeof(float) <= kLdsBudget) return lb;
    return 0;
}

template <int LB>
static void scatter_bl(Plan *p, const float *imgI, int B, int Bp, hipStream_t s)
{
    const int WLp = p->WL | 1;
    dim3 grid(p->D, p->nbk, (B + LB - 1) / LB);
    const size_t lds = sizeof(float) * (size_t)LB * WLp;
    if (p->use_sir)
        hipLaunchKernelGGL((k_scatter_bl<true, LB>), grid, dim3(256), lds, s, p->d_rec, p->d_rec_off, p->d_rec_k,
                           imgI, p->d_hpart, p->d_sir, p->d_swin, p->Ls, p->geom, p->Wb, p->nbk, p->lo_off, p->WL,
                           WLp, Bp, B);
    else
        hipLaunchKernelGGL((k_scatter_bl<false, LB>), grid, dim3(256), lds, s, p->d_rec, p->d_rec_off, p->d_rec_k,
                           imgI, p->d_hpart, p->d_sir, p->d_swin, p->Ls, p->geom, p->Wb, p->nbk, p->lo_off, p->WL,
                           WLp, Bp, B);
}

// scatter launch choice: > 0 record-lane kernel with that LDS batch chunk, < 0
// batch-lane kernel with lane groups of -value sinograms
// SIR forward in pull form (records of build_sir_pull): scatter choices of their own, the
// bins written directly (no bucket windows).  kSirPull: for one sinogram the lane-group
// kernel (the default there), else the batch-lane kernel with 24 samples and 2 records in
// flight per wave; kSirPull + 1 .. 3: batch-lane kernels with (16 samples, 4 records),
// (16, 2), (32, 1); kSirBand (batches; the default there): every record visited once by
// the wave of its first-sample band, the overlapping wave rows summed in a second pass
constexpr int kSirPull = 1000, kSirPullVariants = 5, kSirBand = kSirPull + 4;

static bool is_sir_pull(int c) { return c >= kSirPull && c < kSirPull + kSirPullVariants; }

static bool sir_pull_ok(const Plan *p, int Bp, int c = kSirPull)
{
    if (c == kSirBand)
        return p->use_sir && p->d_srec && p->d_sirw && Bp > 1 && p->d_sband && p->sband_nfb >= 1 &&
               p->bcap >= 1 && batch_stride(p->bcap) >= Bp;
    return p->use_sir && p->d_srec && p->d_sirw && is_sir_pull(c) && (c == kSirPull || Bp > 1);
}

static int scatter_choice(const Plan *p, int Bp)
{
    if (p->tune.Bp == Bp && sir_pull_ok(p, Bp, p->tune.scatter_bc)) return p->tune.scatter_bc;
    if (!(p->tune.Bp == Bp && p->tune.scatter_bc != 0) && sir_pull_ok(p, Bp))
        return sir_pull_ok(p, Bp, kSirBand) ? kSirBand : kSirPull;
    if (p->tune.Bp == Bp && p->tune.scatter_bc != 0 && !is_sir_pull(p->tune.scatter_bc)) {
        const int c = p->tune.scatter_bc;
        if (c > 0) return std::min(c, scatter_chunk_max(p, Bp));
        if (-c <= scatter_lb_max(p, Bp)) return c;
    }
    const int lb = scatter_lb_max(p, Bp);
    return lb ? -lb : scatter_chunk_max(p, Bp);
}

// pull forward choice: > 0 bins per thread pass BC (1, 4, 8, 16); < 0 batch-lane,
// -(LB + 100 * mode) with lane groups LB (16, 32, 64) and mode 0: plain order,
// 8 pairs in flight; 1: XCD-aware order, 8 in flight; 2: XCD-aware, 16 in flight
static bool pull_choice_ok(const Plan *p, int c, int Bp)
{
    if (c == 5 || c == 6 || c == 7 || c == 9 || c == 10) return Bp == 1 && p->d_tfrec && p->d_img_slot;
    if (c >= 11 && c <= 13) return Bp == 1 && p->d_tfsrec && p->d_img_slot;
    if (c > 0) return c == 1 || ((c == 2 || c == 3) && Bp == 1) || ((c == 4 || c == 8 || c == 16) && Bp % c == 0);
    if (c < 0) {
        const int lb = (-c) % 100, mode = (-c) / 100;
        int cap = 16;
        while (cap < Bp) cap *= 2;
        if (mode >= 9)   // raw-buffer image rows: 32-bit byte offsets
            return Bp >= 16 && lb == 64 && mode <= 11 && sizeof(float) * (size_t)p->P * Bp < ((size_t)1 << 31);
        if (mode >= 3)   // wave-uniform runs: lane groups of 64 only
            return Bp >= 16 && lb == 64 && mode <= 8;
        return Bp >= 16 && (lb == 16 || lb == 32 || lb == 64) && lb <= cap && mode <= 2;
    }
    return false;
}

static int default_pull_choice(int Bp)
{
    if (Bp >= 64) return -964;
    if (Bp >= 32) return -32;
    if (Bp >= 16) return -16;
    if (Bp % 8 == 0) return 8;
    if (Bp % 4 == 0) return 4;
    return Bp == 1 ? 2 : 1;
}

constexpr int kPullRange = 32;   // bins per lane group (batch-lane pull)

// one sinogram's image in slot order (0 in the pad slots): the record stream then reads
// pixel values of the same device tile from the same few cache lines
__global__ __launch_bounds__(256) void k_img_to_slots(const int32_t *__restrict__ pix, const float *__restrict__ img,
                                                      float *__restrict__ out, int Pd)
{
    const int sl = blockIdx.x * 256 + threadIdx.x;
    if (sl >= Pd) return;
    const int px = pix[sl];
    out[sl] = px >= 0 ? img[px] : 0.0f;
}

// The 4096 slot values of a 4 x 4-tile block (tile k of the block: slots 256 k ..) into
// LDS, 0 for tiles past the grid; all 16 loads of a thread in flight before the writes
__device__ __forceinline__ void stage_block_image(float *v, const float *__restrict__ img_slot, int bx, int bz, int ntx,
                                                  int ntz, int t)
{
    float x[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int tx = 4 * bx + (k >> 2), tz = 4 * bz + (k & 3);
        x[k] = (tx < ntx && tz < ntz) ? img_slot[(size_t)(tx * ntz + tz) * 256 + t] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) v[t + 256 * k] = x[k];
}

// Tile forward for one sinogram, pass 1: workgroup = one 4 x 4-tile block (its 4096 slot
// values staged in LDS) x a group of detectors; thread per window row r of a (block,
// detector) pair: partial[r] = sum of (1 - u) v over the records of row r + sum of u v over
// the records of row r - 1 (rows = bins jmin + r; the window has nrow + 1 partials).
// The pair's records (<= 4096) and row offsets (<= kTfMaxRows) are copied to LDS with
// coalesced loads; the next detector's are loaded into registers while this one's rows
// are summed, so the rows read records and pixel values from LDS only.
// G detectors per workgroup (choices 5, 6, 7: G = 8, 4, 2)
template <int G>
__global__ __launch_bounds__(256) void k_fwd_tile1(const float *__restrict__ img_slot, const uint32_t *__restrict__ rec,
                                                   const uint16_t *__restrict__ roff, const int4 *__restrict__ desc,
                                                   float *__restrict__ part, int nbz, int ntz, int ntx, int D)
{
    constexpr int RPT = 4096 / 256, OPT = kTfMaxRows / 256;   // records / offsets per thread
    __shared__ float v[4096];
    __shared__ uint32_t rb[4096];
    __shared__ uint16_t ob[kTfMaxRows];
    __shared__ float edge[OPT][4];   // u part of the last row of each wave, per row slot
    __shared__ int4 ed[G + 1];
    const int T = blockIdx.x, bx = T / nbz, bz = T - bx * nbz;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    stage_block_image(v, img_slot, bx, bz, ntx, ntz, t);
    const int d0 = blockIdx.y * G, d1 = min(D, d0 + G);
    // the group's descriptors (and the next entry, for the run lengths) up front: the
    // record loads of a detector then wait on no other global load
    if (t <= d1 - d0) ed[t] = desc[(size_t)T * D + d0 + t];
    __syncthreads();
    // records of (T, d) are [e.x, e.x + n), n = the next entry's start - e.x (the table ends
    // with a sentinel entry); offsets [e.y, e.y + nrow].  Two register sets: the loads of
    // the next two detectors are in flight while one detector's rows are summed.
    uint32_t ra[RPT], rb2[RPT];
    uint16_t oa[OPT], ob2[OPT];
    auto fetch = [&](int di, uint32_t (&rr)[RPT], uint16_t (&orr)[OPT]) __attribute__((always_inline)) {
        const int4 e = ed[di];
        const int n = ed[di + 1].x - e.x;
        const uint32_t *r0 = rec + e.x;
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
            const int k = t + 256 * i;
            rr[i] = k < n ? r0[k] : 0u;
        }
        const uint16_t *o0 = roff + e.y;
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int k = t + 256 * i;
            orr[i] = k <= e.w ? o0[k] : (uint16_t)0;
        }
    };
    const float su = 1.0f / 1048576.0f;
    auto step = [&](int d, uint32_t (&rr)[RPT], uint16_t (&orr)[OPT]) __attribute__((always_inline)) {
        const int4 e = ed[d - d0];
        __syncthreads();   // the previous detector's rows done with rb / ob / edge (and v written)
#pragma unroll
        for (int i = 0; i < RPT; ++i) rb[t + 256 * i] = rr[i];
#pragma unroll
        for (int i = 0; i < OPT; ++i) ob[t + 256 * i] = orr[i];
        __syncthreads();
        if (d + 2 < d1) fetch(d + 2 - d0, rr, orr);
        // thread per row r < nrow: its records [b, c) give (1 - u) v to row r (own) and u v
        // to row r + 1 (up); partial[r] = own[r] + up[r - 1], up[r - 1] from the previous
        // lane, or through LDS across waves and row slots.  Records in chunks of 8 whose LDS
        // reads are all issued before the sums (lanes past c read record b with weight 0).
        float own[OPT], upl[OPT];
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = t + 256 * i;
            float a0 = 0.0f, a1 = 0.0f;
            if (r < e.w) {
                const int b = ob[r], c = ob[r + 1];
                for (int k0 = b; k0 < c; k0 += 8) {
                    uint32_t q[8];
                    float val[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) q[j] = rb[k0 + j < c ? k0 + j : b];
#pragma unroll
                    for (int j = 0; j < 8; ++j) val[j] = v[q[j] & 4095u];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float u = (float)(q[j] >> 12) * su, x = k0 + j < c ? val[j] : 0.0f;
                        a0 = fmaf(1.0f - u, x, a0);
                        a1 = fmaf(u, x, a1);
                    }
                }
            }
            own[i] = a0;
            upl[i] = __shfl_up(a1, 1);
            if (lane == 63) edge[i][wv] = a1;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = t + 256 * i;
            if (r <= e.w) {
                const float prev = lane > 0 ? upl[i] : (wv > 0 ? edge[i][wv - 1] : (i > 0 ? edge[i - 1][3] : 0.0f));
                part[e.y + r] = (r < e.w ? own[i] : 0.0f) + prev;
            }
        }
    };
    if (d0 < d1) fetch(0, ra, oa);
    if (d0 + 1 < d1) fetch(1, rb2, ob2);
    for (int d = d0; d < d1; d += 2) {
        step(d, ra, oa);
        if (d + 1 < d1) step(d + 1, rb2, ob2);
    }
}

// pass 2: bins[d][j] = sum over blocks T (in order) of their partial covering j; a
// workgroup takes 1024 bins of one detector, 4 per thread.  The blocks whose windows meet
// those bins are listed first (in block order), then read four at a time so that 16 loads
// per thread are in flight before the sums (which keep the block order).
__global__ __launch_bounds__(256) void k_fwd_tile1_sum(const float *__restrict__ part, const int4 *__restrict__ desc,
                                                       float *__restrict__ bins, int NB, int nblk, int D)
{
    __shared__ int4 w[256];
    __shared__ int cov[256];
    __shared__ int ncov, wcnt[4];
    const int d = blockIdx.y;
    const int j0 = blockIdx.x * 1024;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int t0 = 0; t0 < nblk; t0 += 256) {
        __syncthreads();
        const int nt = min(256, nblk - t0);
        if ((int)threadIdx.x < nt) w[threadIdx.x] = desc[(size_t)(t0 + threadIdx.x) * D + d];
        __syncthreads();
        {   // compact the blocks whose window meets these bins, keeping their order
            const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
            bool c = false;
            if (t < nt) {
                const int4 e = w[t];
                c = e.w > 0 && e.z <= j0 + 1023 && e.z + e.w >= j0;
            }
            const unsigned long long m = __ballot(c);
            if (lane == 0) wcnt[wv] = __popcll(m);
            __syncthreads();
            int off = __popcll(m & ((1ull << lane) - 1ull));
            for (int i = 0; i < wv; ++i) off += wcnt[i];
            if (c) cov[off] = t;
            if (t == 0) ncov = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        }
        __syncthreads();
        const int n = ncov;
        for (int i0 = 0; i0 < n; i0 += 4) {
            float val[4][4];
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) {
                const int i = i0 + ii;
                const int4 e = i < n ? w[cov[i]] : make_int4(0, 0, 0, -1);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = j0 + (int)threadIdx.x + 256 * q - e.z;
                    val[ii][q] = (r >= 0 && r <= e.w) ? part[e.y + r] : 0.0f;
                }
            }
#pragma unroll
            for (int ii = 0; ii < 4; ++ii)
                if (i0 + ii < n) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] += val[ii][q];
                }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = j0 + (int)threadIdx.x + 256 * q;
        if (j < NB) bins[(size_t)d * NB + j] = acc[q];
    }
}

// Tile forward for one sinogram, pass 1, lean form of k_fwd_tile1 (same tables, same
// partials): the descriptors are read as scalars, the record and row-offset loads of a
// detector are unconditional (indices clamped to the pair's last entry; entries past it
// are never read back), a row's records are read from LDS at immediate offsets in chunks
// of C, and a row sums s_x = sum v and s_t = sum t v (t = u 2^20, the record's integer u),
// so that its own part is s_x - s_t 2^-20 and its up part s_t 2^-20: four VALU
// operations per record for the sums instead of six.
template <int G, int C>
__global__ __launch_bounds__(256) void k_fwd_tile1r(const float *__restrict__ img_slot, const uint32_t *__restrict__ rec,
                                                    const uint16_t *__restrict__ roff, const int4 *__restrict__ desc,
                                                    float *__restrict__ part, int nbz, int ntz, int ntx, int D)
{
    constexpr int RPT = 4096 / 256, OPT = kTfMaxRows / 256;
    __shared__ float v[4096];
    __shared__ uint32_t rb[4096 + C];   // C entries of slack: a chunk may run past the pair's last record
    __shared__ uint16_t ob[kTfMaxRows];
    __shared__ float edge[OPT][4];
    const int T = blockIdx.x, bx = T / nbz, bz = T - bx * nbz;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    stage_block_image(v, img_slot, bx, bz, ntx, ntz, t);
    if (t < C) rb[4096 + t] = 0u;
    const int d0 = blockIdx.y * G, d1 = min(D, d0 + G);
    const int4 *dp = desc + (size_t)T * D;   // entry d + 1 always exists (sentinel after the last pair)
    uint32_t ra[RPT], rb2[RPT];
    uint16_t oa[OPT], ob2[OPT];
    auto fetch = [&](int d, uint32_t (&rr)[RPT], uint16_t (&orr)[OPT]) __attribute__((always_inline)) {
        const int4 e = dp[d];
        const int nl = max(dp[d + 1].x - e.x - 1, 0);   // the pair's last record (the next entry's first if none)
        const uint32_t *r0 = rec + e.x;
#pragma unroll
        for (int i = 0; i < RPT; ++i) rr[i] = r0[min(t + 256 * i, nl)];
        const uint16_t *o0 = roff + e.y;
#pragma unroll
        for (int i = 0; i < OPT; ++i) orr[i] = o0[min(t + 256 * i, e.w)];
    };
    const float su = 1.0f / 1048576.0f;
    auto step = [&](int d, uint32_t (&rr)[RPT], uint16_t (&orr)[OPT]) __attribute__((always_inline)) {
        const int4 e = dp[d];
        __syncthreads();   // the previous detector's rows done with rb / ob / edge (and v written)
#pragma unroll
        for (int i = 0; i < RPT; ++i) rb[t + 256 * i] = rr[i];
#pragma unroll
        for (int i = 0; i < OPT; ++i) ob[t + 256 * i] = orr[i];
        __syncthreads();
        if (d + 2 < d1) fetch(d + 2, rr, orr);
        float own[OPT], upl[OPT];
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = t + 256 * i;
            float sx = 0.0f, st = 0.0f;
            if (r < e.w) {
                const int b = ob[r], c = ob[r + 1];
                for (int k0 = b; k0 < c; k0 += C) {
                    const int rem = c - k0;
                    uint32_t q[C];
                    float val[C];
#pragma unroll
                    for (int j = 0; j < C; ++j) q[j] = rb[k0 + j];
#pragma unroll
                    for (int j = 0; j < C; ++j) val[j] = v[q[j] & 4095u];
#pragma unroll
                    for (int j = 0; j < C; ++j) {
                        const float x = j < rem ? val[j] : 0.0f;
                        sx += x;
                        st = fmaf((float)(q[j] >> 12), x, st);
                    }
                }
            }
            const float a1 = st * su;
            own[i] = sx - a1;
            upl[i] = __shfl_up(a1, 1);
            if (lane == 63) edge[i][wv] = a1;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = t + 256 * i;
            if (r <= e.w) {
                const float prev = lane > 0 ? upl[i] : (wv > 0 ? edge[i][wv - 1] : (i > 0 ? edge[i - 1][3] : 0.0f));
                part[e.y + r] = (r < e.w ? own[i] : 0.0f) + prev;
            }
        }
    };
    if (d0 < d1) fetch(d0, ra, oa);
    if (d0 + 1 < d1) fetch(d0 + 1, rb2, ob2);
    for (int d = d0; d < d1; d += 2) {
        step(d, ra, oa);
        if (d + 1 < d1) step(d + 1, rb2, ob2);
    }
}

// Tile forward for one sinogram, pass 1, sorted-row form (the partials of k_fwd_tile1): each
// pair's rows come in order of decreasing record count (d_tfsrec / d_tfsro), so a block of
// 64 consecutive sorted positions (lane = position) has nearly equal record counts and its
// lanes run nearly the same number of iterations (the natural row order wastes about half
// of them).  The blocks of a pair go to the four waves in snake order (0 1 2 3 3 2 1 0 ..),
// which balances their descending costs; each lane sums s_x = sum v and s_t = sum t v over
// its row's records (as k_fwd_tile1r, records in the same order: identical sums), the rows'
// own and up parts go to LDS by row, and the window partials are written in row order.
template <int G, int C>
__global__ __launch_bounds__(256) void k_fwd_tile1s(const float *__restrict__ img_slot, const uint32_t *__restrict__ srec,
                                                    const uint32_t *__restrict__ sro, const int4 *__restrict__ desc,
                                                    float *__restrict__ part, int nbz, int ntz, int ntx, int D)
{
    constexpr int RPT = 4096 / 256, OPT = kTfMaxRows / 256, KB = kTfMaxRows / 64 / 4;   // blocks per wave
    // one array (40 KB: four workgroups per CU): slot values, records (after the row sums:
    // own part by row at [0, 2048), up part at [2048, 4096)), sorted positions; a chunk
    // that runs past the last record reads position entries, with weight 0
    __shared__ uint32_t smem[4096 + 4096 + kTfMaxRows];
    float *v = reinterpret_cast<float *>(smem);
    uint32_t *rb = smem + 4096, *pb = smem + 8192;
    const int T = blockIdx.x, bx = T / nbz, bz = T - bx * nbz;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    stage_block_image(v, img_slot, bx, bz, ntx, ntz, t);
    const int d0 = blockIdx.y * G, d1 = min(D, d0 + G);
    const int4 *dp = desc + (size_t)T * D;   // entry d + 1 always exists (sentinel after the last pair)
    uint32_t ra[RPT], rb2[RPT], oa[OPT], ob2[OPT];
    auto fetch = [&](int d, uint32_t (&rr)[RPT], uint32_t (&orr)[OPT]) __attribute__((always_inline)) {
        const int4 e = dp[d];
        const int nl = max(dp[d + 1].x - e.x - 1, 0);   // the pair's last record (the next entry's first if none)
        const uint32_t *r0 = srec + e.x;
#pragma unroll
        for (int i = 0; i < RPT; ++i) rr[i] = r0[min(t + 256 * i, nl)];
        const uint32_t *o0 = sro + e.y;
#pragma unroll
        for (int i = 0; i < OPT; ++i) orr[i] = o0[min(t + 256 * i, e.w)];
    };
    const float su = 1.0f / 1048576.0f;
    auto step = [&](int d, uint32_t (&rr)[RPT], uint32_t (&orr)[OPT]) __attribute__((always_inline)) {
        const int4 e = dp[d];
        __syncthreads();   // the previous detector done with rb / pb (and v written)
#pragma unroll
        for (int i = 0; i < RPT; ++i) rb[t + 256 * i] = rr[i];
#pragma unroll
        for (int i = 0; i < OPT; ++i) pb[t + 256 * i] = orr[i];
        __syncthreads();
        if (d + 2 < d1) fetch(d + 2, rr, orr);
        const int nb = (e.w + 63) >> 6;   // blocks of 64 sorted positions
        float own[KB], up[KB];
        int row[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            const int blk = 4 * k + ((k & 1) ? 3 - wv : wv);
            float sx = 0.0f, st = 0.0f;
            row[k] = -1;
            if (blk < nb) {
                const int p = 64 * blk + lane;
                if (p < e.w) {
                    const uint32_t a = pb[p];
                    const int b = a & 0xffffu, c = pb[p + 1] & 0xffffu;
                    row[k] = a >> 16;
                    for (int k0 = b; k0 < c; k0 += C) {
                        const int rem = c - k0;
                        uint32_t q[C];
                        float val[C];
#pragma unroll
                        for (int j = 0; j < C; ++j) q[j] = rb[k0 + j];
#pragma unroll
                        for (int j = 0; j < C; ++j) val[j] = v[q[j] & 4095u];
#pragma unroll
                        for (int j = 0; j < C; ++j) {
                            const float x = j < rem ? val[j] : 0.0f;
                            sx += x;
                            st = fmaf((float)(q[j] >> 12), x, st);
                        }
                    }
                }
            }
            up[k] = st * su;
            own[k] = sx - up[k];
        }
        __syncthreads();   // every record read: rb takes the row parts
#pragma unroll
        for (int k = 0; k < KB; ++k)
            if (row[k] >= 0) {
                rb[row[k]] = __float_as_uint(own[k]);
                rb[2048 + row[k]] = __float_as_uint(up[k]);
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = t + 256 * i;
            if (r <= e.w)
                part[e.y + r] = (r < e.w ? __uint_as_float(rb[r]) : 0.0f) + (r > 0 ? __uint_as_float(rb[2047 + r]) : 0.0f);
        }
    };
    if (d0 < d1) fetch(d0, ra, oa);
    if (d0 + 1 < d1) fetch(d0 + 1, rb2, ob2);
    for (int d = d0; d < d1; d += 2) {
        step(d, ra, oa);
        if (d + 1 < d1) step(d + 1, rb2, ob2);
    }
}

// PASS1 < 0: k_fwd_tile1<G>; PASS1 > 0: k_fwd_tile1r<G, PASS1>; PASS1 <= -2: k_fwd_tile1s<G, -PASS1 - 1>
template <int G, int PASS1>
static void fwd_tile1(Plan *p, const float *imgI, hipStream_t s)
{
    hipLaunchKernelGGL(k_img_to_slots, dim3((p->Pd + 255) / 256), dim3(256), 0, s, p->d_pix, imgI, p->d_img_slot,
                       p->Pd);
    const int ntz = (p->Nz + 15) / 16, ntx = (p->Nx + 15) / 16, nblk = p->tf_nbz * p->tf_nbx;
    if constexpr (PASS1 <= -2)
        hipLaunchKernelGGL((k_fwd_tile1s<G, -PASS1 - 1>), dim3(nblk, (p->D + G - 1) / G), dim3(256), 0, s,
                           (const float *)p->d_img_slot, p->d_tfsrec, p->d_tfsro, p->d_tfdesc, p->d_tfpart, p->tf_nbz,
                           ntz, ntx, p->D);
    else if constexpr (PASS1 < 0)
        hipLaunchKernelGGL(k_fwd_tile1<G>, dim3(nblk, (p->D + G - 1) / G), dim3(256), 0, s,
                           (const float *)p->d_img_slot, p->d_tfrec, p->d_tfroff, p->d_tfdesc, p->d_tfpart, p->tf_nbz,
                           ntz, ntx, p->D);
    else
        hipLaunchKernelGGL((k_fwd_tile1r<G, PASS1>), dim3(nblk, (p->D + G - 1) / G), dim3(256), 0, s,
                           (const float *)p->d_img_slot, p->d_tfrec, p->d_tfroff, p->d_tfdesc, p->d_tfpart, p->tf_nbz,
                           ntz, ntx, p->D);
    hipLaunchKernelGGL(k_fwd_tile1_sum, dim3((p->geom.NB + 1023) / 1024, p->D), dim3(256), 0, s,
                       (const float *)p->d_tfpart, p->d_tfdesc, p->d_bins, p->geom.NB, nblk, p->D);
}

template <int NG>
static void fwd_pull_b1(Plan *p, const float *imgI, hipStream_t s)
{
    const int nblk = (p->geom.NB + 256 * NG - 1) / (256 * NG);   // 4 waves x 64 NG bins per workgroup
    const int xm = p->D % 8 == 0 ? p->D / 8 : 0;
    const dim3 grid = xm ? dim3(nblk * p->D) : dim3(nblk, p->D);
    const int2 *rec = p->d_rec;
    if (p->d_rec1 && p->d_img_slot) {   // slot-indexed records over the image in slot order
        hipLaunchKernelGGL(k_img_to_slots, dim3((p->Pd + 255) / 256), dim3(256), 0, s, p->d_pix, imgI,
                           p->d_img_slot, p->Pd);
        rec = p->d_rec1;
        imgI = p->d_img_slot;
    }
    hipLaunchKernelGGL(k_fwd_pull_b1<NG>, grid, dim3(256), 0, s, rec, p->d_rec_off, imgI, p->d_bins, p->geom.NB, xm);
}

template <int BC>
static void fwd_pull(Plan *p, const float *imgI, int Bp, hipStream_t s)
{
    dim3 grid((p->geom.NB + 255) / 256, p->D, Bp / BC);
    hipLaunchKernelGGL(k_fwd_pull<BC>, grid, dim3(256), 0, s, p->d_rec, p->d_rec_off, imgI, p->d_bins, p->geom.NB, Bp);
}

// xcd: XCD-aware workgroup order (needs D % 8 == 0): a 1-D grid of (detector,
// bin block) pairs, D / 8 detectors per XCD passed as xcd_map (0 = plain order)
template <int LB, int U>
static void fwd_pull_bl(Plan *p, const float *imgI, int Bp, bool xcd, hipStream_t s)
{
    constexpr int RPW = 64 / LB;
    const int ngrp = (p->geom.NB + kPullRange - 1) / kPullRange;
    const int nblk = (ngrp + 4 * RPW - 1) / (4 * RPW);
    if (xcd && p->D % 8 == 0) {
        dim3 grid(nblk * p->D, 1, (Bp + LB - 1) / LB);
        hipLaunchKernelGGL((k_fwd_pull_bl<LB, U>), grid, dim3(256), 0, s, p->d_rec, p->d_rec_off, imgI, p->d_bins,
                           p->geom.NB, Bp, kPullRange, p->D / 8);
        return;
    }
    dim3 grid(nblk, p->D, (Bp + LB - 1) / LB);
    hipLaunchKernelGGL((k_fwd_pull_bl<LB, U>), grid, dim3(256), 0, s, p->d_rec, p->d_rec_off, imgI, p->d_bins,
                       p->geom.NB, Bp, kPullRange, 0);
}

template <int U>
static void fwd_pull_wu(Plan *p, const float *imgI, int Bp, bool xcd, hipStream_t s, int JR = kPullRange)
{
    const int ngrp = (p->geom.NB + JR - 1) / JR;
    const int nblk = (ngrp + 3) / 4;
    if (xcd && p->D % 8 == 0) {
        dim3 grid(nblk * p->D, 1, (Bp + 63) / 64);
        hipLaunchKernelGGL((k_fwd_pull_wu<U>), grid, dim3(256), 0, s, p->d_rec, p->d_rec_off, imgI, p->d_bins,
                           p->geom.NB, Bp, JR, p->D / 8);
        return;
    }
    dim3 grid(nblk, p->D, (Bp + 63) / 64);
    hipLaunchKernelGGL((k_fwd_pull_wu<U>), grid, dim3(256), 0, s, p->d_rec, p->d_rec_off, imgI, p->d_bins,
                       p->geom.NB, Bp, JR, 0);
}

template <int U>
static void fwd_pull_wt(Plan *p, const float *imgI, int Bp, hipStream_t s, int JR)
{
    const int ngrp = (p->geom.NB + JR - 1) / JR;
    const int nblk = (ngrp + 3) / 4;
    const int bytes = (int)(sizeof(float) * (size_t)p->P * Bp);
    const int xm = p->D % 8 == 0 ? p->D / 8 : 0;
    const dim3 grid = xm ? dim3(nblk * p->D, 1, (Bp + 63) / 64) : dim3(nblk, p->D, (Bp + 63) / 64);
    hipLaunchKernelGGL((k_fwd_pull_wt<U>), grid, dim3(256), 0, s, p->d_rec, p->d_rec_off, imgI, p->d_bins,
                       p->geom.NB, Bp, JR, xm, bytes);
}

static int launch_pull(Plan *p, const float *imgI, int Bp, hipStream_t s)
{
    // one sinogram: the tile form over sorted rows (2 detectors per workgroup) when its tables exist
    const int dflt = (Bp == 1 && pull_choice_ok(p, 11, Bp)) ? 11 : default_pull_choice(Bp);
    int c = (p->tune.Bp == Bp && p->tune.scatter_bc != 0) ? p->tune.scatter_bc : dflt;
    if (!pull_choice_ok(p, c, Bp)) c = dflt;
    switch (c) {
    case 1: fwd_pull<1>(p, imgI, Bp, s); break;
    case 2: fwd_pull_b1<1>(p, imgI, s); break;
    case 5: fwd_tile1<8, -1>(p, imgI, s); break;
    case 6: fwd_tile1<4, -1>(p, imgI, s); break;
    case 7: fwd_tile1<2, -1>(p, imgI, s); break;
    case 9: fwd_tile1<2, 4>(p, imgI, s); break;
    case 10: fwd_tile1<4, 4>(p, imgI, s); break;
    case 11: fwd_tile1<2, -5>(p, imgI, s); break;
    case 12: fwd_tile1<2, -3>(p, imgI, s); break;
    case 13: fwd_tile1<4, -3>(p, imgI, s); break;
    case 3: fwd_pull_b1<4>(p, imgI, s); break;
    case 4: fwd_pull<4>(p, imgI, Bp, s); break;
    case 8: fwd_pull<8>(p, imgI, Bp, s); break;
    case 16: fwd_pull<16>(p, imgI, Bp, s); break;
    case -16: fwd_pull_bl<16, 8>(p, imgI, Bp, false, s); break;
    case -32: fwd_pull_bl<32, 8>(p, imgI, Bp, false, s); break;
    case -64: fwd_pull_bl<64, 8>(p, imgI, Bp, false, s); break;
    case -116: fwd_pull_bl<16, 8>(p, imgI, Bp, true, s); break;
    case -132: fwd_pull_bl<32, 8>(p, imgI, Bp, true, s); break;
    case -164: fwd_pull_bl<64, 8>(p, imgI, Bp, true, s); break;
    case -216: fwd_pull_bl<16, 16>(p, imgI, Bp, true, s); break;
    case -232: fwd_pull_bl<32, 16>(p, imgI, Bp, true, s); break;
    case -264: fwd_pull_bl<64, 16>(p, imgI, Bp, true, s); break;
    case -364: fwd_pull_wu<8>(p, imgI, Bp, false, s); break;
    case -464: fwd_pull_wu<8>(p, imgI, Bp, true, s); break;
    case -564: fwd_pull_wu<16>(p, imgI, Bp, true, s); break;
    case -664: fwd_pull_wu<16>(p, imgI, Bp, true, s, 16); break;
    case -764: fwd_pull_wu<16>(p, imgI, Bp, true, s, 64); break;
    case -864: fwd_pull_wu<8>(p, imgI, Bp, true, s, 16); break;
    case -964: fwd_pull_wt<16>(p, imgI, Bp, s, 16); break;
    case -1064: fwd_pull_wt<8>(p, imgI, Bp, s, 16); break;
    default: fwd_pull_wt<16>(p, imgI, Bp, s, 32); break;
    }
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

// whether the forward pairs stage writes the bins themselves (pull forms) rather than
// overlapping bucket windows
static bool forward_writes_bins(const Plan *p, int Bp) { return p->pull || is_sir_pull(scatter_choice(p, Bp)); }

template <int TS, int U>
static void sir_pull_bl(Plan *p, const float *imgI, int Bp, hipStream_t s)
{
    const int span = p->geom.nT + 2 * p->geom.K;
    const int nblk = (span + 1 + 4 * TS - 1) / (4 * TS);   // 4 waves x TS samples per workgroup
    hipLaunchKernelGGL((k_fwd_sir_pull_bl<TS, U>), dim3(nblk, p->D, (Bp + 63) / 64), dim3(256), 0, s, p->d_srec,
                       p->d_srec_off, imgI, p->d_sirw, p->d_bins, p->geom.NB, span, p->sp_lo, p->sp_n, p->sir_maxlen,
                       Bp);
}

static int launch_scatter(Plan *p, const float *imgI, int B, int Bp, hipStream_t s)
{
    if (p->pull) return launch_pull(p, imgI, Bp, s);
    if (const int sc = scatter_choice(p, Bp); is_sir_pull(sc)) {
        if (Bp == 1) {
            const int span = p->geom.nT + 2 * p->geom.K;
            const int nblk = (span + 1 + 59) / 60;   // 4 waves x 15 samples per workgroup
            hipLaunchKernelGGL(k_fwd_sir_pull1, dim3(nblk, p->D), dim3(256), 0, s, p->d_srec, p->d_srec_off, imgI,
                               p->d_sirw, p->d_bins, p->geom.NB, span, p->sp_lo, p->sp_n, p->sir_maxlen);
        } else if (sc == kSirBand) {
            constexpr int TS = kSirBandTS;
            const int span = p->geom.nT + 2 * p->geom.K;
            hipLaunchKernelGGL((k_fwd_sir_band<TS>), dim3((p->sband_nwave + 3) / 4, p->D, (Bp + 63) / 64), dim3(256),
                               0, s, p->d_srec, p->d_srec_off, imgI, p->d_sirw, p->d_sband, span, p->sp_lo, p->sp_n,
                               p->sband_nfb, p->sband_nwave, Bp);
            hipLaunchKernelGGL((k_fwd_sir_band_sum<TS>),
                               dim3((unsigned)(((size_t)p->geom.NB * Bp + 255) / 256), p->D), dim3(256), 0, s,
                               (const float *)p->d_sband, p->d_bins, p->geom.NB, p->sp_lo, p->sband_nfb,
                               p->sband_nwave, Bp);
        } else if (sc == kSirPull + 1) sir_pull_bl<16, 4>(p, imgI, Bp, s);
        else if (sc == kSirPull + 2) sir_pull_bl<16, 2>(p, imgI, Bp, s);
        else if (sc == kSirPull + 3) sir_pull_bl<32, 1>(p, imgI, Bp, s);
        else sir_pull_bl<24, 2>(p, imgI, Bp, s);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    const int WLp = p->WL | 1;
    if (p->WL < p->Wb + 1 || p->lo_off > 0 || p->hi_off < 1 || p->WL != p->Wb + p->hi_off - p->lo_off)
        return set_error(MBREC_ERR_ARG, "mbrec:internal", "scatter window not set up (WL=%d)", p->WL);
    const int ch = scatter_choice(p, Bp);
    if (ch < 0) {
        if (ch == -64) scatter_bl<64>(p, imgI, B, Bp, s);
        else if (ch == -32) scatter_bl<32>(p, imgI, B, Bp, s);
        else scatter_bl<16>(p, imgI, B, Bp, s);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    const int Bc = ch;
    if ((size_t)Bc * WLp * sizeof(float) > kLdsBudget)
        return set_error(MBREC_ERR_ARG, "mbrec:size", "bucket window too long for LDS (WL=%d)", p->WL);
    dim3 grid(p->D, p->nbk, (B + Bc - 1) / Bc);
    const size_t lds = sizeof(float) * (size_t)Bc * WLp;
    if (p->use_sir)
        hipLaunchKernelGGL((k_scatter_bk<true, false>), grid, dim3(256), lds, s, p->d_rec, p->d_rec_off, p->d_rec_k,
                           imgI, p->d_hpart, p->d_sir, p->d_swin, p->Ls, p->geom, p->Wb, p->nbk, p->lo_off, p->WL,
                           WLp, Bp, B, Bc);
    else if (Bp % 4 == 0)
        hipLaunchKernelGGL((k_scatter_bk<false, true>), grid, dim3(256), lds, s, p->d_rec, p->d_rec_off, p->d_rec_k,
                           imgI, p->d_hpart, p->d_sir, p->d_swin, p->Ls, p->geom, p->Wb, p->nbk, p->lo_off, p->WL,
                           WLp, Bp, B, Bc);
    else
        hipLaunchKernelGGL((k_scatter_bk<false, false>), grid, dim3(256), lds, s, p->d_rec, p->d_rec_off, p->d_rec_k,
                           imgI, p->d_hpart, p->d_sir, p->d_swin, p->Ls, p->geom, p->Wb, p->nbk, p->lo_off, p->WL,
                           WLp, Bp, B, Bc);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

template <int BC, int TY, int NW>
static void bins_to_y_bc_nw(Plan *p, float *y, int B, int Bp, float scale, hipStream_t s, const int2 *win)
{
    const int nj = 3 * (TY + p->taps.nw) + 3;
    const size_t lds = sizeof(float) * (3 * MBREC_MAX_TAPS + (size_t)BC * (nj | 1) + 2 * BC + 2);
    dim3 grid((p->nT + TY - 1) / TY, p->D, (Bp + BC - 1) / BC);
    if (forward_writes_bins(p, Bp))
        hipLaunchKernelGGL((k_bins_to_y_bc<BC, TY, true, NW>), grid, dim3(256), lds, s, p->d_bins, y, p->taps,
                           p->geom, p->Wb, p->nbk, p->lo_off, p->WL, Bp, p->D, B, scale, win);
    else
        hipLaunchKernelGGL((k_bins_to_y_bc<BC, TY, false, NW>), grid, dim3(256), lds, s, p->d_hpart, y, p->taps,
                           p->geom, p->Wb, p->nbk, p->lo_off, p->WL, Bp, p->D, B, scale, win);
}

template <int BC, int TY>
static void bins_to_y_bc(Plan *p, float *y, int B, int Bp, float scale, hipStream_t s, const int2 *win)
{
    switch (p->taps.nw) {   // 2 floor(R) + 3 taps per phase
    case 5: bins_to_y_bc_nw<BC, TY, 5>(p, y, B, Bp, scale, s, win); break;
    case 7: bins_to_y_bc_nw<BC, TY, 7>(p, y, B, Bp, scale, s, win); break;
    case 9: bins_to_y_bc_nw<BC, TY, 9>(p, y, B, Bp, scale, s, win); break;
    default: bins_to_y_bc_nw<BC, TY, 0>(p, y, B, Bp, scale, s, win); break;
    }
}

static int launch_bins_to_y(Plan *p, float *y, int B, int Bp, float scale, hipStream_t s,
                            const int2 *win = nullptr)
{
    if (Bp >= 16) {
        // batches of 64 and more: 32 sinograms x 64 samples per workgroup (28 KB of LDS, five
        // workgroups per CU: 113 us at 64 x 400^2 against 134 us for 64 x 64, 56 KB, two per CU)
        if (Bp >= 64) bins_to_y_bc<32, 64>(p, y, B, Bp, scale, s, win);
        else if (Bp >= 32) bins_to_y_bc<32, 128>(p, y, B, Bp, scale, s, win);
        else bins_to_y_bc<16, 128>(p, y, B, Bp, scale, s, win);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    const int nj = 3 * (kYTile + p->taps.nw) + 3;
    // LDS chunk of sinograms: at most kHatLds bytes, so several workgroups share a CU
    const int Bc = std::max(1, std::min(B, (int)(kHatLds / (sizeof(float) * (size_t)(nj | 1)))));
    dim3 grid((p->nT + kYTile - 1) / kYTile, p->D, (B + Bc - 1) / Bc);
    const size_t lds = sizeof(float) * (size_t)Bc * (nj | 1);
    if (forward_writes_bins(p, Bp)) {
        switch (p->taps.nw) {   // 2 floor(R) + 3 taps per phase
        case 5:
            hipLaunchKernelGGL((k_bins_to_y<true, 5>), grid, dim3(256), lds, s, p->d_bins, y, p->taps, p->geom, p->Wb,
                               p->nbk, p->lo_off, p->WL, Bp, p->D, B, Bc, scale, win);
            break;
        case 7:
            hipLaunchKernelGGL((k_bins_to_y<true, 7>), grid, dim3(256), lds, s, p->d_bins, y, p->taps, p->geom, p->Wb,
                               p->nbk, p->lo_off, p->WL, Bp, p->D, B, Bc, scale, win);
            break;
        case 9:
            hipLaunchKernelGGL((k_bins_to_y<true, 9>), grid, dim3(256), lds, s, p->d_bins, y, p->taps, p->geom, p->Wb,
                               p->nbk, p->lo_off, p->WL, Bp, p->D, B, Bc, scale, win);
            break;
        default:
            hipLaunchKernelGGL((k_bins_to_y<true>), grid, dim3(256), lds, s, p->d_bins, y, p->taps, p->geom, p->Wb,
                               p->nbk, p->lo_off, p->WL, Bp, p->D, B, Bc, scale, win);
            break;
        }
    } else
        hipLaunchKernelGGL(k_bins_to_y<false>, grid, dim3(256), sizeof(float) * (size_t)Bc * (nj | 1), s, p->d_hpart,
                           y, p->taps, p->geom, p->Wb, p->nbk, p->lo_off, p->WL, Bp, p->D, B, Bc, scale, win);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

static int launch_fir(Plan *p, const float *x, float *out, int B, bool corr, float scale, hipStream_t s,
                      const int2 *win = nullptr)
{
    const int L = p->Lir;
    const int off = L / 2;            // odd L: the same offset serves the convolution and the flipped-IR correlation
    const int Lp = fir16_padded_taps(L);
    const size_t lds16 = 2 * sizeof(float) * (size_t)fir16_window(Lp);
    if (!p->ir_per_det && L <= 4 * kFmNS - 15 && p->d_ir64 && p->fir_mfma) {
        const int rows = B * p->D;
        const dim3 grid((p->nT + kFmTT - 1) / kFmTT, (rows + 15) / 16);
        hipLaunchKernelGGL(k_fir_mfma, grid, dim3(256), 0, s, x, out, corr ? p->d_irf64 : p->d_ir64, L, off, p->D,
                           p->nT, rows, scale, win);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    if (lds16 <= kLdsBudget && p->d_ir64) {
        const int rows = B * p->D, ntile = (p->nT + kFirSpan - 1) / kFirSpan;
        const size_t nblk = ((size_t)rows * ntile + 1) / 2;
        // few tiles (fewer than two per CU): one tile per workgroup, the taps split between its halves
        const size_t ldsp = sizeof(float) * (size_t)fir16_window(Lp) + sizeof(double) * 128 * kFirOut;
        if ((size_t)rows * ntile < 512 && ldsp <= kLdsBudget) {
            const dim3 gp((unsigned)((size_t)rows * ntile));
            if (Lp == 208)
                hipLaunchKernelGGL((k_fir16<208, true>), gp, dim3(256), ldsp, s, x, out, corr ? p->d_irf64 : p->d_ir64,
                                   L, p->ir_per_det, off, p->D, p->nT, rows, scale, win);
            else
                hipLaunchKernelGGL((k_fir16<0, true>), gp, dim3(256), ldsp, s, x, out, corr ? p->d_irf64 : p->d_ir64,
                                   L, p->ir_per_det, off, p->D, p->nT, rows, scale, win);
            MBREC_HIP(hipGetLastError());
            return MBREC_OK;
        }
        if (Lp == 208)   // the 201-tap IR of the default IR design
            hipLaunchKernelGGL(k_fir16<208>, dim3((unsigned)nblk), dim3(256), lds16, s, x, out,
                               corr ? p->d_irf64 : p->d_ir64, L, p->ir_per_det, off, p->D, p->nT, rows, scale, win);
        else
            hipLaunchKernelGGL(k_fir16<0>, dim3((unsigned)nblk), dim3(256), lds16, s, x, out,
                               corr ? p->d_irf64 : p->d_ir64, L, p->ir_per_det, off, p->D, p->nT, rows, scale, win);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    dim3 grid((p->nT + kFirTile - 1) / kFirTile, B * p->D);
    const size_t lds = sizeof(float) * (size_t)(kFirTile + 2 * L + 2);
    hipLaunchKernelGGL(k_fir, grid, dim3(256), lds, s, x, out, corr ? p->d_irf : p->d_ir, L, p->ir_per_det, off,
                       p->D, p->nT, scale, win);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

template <int BC>
static void z_to_bins_bc(Plan *p, const float *z, int B, int Bp, hipStream_t s, const int2 *win)
{
    const int nz = kBinTile / 3 + p->taps.nw + 2;
    const size_t lds = sizeof(float) * (3 * MBREC_MAX_TAPS + (size_t)BC * (nz | 1) + 2 * BC + 2);
    dim3 grid((p->geom.NB + kBinTile - 1) / kBinTile, p->D, (Bp + BC - 1) / BC);
    switch (p->taps.nw) {   // 2 floor(R) + 3 taps per phase
    case 5:
        hipLaunchKernelGGL((k_z_to_bins_bc<BC, 5>), grid, dim3(256), lds, s, z, p->d_bins, p->taps, p->geom, p->D, B,
                           Bp, win);
        break;
    case 7:
        hipLaunchKernelGGL((k_z_to_bins_bc<BC, 7>), grid, dim3(256), lds, s, z, p->d_bins, p->taps, p->geom, p->D, B,
                           Bp, win);
        break;
    case 9:
        hipLaunchKernelGGL((k_z_to_bins_bc<BC, 9>), grid, dim3(256), lds, s, z, p->d_bins, p->taps, p->geom, p->D, B,
                           Bp, win);
        break;
    default:
        hipLaunchKernelGGL((k_z_to_bins_bc<BC, 0>), grid, dim3(256), lds, s, z, p->d_bins, p->taps, p->geom, p->D, B,
                           Bp, win);
        break;
    }
}

static int launch_z_to_bins(Plan *p, const float *z, int B, int Bp, hipStream_t s, const int2 *win = nullptr)
{
    if (Bp == 1) {   // one sinogram: the chunk form with a chunk of one (taps in registers, same sums)
        z_to_bins_bc<1>(p, z, B, Bp, s, win);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    if (Bp >= 16) {
        if (Bp >= 64) z_to_bins_bc<64>(p, z, B, Bp, s, win);
        else if (Bp >= 32) z_to_bins_bc<32>(p, z, B, Bp, s, win);
        else z_to_bins_bc<16>(p, z, B, Bp, s, win);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    const int nz = kBinTile / 3 + p->taps.nw + 2;
    const int Bc = std::max(1, std::min(B, (int)(kHatLds / (sizeof(float) * (size_t)(nz | 1)))));
    dim3 grid((p->geom.NB + kBinTile - 1) / kBinTile, p->D, (B + Bc - 1) / Bc);
    hipLaunchKernelGGL(k_z_to_bins, grid, dim3(256), sizeof(float) * (size_t)Bc * (nz | 1), s, z, p->d_bins, p->taps,
                       p->geom, p->D, B, Bp, Bc, win);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

template <int BC, int U>
static void gather_bc(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    dim3 grid(p->Pd / 64, (Bp + BC - 1) / BC);
    if (p->use_sir)
        hipLaunchKernelGGL((k_gather<BC, 1, true>), grid, dim3(256), 0, s, p->d_tq, p->d_pix, p->d_bins, img,
                           p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd, p->P, p->D, B, Bp, scale);
    else
        hipLaunchKernelGGL((k_gather<BC, U, false>), grid, dim3(256), 0, s, p->d_tq, p->d_pix, p->d_bins, img,
                           p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd, p->P, p->D, B, Bp, scale);
}

template <int LB, int U>
static void gather_bl(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    constexpr int RPW = 64 / LB;
    dim3 grid(p->Pd / (4 * RPW), (Bp + LB - 1) / LB);
    if (p->use_sir)
        hipLaunchKernelGGL((k_gather_bl<true, LB, 1>), grid, dim3(256), 0, s, p->d_tq, p->d_pix, p->d_bins, img,
                           p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd, p->P, p->D, B, Bp, scale);
    else
        hipLaunchKernelGGL((k_gather_bl<false, LB, U>), grid, dim3(256), 0, s, p->d_tq, p->d_pix, p->d_bins, img,
                           p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd, p->P, p->D, B, Bp, scale);
}

template <int LB>
static size_t tile_gather_lds(const Plan *p)
{
    return sizeof(float) * (size_t)p->tile_wmax * LB + sizeof(int32_t) * 256;
}

template <int LB>
static void gather_tile(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    dim3 grid(p->Pd / 256, (Bp + LB - 1) / LB);
    hipLaunchKernelGGL(k_gather_tile<LB>, grid, dim3(256), tile_gather_lds<LB>(p), s, p->d_tq, p->d_twin, p->d_pix,
                       p->d_bins, img, p->geom, p->Pd, p->P, p->D, B, Bp, p->tile_wmax, scale);
}

template <int LB, int U>
static void gather_bl_xcd(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    constexpr int RPW = 64 / LB;
    const int nsb = p->Pd / (4 * RPW);
    dim3 grid(nsb * 8 * ((Bp + LB - 1) / LB));
    hipLaunchKernelGGL((k_gather_bl_xcd<LB, U>), grid, dim3(256), 0, s, p->d_tq, p->d_pix, p->d_bins, p->d_gpart,
                       p->geom, p->Pd, p->P, p->D, B, Bp, nsb);
    const size_t n = (size_t)B * p->P;
    hipLaunchKernelGGL(k_sum_parts<8>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                       (const float *)p->d_gpart, img, n, scale);
}

template <int U, bool XG>
static void gather_wu(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    const int nsb = p->Pd / 4, nbc = (Bp + 63) / 64;
    const dim3 grid = XG ? dim3(8 * nsb * nbc) : dim3(nsb, nbc);
    hipLaunchKernelGGL((k_gather_wu<U, XG>), grid, dim3(256), 0, s, p->d_tq, p->d_pix, p->d_bins,
                       XG ? p->d_gpart : img, p->geom, p->Pd, p->P, p->D, B, Bp, nsb, scale);
    if (XG) {
        const size_t n = (size_t)B * p->P;
        hipLaunchKernelGGL(k_sum_parts<8>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int U, bool XG>
static void gather_wt(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    const int nsb = p->Pd / 4, nbc = (Bp + 63) / 64;
    const dim3 grid = XG ? dim3(8 * nsb * nbc) : dim3(nsb, nbc);
    const int bytes = (int)(sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp);
    hipLaunchKernelGGL((k_gather_wt<U, XG>), grid, dim3(256), 0, s, p->d_ptab, p->d_pix, p->d_bins,
                       XG ? p->d_gpart : img, p->Pd, p->P, p->D, B, Bp, nsb, scale, bytes);
    if (XG) {
        const size_t n = (size_t)B * p->P;
        hipLaunchKernelGGL(k_sum_parts<8>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int G, int NCH>
static void gather_ta_n(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    const int nbc = (Bp + 63) / 64;
    const size_t lds = 0;   // static LDS
    const int bytes = (int)(sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp);
    const int tbytes = (int)(sizeof(int2) * (size_t)(p->Pd / 256) * p->D * 256);
    const int nwg = p->d_tord ? p->tile_launch : p->Pd / 256;
    hipLaunchKernelGGL((k_gather_ta<G, NCH>), dim3(nwg, G * nbc), dim3(1024), lds, s, p->d_ttab, p->d_twin, p->d_pix,
                       p->d_bins, G > 1 ? p->d_gpart : img, p->D, p->geom.NB, p->P, B, Bp, scale, bytes, tbytes,
                       (const int32_t *)p->d_tord);
    if (G > 1) {
        const size_t n = (size_t)B * p->P;
        hipLaunchKernelGGL(k_sum_parts<G>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int G>
static void gather_ta(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    switch (p->tile_zr / 64) {   // 64-row rounds of the staged window (validity keeps it in 1..3)
    case 1: gather_ta_n<G, 1>(p, img, B, Bp, scale, s); break;
    case 2: gather_ta_n<G, 2>(p, img, B, Bp, scale, s); break;
    default: gather_ta_n<G, 3>(p, img, B, Bp, scale, s); break;
    }
}

template <int G, int NCH>
static void gather_tas_n(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    const int nbc = (Bp + 63) / 64;
    const int bytes = (int)(sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp);
    const int tbytes = (int)(sizeof(int4) * (size_t)p->D * 256);   // one tile's rows
    hipLaunchKernelGGL((k_gather_tas<G, NCH>), dim3(p->Pd / 256, G * nbc), dim3(1024), 0, s, p->d_stab, p->d_stwin,
                       p->d_pix, p->d_bins, p->d_sirw, G > 1 ? p->d_gpart : img, p->D, p->geom.NB, p->P, B, Bp,
                       scale, bytes, tbytes);
    if (G > 1) {
        const size_t n = (size_t)B * p->P;
        hipLaunchKernelGGL(k_sum_parts<G>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int G>
static void gather_tas(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    switch (p->stile_zr / 64) {   // 64-row rounds of the staged window (validity keeps it in 1..4)
    case 1: gather_tas_n<G, 1>(p, img, B, Bp, scale, s); break;
    case 2: gather_tas_n<G, 2>(p, img, B, Bp, scale, s); break;
    case 3: gather_tas_n<G, 3>(p, img, B, Bp, scale, s); break;
    default: gather_tas_n<G, 4>(p, img, B, Bp, scale, s); break;
    }
}

template <int G>
static void gather_tp1(Plan *p, float *img, float scale, hipStream_t s)
{
    hipLaunchKernelGGL((k_gather_tp1<G>), dim3(p->Pd / 256, G), dim3(256), 0, s, p->d_tq1, p->d_twin, p->d_pix,
                       p->d_bins, G > 1 ? p->d_gpart : img, p->D, p->geom.NB, p->P, scale, p->tile_zr);
    if (G > 1) {
        const size_t n = (size_t)p->P;
        hipLaunchKernelGGL(k_sum_parts<G>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int G, int PF = 2, int DPB = 1>
static void gather_ta1(Plan *p, float *img, float scale, hipStream_t s)
{
    hipLaunchKernelGGL((k_gather_ta1<G, PF, DPB>), dim3(p->Pd / 256, G), dim3(256), 0, s, p->d_tq, p->d_twin, p->d_pix,
                       p->d_bins, G > 1 ? p->d_gpart : img, p->geom, p->D, p->Pd, p->P, scale, p->tile_zr);
    if (G > 1) {
        const size_t n = (size_t)p->P;
        hipLaunchKernelGGL(k_sum_parts<G>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int G>
static void gather_tas1(Plan *p, float *img, float scale, hipStream_t s)
{
    hipLaunchKernelGGL((k_gather_tas1<G>), dim3(p->Pd / 256, G), dim3(256), 0, s, p->d_stab, p->d_stwin, p->d_pix,
                       p->d_bins, p->d_sirw, G > 1 ? p->d_gpart : img, p->D, p->geom.NB, p->P, scale, p->stile_zr);
    if (G > 1) {
        const size_t n = (size_t)p->P;
        hipLaunchKernelGGL(k_sum_parts<G>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

template <int G>
static void gather_tt(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    const int nbc = (Bp + 63) / 64;
    const size_t lds = sizeof(float) * 2 * (size_t)(p->tile_zr + 2) * 64;
    const int bytes = (int)(sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp);
    hipLaunchKernelGGL((k_gather_tt<G>), dim3(p->Pd / 256, G * nbc), dim3(1024), lds, s, p->d_ttab, p->d_twin,
                       p->d_pix, p->d_bins, G > 1 ? p->d_gpart : img, p->D, p->geom.NB, p->P, B, Bp, scale, bytes,
                       p->tile_zr);
    if (G > 1) {
        const size_t n = (size_t)B * p->P;
        hipLaunchKernelGGL(k_sum_parts<G>, dim3(std::min<size_t>(4096, (n + 255) / 256)), dim3(256), 0, s,
                           (const float *)p->d_gpart, img, n, scale);
    }
}

constexpr int kTilePF = 16;   // 16-byte prefetch registers per thread (pipelined tile gather)

template <int LB>
static void gather_tile2(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    dim3 grid(p->Pd / 256, (Bp + LB - 1) / LB);
    const size_t lds = sizeof(float) * (size_t)p->tile_wmax * LB + 2 * sizeof(int32_t) * 256;
    hipLaunchKernelGGL((k_gather_tile2<LB, kTilePF>), grid, dim3(256), lds, s, p->d_tq, p->d_twin, p->d_pix,
                       p->d_bins, img, p->geom, p->Pd, p->P, p->D, B, Bp, p->tile_wmax, scale);
}

// gather variants: record-lane (lb = 0; sinograms per pass BC, detector unroll U),
// batch-lane (lane groups of lb sinograms, detector unroll U) and tile (u = -1:
// LDS-staged bin windows, lane groups of lb sinograms)
struct GatherVariant {
    int bc, u, lb;
};
static const GatherVariant kGatherVariants[] = {{1, 8, 0},   {1, 4, 0},   {4, 4, 0},   {4, 2, 0},  {8, 2, 0},
                                                {8, 1, 0},   {16, 1, 0},  {0, 2, 16},  {0, 2, 32}, {0, 1, 64},
                                                {0, 2, 64},  {0, 4, 64},  {0, -1, 16}, {0, -1, 32}, {0, -1, 64},
                                                {0, -2, 16}, {0, -2, 32}, {0, -2, 64}, {0, -3, 16},
                                                {0, -3, 32}, {0, -3, 64}, {0, -4, 64}, {0, -5, 8},
                                                {0, -5, 16}, {0, -6, 8},  {0, -6, 16}, {0, -7, 8},
                                                {0, -7, 16}, {0, -8, 8},  {0, -8, 16}, {0, -9, 1},
                                                {0, -9, 2},  {0, -9, 4},  {0, -9, 8},  {0, -10, 1},
                                                {0, -10, 2}, {0, -10, 4}, {0, -11, 1}, {0, -11, 2},
                                                {0, -11, 4}, {0, -12, 1}, {0, -12, 2}, {0, -12, 4},
                                                {0, -12, 8}, {0, -13, 1}, {0, -13, 2}, {0, -13, 4},
                                                {0, -13, 8}, {0, -13, 16}, {0, -14, 4}, {0, -14, 8},
                                                {0, -14, 16}, {0, -15, 4}, {0, -15, 8}, {0, -15, 16}};
constexpr int kNumGatherVariants = sizeof(kGatherVariants) / sizeof(kGatherVariants[0]);

static bool gather_variant_ok(const Plan *p, int v, int Bp)
{
    const GatherVariant &g = kGatherVariants[v];
    if (g.u == -15)   // packed tile-major pair table (built when the tile windows fit 256 rows)
        return !p->use_sir && p->d_tq1 && p->d_twin && p->tile_zr >= 64 && p->tile_zr <= 256 && Bp == 1 &&
               p->d_gpart;
    if (g.u == -13 || g.u == -14)   // tile gather for one sinogram without SIR: tile windows of at most 256 rows
        return !p->use_sir && p->d_twin && p->tile_zr >= 64 && p->tile_zr <= 256 && Bp == 1 &&
               (g.lb == 1 || p->d_gpart);
    if (g.u == -12)   // SIR tile gather for one sinogram: the SIR tile tables, G partial images
        return p->use_sir && p->d_stab && p->stile_zr >= 64 && p->stile_zr <= 256 && Bp == 1 &&
               (g.lb == 1 || p->d_gpart);
    if (g.u == -11) {   // SIR tile gather: tables built, window within 4 rounds of 64 rows, 32-bit offsets
        if (!p->use_sir || !p->d_stab || p->stile_zr < 64 || p->stile_zr > 256 || Bp < 16) return false;
        if (sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp >= ((size_t)1 << 31)) return false;
        if (p->D > kTaMaxD) return false;
        return g.lb == 1 || p->d_gpart;
    }
    if (g.u == -10) {   // tile gather with asynchronous staging: three buffers of at most 3 rounds of 64 rows
        if (p->use_sir || !p->d_ttab || Bp < 16 || p->tile_zr > 192) return false;
        if (sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp >= ((size_t)1 << 31)) return false;
        if (sizeof(int2) * (size_t)(p->Pd / 256) * p->D * 256 >= ((size_t)1 << 31)) return false;
        if (p->D > kTaMaxD) return false;
        return g.lb == 1 || p->d_gpart;
    }
    if (g.u == -9) {   // LDS tile gather from the tile table: SIR-free, window fits the registers and LDS
        if (p->use_sir || !p->d_ttab || Bp < 16 || p->tile_zr > kTileRowsMax) return false;
        if (sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp >= ((size_t)1 << 31)) return false;
        return g.lb == 1 || p->d_gpart;
    }
    if (g.u <= -7) {   // pair-table gather (u = -8: XCD-grouped): SIR-free, 32-bit bin offsets
        if (p->use_sir || !p->d_ptab || Bp < 16 || sizeof(float) * ((size_t)p->D * p->geom.NB + 2) * Bp >= ((size_t)1 << 31))
            return false;
        return g.u == -7 || p->d_gpart;
    }
    if (g.u <= -5) {   // wave-uniform gather (u = -6: XCD-grouped): SIR-free, 32-bit bin offsets
        if (p->use_sir || Bp < 16 || (size_t)p->D * p->geom.NB * Bp >= ((size_t)1 << 31)) return false;
        return g.u == -5 || p->d_gpart;
    }
    if (g.u <= -3) {   // XCD-grouped gather: SIR-free plans, partial images
        if (p->use_sir || Bp < 16 || !p->d_gpart) return false;
        int cap = 16;
        while (cap < Bp) cap *= 2;
        return g.lb <= cap;
    }
    if (g.u < 0) {   // tile gather: SIR-free plans, window fits LDS (and the prefetch registers)
        if (p->use_sir || !p->d_twin || Bp < 16 || Bp % 4) return false;
        const size_t lds = sizeof(float) * (size_t)p->tile_wmax * g.lb + 2048;
        if (g.u == -2 && (size_t)p->tile_wmax * (g.lb / 4) > (size_t)256 * kTilePF) return false;
        return lds <= kLdsBudget && g.lb <= std::max(16, Bp);
    }
    if (g.lb) {
        int cap = 16;
        while (cap < Bp) cap *= 2;
        return Bp >= 16 && g.lb <= cap;
    }
    return g.bc == 1 ? Bp == 1 : Bp % g.bc == 0;
}

static int default_gather_variant(const Plan *p, int Bp)
{
    if (Bp == 1)   // the tile forms (SIR in 8 detector groups; without SIR the packed-table form
                   // in 16, then the decoding form in 8), else record-lane
        for (int v : {43, 54, 47, 0})
            if (gather_variant_ok(p, v, Bp)) return v;
    if (p->use_sir && Bp >= 16)
        for (int v : {39, 38, 37})   // SIR tile gather in four detector groups, two, one
            if (gather_variant_ok(p, v, Bp)) return v;
    if (Bp >= 64) {
        // LDS tile gathers in two detector groups (asynchronous, then register staging),
        // else the wave-uniform XCD gather
        for (int v : {35, 31, 24})
            if (gather_variant_ok(p, v, Bp)) return v;
        return 11;
    }
    if (Bp >= 32) return 8;
    if (Bp >= 16) return 7;
    if (Bp % 8 == 0) return 4;
    return 2;
}

static int launch_gather(Plan *p, float *img, int B, int Bp, float scale, hipStream_t s)
{
    int v = (p->tune.Bp == Bp && p->tune.gather >= 0) ? p->tune.gather : default_gather_variant(p, Bp);
    if (!gather_variant_ok(p, v, Bp)) v = default_gather_variant(p, Bp);
    switch (v) {
    case 0: gather_bc<1, 8>(p, img, B, Bp, scale, s); break;
    case 1: gather_bc<1, 4>(p, img, B, Bp, scale, s); break;
    case 2: gather_bc<4, 4>(p, img, B, Bp, scale, s); break;
    case 3: gather_bc<4, 2>(p, img, B, Bp, scale, s); break;
    case 4: gather_bc<8, 2>(p, img, B, Bp, scale, s); break;
    case 5: gather_bc<8, 1>(p, img, B, Bp, scale, s); break;
    case 6: gather_bc<16, 1>(p, img, B, Bp, scale, s); break;
    case 7: gather_bl<16, 2>(p, img, B, Bp, scale, s); break;
    case 8: gather_bl<32, 2>(p, img, B, Bp, scale, s); break;
    case 9: gather_bl<64, 1>(p, img, B, Bp, scale, s); break;
    case 10: gather_bl<64, 2>(p, img, B, Bp, scale, s); break;
    case 11: gather_bl<64, 4>(p, img, B, Bp, scale, s); break;
    case 12: gather_tile<16>(p, img, B, Bp, scale, s); break;
    case 13: gather_tile<32>(p, img, B, Bp, scale, s); break;
    case 14: gather_tile<64>(p, img, B, Bp, scale, s); break;
    case 15: gather_tile2<16>(p, img, B, Bp, scale, s); break;
    case 16: gather_tile2<32>(p, img, B, Bp, scale, s); break;
    case 17: gather_tile2<64>(p, img, B, Bp, scale, s); break;
    case 18: gather_bl_xcd<16, 2>(p, img, B, Bp, scale, s); break;
    case 19: gather_bl_xcd<32, 2>(p, img, B, Bp, scale, s); break;
    case 20: gather_bl_xcd<64, 2>(p, img, B, Bp, scale, s); break;
    case 21: gather_bl_xcd<64, 4>(p, img, B, Bp, scale, s); break;
    case 22: gather_wu<8, false>(p, img, B, Bp, scale, s); break;
    case 23: gather_wu<16, false>(p, img, B, Bp, scale, s); break;
    case 24: gather_wu<8, true>(p, img, B, Bp, scale, s); break;
    case 25: gather_wu<16, true>(p, img, B, Bp, scale, s); break;
    case 26: gather_wt<8, false>(p, img, B, Bp, scale, s); break;
    case 27: gather_wt<16, false>(p, img, B, Bp, scale, s); break;
    case 28: gather_wt<8, true>(p, img, B, Bp, scale, s); break;
    case 29: gather_wt<16, true>(p, img, B, Bp, scale, s); break;
    case 30: gather_tt<1>(p, img, B, Bp, scale, s); break;
    case 31: gather_tt<2>(p, img, B, Bp, scale, s); break;
    case 32: gather_tt<4>(p, img, B, Bp, scale, s); break;
    case 33: gather_tt<8>(p, img, B, Bp, scale, s); break;
    case 34: gather_ta<1>(p, img, B, Bp, scale, s); break;
    case 35: gather_ta<2>(p, img, B, Bp, scale, s); break;
    case 36: gather_ta<4>(p, img, B, Bp, scale, s); break;
    case 37: gather_tas<1>(p, img, B, Bp, scale, s); break;
    case 38: gather_tas<2>(p, img, B, Bp, scale, s); break;
    case 39: gather_tas<4>(p, img, B, Bp, scale, s); break;
    case 40: gather_tas1<1>(p, img, scale, s); break;
    case 41: gather_tas1<2>(p, img, scale, s); break;
    case 42: gather_tas1<4>(p, img, scale, s); break;
    case 43: gather_tas1<8>(p, img, scale, s); break;
    case 44: gather_ta1<1>(p, img, scale, s); break;
    case 45: gather_ta1<2>(p, img, scale, s); break;
    case 46: gather_ta1<4>(p, img, scale, s); break;
    case 47: gather_ta1<8>(p, img, scale, s); break;
    case 48: gather_ta1<16>(p, img, scale, s); break;
    case 49: gather_ta1<4, 4, 2>(p, img, scale, s); break;
    case 50: gather_ta1<8, 4, 2>(p, img, scale, s); break;
    case 51: gather_ta1<16, 4, 2>(p, img, scale, s); break;
    case 52: gather_tp1<4>(p, img, scale, s); break;
    case 53: gather_tp1<8>(p, img, scale, s); break;
    default: gather_tp1<16>(p, img, scale, s); break;
    }
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

template <int KZ, int KX>
static void prefilter_passes(Plan *p, const float *in, float *out, float *tmp, int B, bool transp,
                             bool interleaved_out, int Bp, hipStream_t s)
{
    const size_t tapb = sizeof(float) * 2 * (size_t)(2 * std::max(p->Kz, p->Kx) + 2);
    const size_t ldz = tapb + sizeof(float) * (size_t)kFzColsB * ((p->Nz + 2 * p->Kz + (KZ ? kFiltR : 0)) | 1);
    const size_t ldz0 = tapb + sizeof(float) * (size_t)kFzColsIL * ((p->Nz + 2 * p->Kz) | 1);
    const size_t ldzi = tapb + sizeof(float) * (size_t)kFzColsIL * ((p->Nz + 2 * p->Kz + (KZ ? kFiltR : 0)) | 1);
    const size_t ldx = tapb + sizeof(float) * (size_t)kFxRowsB * (p->Nx + 2 * p->Kx + (KX ? kFiltR : 0));
    const dim3 gx((p->Nz + kFxRowsB - 1) / kFxRowsB, B);
    // one sinogram: 4 z rows per workgroup (100 workgroups at 400^2 instead of 25)
    constexpr int R1 = 4;
    const dim3 gx1((p->Nz + R1 - 1) / R1, 1);
    const size_t ldx1 = tapb + sizeof(float) * (size_t)R1 * (p->Nx + 2 * p->Kx + (KX ? kFiltR : 0));
    const bool onex = B == 1 && KX > 0;
    const int tr = transp ? 1 : 0;
    const dim3 gzp(((size_t)B * p->Nx + kFzColsB - 1) / kFzColsB);
    // one sinogram: 4 columns per workgroup (64 threads each), so that a 400-column image
    // spreads over 100 workgroups instead of 25
    constexpr int C1 = 4;
    const dim3 gz1((p->Nx + C1 - 1) / C1);
    const size_t ldz1 = tapb + sizeof(float) * (size_t)C1 * ((p->Nz + 2 * p->Kz + (KZ ? kFiltR : 0)) | 1);
    const bool one = B == 1 && KZ > 0;
    // float4 stores of the planar z pass: columns of Nz % 4 == 0 samples from a 16-byte aligned base
    auto vec_ok = [&](const float *o) {
        return (KZ > 0 && p->Nz % 4 == 0 && ((uintptr_t)o & 15) == 0) ? 1 : 0;
    };
    if (!transp) {   // forward: x pass then z pass (the z pass writes the interleaved layout)
        if (onex)
            hipLaunchKernelGGL((k_filt_x<KX, R1>), gx1, dim3(256), ldx1, s, in, tmp, p->d_kx, p->Kx, p->Nz, p->Nx, 0);
        else
            hipLaunchKernelGGL((k_filt_x<KX, kFxRowsB>), gx, dim3(256), ldx, s, in, tmp, p->d_kx, p->Kx, p->Nz, p->Nx,
                               0);
        // interleaved output [P][Bp]: for one sinogram it is the planar layout, which the
        // register-blocked planar pass writes with every thread busy (the interleaved pass
        // gives one of its 16 column lanes work there); for batches the interleaved pass,
        // register-blocked (32 us at 64 x 400^2 against 56 us for the runtime-tap loop,
        // whose two LDS reads per FMA bound it)
        if (one)
            hipLaunchKernelGGL((k_filt_z<false, KZ, C1>), gz1, dim3(256), ldz1, s, (const float *)tmp, out, p->d_kz,
                               p->Kz, p->Nz, p->Nx, B, Bp, 0, vec_ok(out));
        else if (interleaved_out)
            hipLaunchKernelGGL((k_filt_z<true, KZ, kFzColsIL>), dim3(p->Nx, (Bp + kFzColsIL - 1) / kFzColsIL),
                               dim3(256), KZ ? ldzi : ldz0, s, (const float *)tmp, out, p->d_kz, p->Kz, p->Nz, p->Nx,
                               B, Bp, 0, 0);
        else
            hipLaunchKernelGGL((k_filt_z<false, KZ, kFzColsB>), gzp, dim3(256), ldz, s, (const float *)tmp, out,
                               p->d_kz, p->Kz, p->Nz, p->Nx, B, Bp, 0, vec_ok(out));
    } else {   // transpose: z^T then x^T
        if (one)
            hipLaunchKernelGGL((k_filt_z<false, KZ, C1>), gz1, dim3(256), ldz1, s, in, tmp, p->d_kz, p->Kz, p->Nz,
                               p->Nx, B, Bp, tr, vec_ok(tmp));
        else
            hipLaunchKernelGGL((k_filt_z<false, KZ, kFzColsB>), gzp, dim3(256), ldz, s, in, tmp, p->d_kz, p->Kz,
                               p->Nz, p->Nx, B, Bp, tr, vec_ok(tmp));
        if (onex)
            hipLaunchKernelGGL((k_filt_x<KX, R1>), gx1, dim3(256), ldx1, s, (const float *)tmp, out, p->d_kx, p->Kx,
                               p->Nz, p->Nx, tr);
        else
            hipLaunchKernelGGL((k_filt_x<KX, kFxRowsB>), gx, dim3(256), ldx, s, (const float *)tmp, out, p->d_kx,
                               p->Kx, p->Nz, p->Nx, tr);
    }
}

static int launch_prefilter(Plan *p, const float *in, float *out, float *tmp, int B, bool transp,
                            bool interleaved_out, int Bp, hipStream_t s)
{
    const size_t tapb = sizeof(float) * 2 * (size_t)(2 * std::max(p->Kz, p->Kx) + 2);
    const size_t ldz = tapb + sizeof(float) * (size_t)kFzCols * ((p->Nz + 2 * p->Kz + kFiltR) | 1);
    const size_t ldx = tapb + sizeof(float) * (size_t)std::max(kFxRows, kFxRowsB) * (p->Nx + 2 * p->Kx + kFiltR);
    if (ldz <= kLdsBudget && ldx <= kLdsBudget) {
        if (transp && interleaved_out)
            return set_error(MBREC_ERR_ARG, "mbrec:internal", "interleaved output of the transposed prefilter");
        // taps per axis known at compile time for the common half-widths (register-blocked passes)
        if (p->Kz == 11 && p->Kx == 11) prefilter_passes<11, 11>(p, in, out, tmp, B, transp, interleaved_out, Bp, s);
        else if (p->Kz == 10 && p->Kx == 10) prefilter_passes<10, 10>(p, in, out, tmp, B, transp, interleaved_out, Bp, s);
        else if (p->Kz == 12 && p->Kx == 12) prefilter_passes<12, 12>(p, in, out, tmp, B, transp, interleaved_out, Bp, s);
        else prefilter_passes<0, 0>(p, in, out, tmp, B, transp, interleaved_out, Bp, s);
        MBREC_HIP(hipGetLastError());
        return MBREC_OK;
    }
    const size_t n = (size_t)p->P * B;
    hipLaunchKernelGGL(k_filt_axis, dim3(nblocks(n)), dim3(256), 0, s, in, tmp, p->d_kz, p->Kz, p->Nz, p->Nx, B, 0,
                       transp ? 1 : 0, 0, Bp);
    hipLaunchKernelGGL(k_filt_axis, dim3(nblocks(n)), dim3(256), 0, s, (const float *)tmp, out, p->d_kx, p->Kx,
                       p->Nz, p->Nx, B, 1, transp ? 1 : 0, interleaved_out ? 1 : 0, Bp);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

// ---------------------------------------------------------------------------
// composite operators
// ---------------------------------------------------------------------------
int op_forward(Plan *p, const float *img, float *sino, int B, bool model, hipStream_t s, const int2 *win)
{
    int rc = ensure_work(p, B);
    if (rc) return rc;
    const int Bp = batch_stride(B);
    if (model && p->Kz >= 0) {
        float *t2 = p->d_img_tmp + (size_t)p->bcap * p->P;
        if ((rc = launch_prefilter(p, img, p->d_imgI, t2, B, false, true, Bp, s))) return rc;
    } else {
        hipLaunchKernelGGL(k_to_interleaved, dim3(nblocks((size_t)p->P * Bp)), dim3(256), 0, s, img, p->d_imgI, p->P,
                           B, Bp);
    }
    if ((rc = launch_scatter(p, p->d_imgI, B, Bp, s))) return rc;
    const float scale = model ? (float)(1.0 / p->sigma) : 1.0f;
    if (model && p->d_ir) {
        if ((rc = launch_bins_to_y(p, p->d_sino_tmp, B, Bp, 1.0f, s, win))) return rc;
        return launch_fir(p, p->d_sino_tmp, sino, B, false, scale, s, win);
    }
    return launch_bins_to_y(p, sino, B, Bp, scale, s, win);
}

int op_transpose(Plan *p, const float *sino, float *img, int B, bool model, hipStream_t s, const int2 *win)
{
    int rc = ensure_work(p, B);
    if (rc) return rc;
    return op_transpose_run(p, sino, img, B, model, s, win);
}

int op_transpose_run(Plan *p, const float *sino, float *img, int B, bool model, hipStream_t s, const int2 *win)
{
    int rc;
    const int Bp = batch_stride(B);
    const float *z = sino;
    if (model && p->d_ir) {
        if ((rc = launch_fir(p, sino, p->d_sino_tmp, B, true, 1.0f, s, win))) return rc;
        z = p->d_sino_tmp;
    }
    if ((rc = launch_z_to_bins(p, z, B, Bp, s, win))) return rc;
    const float scale = model ? (float)(1.0 / p->sigma) : 1.0f;
    if (model && p->Kz >= 0) {
        float *t1 = p->d_img_tmp, *t2 = p->d_img_tmp + (size_t)p->bcap * p->P;
        if ((rc = launch_gather(p, t1, B, Bp, scale, s))) return rc;
        return launch_prefilter(p, t1, img, t2, B, true, false, Bp, s);
    }
    return launch_gather(p, img, B, Bp, scale, s);
}

// ---------------------------------------------------------------------------
// profiling: average duration of one pipeline stage, HIP events on stream s
// (0 scatter, 1 bins->samples, 2 IR filter, 3 samples->bins, 4 gather)
// ---------------------------------------------------------------------------
// Inputs of the stage timer and the autotuner: a dense pattern in [0.25, 1.25) (timings
// of zero inputs would not count any work a kernel skips for zero values)
__global__ __launch_bounds__(256) void k_timing_fill(float *__restrict__ a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        a[i] = 0.25f + (float)((i * 2654435761u) & 0xffff) * (1.0f / 65536.0f);
}

static void timing_fill(float *a, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_timing_fill, dim3(2048), dim3(256), 0, s, a, n);
}

int time_stage(Plan *p, int which, int B, int reps, hipStream_t s, float *ms)
{
    int rc = ensure_work(p, B);
    if (rc) return rc;
    const int Bp = batch_stride(B);
    float *sino = nullptr;
    MBREC_HIP(hipMalloc((void **)&sino, sizeof(float) * (size_t)B * p->D * p->nT));
    timing_fill(sino, (size_t)B * p->D * p->nT, s);
    timing_fill(p->d_imgI, (size_t)p->P * Bp, s);
    timing_fill(p->d_bins, (size_t)p->D * p->geom.NB * Bp, s);
    MBREC_HIP(hipGetLastError());
    auto run = [&]() -> int {
        switch (which) {
        case 0: return launch_scatter(p, p->d_imgI, B, Bp, s);
        case 1: return launch_bins_to_y(p, sino, B, Bp, 1.0f, s);
        case 2: return p->d_ir ? launch_fir(p, sino, p->d_sino_tmp, B, true, 1.0f, s) : MBREC_OK;
        case 3: return launch_z_to_bins(p, sino, B, Bp, s);
        case 4: return launch_gather(p, p->d_img_tmp, B, Bp, 1.0f, s);
        default: return set_error(MBREC_ERR_ARG, "mbrec:args", "unknown stage %d", which);
        }
    };
    hipEvent_t a, b;
    MBREC_HIP(hipEventCreate(&a));
    MBREC_HIP(hipEventCreate(&b));
    if ((rc = run())) return rc;   // warm-up
    MBREC_HIP(hipEventRecord(a, s));
    for (int i = 0; i < reps; ++i)
        if ((rc = run())) return rc;
    MBREC_HIP(hipEventRecord(b, s));
    MBREC_HIP(hipEventSynchronize(b));
    float t = 0.0f;
    MBREC_HIP(hipEventElapsedTime(&t, a, b));
    *ms = t / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    MBREC_HIP(hipFree(sino));
    return MBREC_OK;
}

// ---------------------------------------------------------------------------
// autotuner: time every valid gather variant and scatter batch chunk for a
// batch of B on this plan (HIP events, reps launches each after one warm-up),
// then time the three fastest of each stage again in two alternating rounds and
// keep the fastest; later launches with the same batch stride use them.  The
// default gather runs untimed for >= 20 ms first: an idle GPU's clocks ramp up
// over its first milliseconds of work, and the candidate timed first (the
// default) was otherwise measured slow (batch 64 at 400^2: 0.92 ms chosen over
// the default's 0.85 ms).
// ---------------------------------------------------------------------------
int autotune(Plan *p, int B, int reps, hipStream_t s, int *gather_out, int *scatter_bc_out)
{
    int rc = ensure_work(p, B);
    if (rc) return rc;
    const int Bp = batch_stride(B);
    timing_fill(p->d_imgI, (size_t)p->P * Bp, s);
    timing_fill(p->d_bins, (size_t)p->D * p->geom.NB * Bp, s);
    MBREC_HIP(hipGetLastError());
    hipEvent_t a, b;
    MBREC_HIP(hipEventCreate(&a));
    MBREC_HIP(hipEventCreate(&b));
    // one timed launch first (it also warms up); a candidate already 4x slower than
    // the best so far is not timed further
    auto timeit = [&](auto &&fn, float &ms, float best_ms) -> int {
        int r;
        MBREC_HIP(hipEventRecord(a, s));
        if ((r = fn())) return r;
        MBREC_HIP(hipEventRecord(b, s));
        MBREC_HIP(hipEventSynchronize(b));
        MBREC_HIP(hipEventElapsedTime(&ms, a, b));
        if (ms > 4.0f * best_ms) return MBREC_OK;
        MBREC_HIP(hipEventRecord(a, s));
        for (int i = 0; i < reps; ++i)
            if ((r = fn())) return r;
        MBREC_HIP(hipEventRecord(b, s));
        MBREC_HIP(hipEventSynchronize(b));
        MBREC_HIP(hipEventElapsedTime(&ms, a, b));
        ms /= reps;
        return MBREC_OK;
    };
    // the three fastest of a first pass (ms, choice), timed again in two alternating
    // rounds with every candidate's better round kept; returns the fastest choice
    auto refine = [&](std::vector<std::pair<float, int>> &res, auto &&set, auto &&fn, int &out) -> int {
        if (res.empty()) return MBREC_OK;
        std::sort(res.begin(), res.end());
        if (res.size() > 3) res.resize(3);
        std::vector<float> t(res.size(), 1e30f);
        for (int round = 0; round < 2; ++round)
            for (size_t k = 0; k < res.size(); ++k) {
                set(res[k].second);
                float ms = 0.0f;
                const int r = timeit(fn, ms, 1e30f);
                if (r) return r;
                t[k] = std::min(t[k], ms);
            }
        size_t bk = 0;
        for (size_t k = 1; k < res.size(); ++k)
            if (t[k] < t[bk]) bk = k;
        out = res[bk].second;
        return MBREC_OK;
    };
    const TuneChoice saved = p->tune;
    p->tune.Bp = Bp;
    // gather
    float best = 1e30f;
    int best_v = default_gather_variant(p, Bp);
    std::vector<int> order;   // the default first: slow candidates are then cut short
    if (gather_variant_ok(p, best_v, Bp)) order.push_back(best_v);
    for (int v = 0; v < kNumGatherVariants; ++v)
        if (v != best_v && gather_variant_ok(p, v, Bp)) order.push_back(v);
    const auto gather_fn = [&]() { return launch_gather(p, p->d_img_tmp, B, Bp, 1.0f, s); };
    if (!order.empty()) {   // clock ramp-up: the first candidate untimed for >= 20 ms (at most 400 launches)
        p->tune.gather = order.front();
        float el = 0.0f;
        MBREC_HIP(hipEventRecord(a, s));
        for (int n = 1; n <= 400 && el < 20.0f; ++n) {
            if ((rc = gather_fn())) {
                p->tune = saved;
                return rc;
            }
            if (n % 4 == 0) {
                MBREC_HIP(hipEventRecord(b, s));
                MBREC_HIP(hipEventSynchronize(b));
                MBREC_HIP(hipEventElapsedTime(&el, a, b));
            }
        }
    }
    std::vector<std::pair<float, int>> res;
    for (int v : order) {
        p->tune.gather = v;
        float ms = 0.0f;
        if ((rc = timeit(gather_fn, ms, best))) {
            p->tune = saved;
            return rc;
        }
        res.emplace_back(ms, v);
        if (ms < best) { best = ms; best_v = v; }
    }
    if ((rc = refine(res, [&](int v) { p->tune.gather = v; }, gather_fn, best_v))) {
        p->tune = saved;
        return rc;
    }
    p->tune.gather = best_v;
    // scatter: record-lane LDS batch chunks, then batch-lane lane groups
    std::vector<int> cand;
    if (p->pull) {
        for (int c : {11, 12, 13, 9, 10, 6, 5, 7, 1, 2, 3, 4, 8, 16, -16, -32, -64, -116, -132, -164, -216, -232, -264, -364, -464, -564, -664,
                      -764, -864, -964, -1064, -1164})
            if (pull_choice_ok(p, c, Bp) && !(c == 1 && Bp > 16)) cand.push_back(c);
    } else {   // SIR: the pull forms first
        for (int c = kSirPull; c < kSirPull + kSirPullVariants; ++c)
            if (sir_pull_ok(p, Bp, c)) cand.push_back(c);
        const int mx = scatter_chunk_max(p, Bp);
        for (int bc = (Bp == 1 ? 1 : 4); bc < mx; bc *= 2) cand.push_back(bc);
        cand.push_back(mx);
        for (int lb = 16; lb <= scatter_lb_max(p, Bp); lb *= 2) cand.push_back(-lb);
    }
    best = 1e30f;
    int best_bc = cand.front();
    const auto scatter_fn = [&]() { return launch_scatter(p, p->d_imgI, B, Bp, s); };
    res.clear();
    for (int c : cand) {
        p->tune.scatter_bc = c;
        float ms = 0.0f;
        if ((rc = timeit(scatter_fn, ms, best))) {
            p->tune = saved;
            return rc;
        }
        res.emplace_back(ms, c);
        if (ms < best) { best = ms; best_bc = c; }
    }
    if ((rc = refine(res, [&](int c) { p->tune.scatter_bc = c; }, scatter_fn, best_bc))) {
        p->tune = saved;
        return rc;
    }
    p->tune.scatter_bc = best_bc;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (gather_out) *gather_out = p->tune.gather;
    if (scatter_bc_out) *scatter_bc_out = p->tune.scatter_bc;
    return MBREC_OK;
}

// explicit launch choice (validated): gather variant index, scatter choice (> 0
// record-lane chunk, < 0 batch-lane lane group); -1 / 0 keep the automatic choice
int set_launch(Plan *p, int B, int gather, int scatter)
{
    const int Bp = batch_stride(B);
    if (gather >= 0) {   // variant validity can depend on the workspace (partial images)
        const int rc = ensure_work(p, B);
        if (rc) return rc;
    }
    if (gather >= kNumGatherVariants || (gather >= 0 && !gather_variant_ok(p, gather, Bp)))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "gather variant %d is not valid for a batch of %d", gather, B);
    if (p->pull) {
        if (scatter != 0 && !pull_choice_ok(p, scatter, Bp))
            return set_error(MBREC_ERR_ARG, "mbrec:args", "forward choice %d is not valid for a batch of %d", scatter,
                             B);
    } else if (is_sir_pull(scatter)) {
        if (!sir_pull_ok(p, Bp, scatter))
            return set_error(MBREC_ERR_ARG, "mbrec:args", "SIR pull forward %d is not available for a batch of %d",
                             scatter, B);
    } else if (scatter > 0 && (scatter > scatter_chunk_max(p, Bp) || (Bp > 1 && scatter % 4)))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "scatter chunk %d is not valid for a batch of %d", scatter, B);
    else if (scatter < 0 && (-scatter > scatter_lb_max(p, Bp) || (-scatter != 16 && -scatter != 32 && -scatter != 64)))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "scatter lane group %d is not valid for a batch of %d", -scatter,
                         B);
    p->tune.Bp = Bp;
    p->tune.gather = gather;
    p->tune.scatter_bc = scatter;
    return MBREC_OK;
}

int num_gather_variants() { return kNumGatherVariants; }

}  // namespace mbrec
