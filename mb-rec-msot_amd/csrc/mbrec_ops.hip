// mbrec_ops.hip -- gfx950 kernels of the forward / transpose model operators.
//
// Pipeline (planar device layouts: image [B][P], sinogram [B][D][nT]):
//   forward   img --prefilter--> coeffs --k_scatter--> bin partials [S][B][D][NB]
//                 --k_bins_to_sino (polyphase hat filter, IR convolution, 1/sigma)--> sino
//   transpose sino --k_sino_to_bins (IR correlation, polyphase hat filter)--> bins [B][D][NB]
//                 --k_gather (1 interpolation per pair, Kahan sum, 1/sigma)--> coeffs
//                 --prefilter^T--> img
// The bin formulation is explained in mbrec_internal.h; it is exact, so the
// composite equals MSOTmodel.m:189-257 (A = crop(IR * MEX(F p)) / sigma) up
// to FP32 rounding.
#include "mbrec_internal.h"

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
    const float ph = (m == 0) ? 0.0f : ((m == 1) ? g.phi1 : g.phi2);
    const float iw = (m == 0) ? g.iw0 : ((m == 1) ? g.iw1 : g.iw2);
    r.nn = base + g.K;
    r.m = m;
    r.u = (fr - ph) * iw;
    return r;
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---------------------------------------------------------------------------
// forward scatter: one workgroup per (detector, pixel split, batch chunk);
// the detector's bin row lives in LDS and receives LDS float atomics, then is
// written once (no global atomics).  Partials of the S splits are summed in a
// fixed order by k_bins_to_sino.
// ---------------------------------------------------------------------------
template <int BC, bool SIR>
__global__ __launch_bounds__(256) void k_scatter(
    const int32_t *__restrict__ tq, const int32_t *__restrict__ pix,
    const float *__restrict__ img, float *__restrict__ hpart,
    const int32_t *__restrict__ sidx, const float *__restrict__ sir,
    const int32_t *__restrict__ swin, int Ls, BinGeom g, int Pd, int P, int D, int B,
    int S, int chunk)
{
    extern __shared__ float h[];
    const int d = blockIdx.x, sp = blockIdx.y, b0 = blockIdx.z * BC;
    const int nb = min(BC, B - b0);
    const int NB = g.NB;
    for (int i = threadIdx.x; i < BC * NB; i += blockDim.x) h[i] = 0.0f;
    __syncthreads();
    const int lo = sp * chunk, hi = min(Pd, lo + chunk);
    const int span = g.nT + 2 * g.K;
    const int32_t *tqd = tq + (size_t)d * Pd;
    for (int slot = lo + (int)threadIdx.x; slot < hi; slot += blockDim.x) {
        const int px = pix[slot];
        if (px < 0) continue;
        const Decoded dc = decode_tau(tqd[slot], g);
        float c[BC];
#pragma unroll
        for (int b = 0; b < BC; ++b) c[b] = (b < nb) ? img[(size_t)(b0 + b) * P + px] : 0.0f;
        if (!SIR) {
            if ((unsigned)dc.nn >= (unsigned)span) continue;
            const int j = 3 * dc.nn + dc.m;
#pragma unroll
            for (int b = 0; b < BC; ++b) {
                atomicAdd(&h[b * NB + j], c[b] * (1.0f - dc.u));
                atomicAdd(&h[b * NB + j + 1], c[b] * dc.u);
            }
        } else {
            const int k = sidx[(size_t)d * Pd + slot];
            const int s0 = swin[2 * k], len = swin[2 * k + 1];
            const float *hk = sir + (size_t)k * Ls;
            const int shift = dc.nn - Ls / 2;
            for (int s = s0; s < s0 + len; ++s) {
                const int nn = shift + s;
                if ((unsigned)nn >= (unsigned)span) continue;
                const float wv = hk[s];
                const int j = 3 * nn + dc.m;
#pragma unroll
                for (int b = 0; b < BC; ++b) {
                    const float v = c[b] * wv;
                    atomicAdd(&h[b * NB + j], v * (1.0f - dc.u));
                    atomicAdd(&h[b * NB + j + 1], v * dc.u);
                }
            }
        }
    }
    __syncthreads();
    for (int b = 0; b < nb; ++b) {
        float *dst = hpart + (((size_t)sp * B + b0 + b) * D + d) * NB;
        for (int j = threadIdx.x; j < NB; j += blockDim.x) dst[j] = h[b * NB + j];
    }
}

// ---------------------------------------------------------------------------
// bins -> samples (M^T), then optional IR convolution (MSOTmodel.m:259-263)
// and scale.  One workgroup per (time tile, detector, batch item).
// ---------------------------------------------------------------------------
constexpr int kTimeTile = 512;

__global__ __launch_bounds__(256) void k_bins_to_sino(
    const float *__restrict__ hpart, int S, float *__restrict__ sino, HatTaps taps, BinGeom g,
    const float *__restrict__ ir, int Lir, int ir_per_det, int D, int B, float scale)
{
    extern __shared__ float sm[];
    const int d = blockIdx.y, b = blockIdx.z;
    const int nT = g.nT, NB = g.NB;
    const int t0 = blockIdx.x * kTimeTile, t1 = min(nT, t0 + kTimeTile);
    const int half = Lir / 2;
    int rlo, rhi;
    if (ir) {
        rlo = max(0, t0 + half - (Lir - 1));
        rhi = min(nT - 1, t1 - 1 + half);
    } else {
        rlo = t0;
        rhi = t1 - 1;
    }
    const int kmax = taps.kmin + taps.nw - 1;
    const int jlo = 3 * (rlo - kmax + g.K);
    const int jhi = 3 * (rhi - taps.kmin + g.K) + 2;
    const int nj = jhi - jlo + 1;
    float *hs = sm;
    float *ys = sm + nj;
    float *irs = ys + (rhi - rlo + 1);
    for (int j = threadIdx.x; j < nj; j += blockDim.x) {
        float acc = 0.0f;
        for (int sp = 0; sp < S; ++sp)
            acc += hpart[(((size_t)sp * B + b) * D + d) * NB + jlo + j];
        hs[j] = acc;
    }
    if (ir) {
        const float *irow = ir + (ir_per_det ? (size_t)d * Lir : 0);
        for (int i = threadIdx.x; i < Lir; i += blockDim.x) irs[i] = irow[i];
    }
    __syncthreads();
    for (int r = rlo + (int)threadIdx.x; r <= rhi; r += blockDim.x) {
        float acc = 0.0f;
        for (int m = 0; m < 3; ++m) {
            const int jb = 3 * (r - taps.kmin + g.K) + m - jlo;
            for (int i = 0; i < taps.nw; ++i) acc = fmaf(taps.w[m][i], hs[jb - 3 * i], acc);
        }
        ys[r - rlo] = acc;
    }
    __syncthreads();
    float *out = sino + ((size_t)b * D + d) * nT;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x) {
        float acc;
        if (ir) {
            // y[t] = sum_r ys[r] IR[t + half - r],  r in [t+half-Lir+1, t+half] cap [rlo, rhi]
            const int ra = max(rlo, t + half - (Lir - 1));
            const int rb = min(rhi, t + half);
            acc = 0.0f;
            for (int r = ra; r <= rb; ++r) acc = fmaf(ys[r - rlo], irs[t + half - r], acc);
        } else {
            acc = ys[t - rlo];
        }
        out[t] = scale * acc;
    }
}

// ---------------------------------------------------------------------------
// samples -> bins (M), with optional IR correlation in front
// (MSOTmodel.m:265-269).  One workgroup per (bin tile, detector, batch item).
// ---------------------------------------------------------------------------
constexpr int kBinTile = 1536;   // 512 samples x 3 phases

__global__ __launch_bounds__(256) void k_sino_to_bins(
    const float *__restrict__ sino, float *__restrict__ bins, HatTaps taps, BinGeom g,
    const float *__restrict__ ir, int Lir, int ir_per_det, int D, int B)
{
    extern __shared__ float sm[];
    const int d = blockIdx.y, b = blockIdx.z;
    const int nT = g.nT, NB = g.NB;
    const int j0 = blockIdx.x * kBinTile, j1 = min(NB, j0 + kBinTile);
    const int nlo = j0 / 3 - g.K, nhi = (j1 - 1) / 3 - g.K;
    const int kmax = taps.kmin + taps.nw - 1;
    const int zlo = max(0, nlo + taps.kmin), zhi = min(nT - 1, nhi + kmax);
    const float *srow = sino + ((size_t)b * D + d) * nT;
    float *zs = sm;
    const int nz = max(0, zhi - zlo + 1);
    if (ir) {
        // z[t] = sum_m s[m] IR[off - t + m], m in [t-off, t-off+Lir-1] cap [0, nT-1]
        const int off = Lir - 1 - Lir / 2;
        const int slo = max(0, zlo - off), shi = min(nT - 1, zhi - off + Lir - 1);
        float *ss = zs + nz;
        float *irs = ss + max(0, shi - slo + 1);
        const float *irow = ir + (ir_per_det ? (size_t)d * Lir : 0);
        for (int i = threadIdx.x; i <= shi - slo; i += blockDim.x) ss[i] = srow[slo + i];
        for (int i = threadIdx.x; i < Lir; i += blockDim.x) irs[i] = irow[i];
        __syncthreads();
        for (int t = zlo + (int)threadIdx.x; t <= zhi; t += blockDim.x) {
            const int ma = max(slo, t - off), mb = min(shi, t - off + Lir - 1);
            float acc = 0.0f;
            for (int m = ma; m <= mb; ++m) acc = fmaf(ss[m - slo], irs[off - t + m], acc);
            zs[t - zlo] = acc;
        }
    } else {
        for (int i = threadIdx.x; i < nz; i += blockDim.x) zs[i] = srow[zlo + i];
    }
    __syncthreads();
    float *brow = bins + ((size_t)b * D + d) * NB;
    for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) {
        const int n = j / 3 - g.K, m = j - 3 * (j / 3);
        float acc = 0.0f;
        for (int i = 0; i < taps.nw; ++i) {
            const int r = n + taps.kmin + i;
            if (r >= zlo && r <= zhi) acc = fmaf(taps.w[m][i], zs[r - zlo], acc);
        }
        brow[j] = acc;
    }
}

// ---------------------------------------------------------------------------
// transpose gather: one workgroup per 64-slot pixel tile; its 4 waves split
// the detectors and are reduced through LDS.  FP32 with Kahan compensation.
// ---------------------------------------------------------------------------
template <int BC, bool SIR>
__global__ __launch_bounds__(256) void k_gather(
    const int32_t *__restrict__ tq, const int32_t *__restrict__ pix,
    const float *__restrict__ bins, float *__restrict__ out,
    const int32_t *__restrict__ sidx, const float *__restrict__ sir,
    const int32_t *__restrict__ swin, int Ls, BinGeom g, int Pd, int P, int D, int B, float scale)
{
    __shared__ float red[4][BC][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int slot = blockIdx.x * 64 + lane;
    const int b0 = blockIdx.y * BC;
    const int nb = min(BC, B - b0);
    const int span = g.nT + 2 * g.K;
    const size_t bstride = (size_t)D * g.NB;
    float acc[BC], cmp[BC];
#pragma unroll
    for (int b = 0; b < BC; ++b) { acc[b] = 0.0f; cmp[b] = 0.0f; }
    for (int d = wv; d < D; d += 4) {
        const Decoded dc = decode_tau(tq[(size_t)d * Pd + slot], g);
        const float *gb = bins + (size_t)b0 * bstride + (size_t)d * g.NB;
        if (!SIR) {
            if ((unsigned)dc.nn < (unsigned)span) {
                const int j = 3 * dc.nn + dc.m;
#pragma unroll
                for (int b = 0; b < BC; ++b) {
                    if (b < nb) {
                        const float g0 = gb[b * bstride + j], g1 = gb[b * bstride + j + 1];
                        kahan_add(acc[b], cmp[b], fmaf(dc.u, g1 - g0, g0));
                    }
                }
            }
        } else {
            const int k = sidx[(size_t)d * Pd + slot];
            const int s0 = swin[2 * k], len = swin[2 * k + 1];
            const float *hk = sir + (size_t)k * Ls;
            const int shift = dc.nn - Ls / 2;
            float part[BC];
#pragma unroll
            for (int b = 0; b < BC; ++b) part[b] = 0.0f;
            for (int s = s0; s < s0 + len; ++s) {
                const int nn = shift + s;
                if ((unsigned)nn >= (unsigned)span) continue;
                const float wvl = hk[s];
                const int j = 3 * nn + dc.m;
#pragma unroll
                for (int b = 0; b < BC; ++b) {
                    if (b < nb) {
                        const float g0 = gb[b * bstride + j], g1 = gb[b * bstride + j + 1];
                        part[b] = fmaf(wvl, fmaf(dc.u, g1 - g0, g0), part[b]);
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < BC; ++b) kahan_add(acc[b], cmp[b], part[b]);
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
                if (b < nb) {
                    const float v = (red[0][b][lane] + red[1][b][lane]) + (red[2][b][lane] + red[3][b][lane]);
                    out[(size_t)(b0 + b) * P + px] = scale * v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// separable circular prefilter (Discretization.m:54-72): one pass per axis.
// forward: out[i] = sum_a k[a] in[(i - a) mod N];  transpose: in[(i + a) mod N]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_filt_axis(
    const float *__restrict__ in, float *__restrict__ out, const float *__restrict__ k, int K,
    int Nz, int Nx, int B, int axis, int transp)
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
    float acc = 0.0f;
    for (int a = -K; a <= K; ++a) {
        int j = i0 + sgn * a;
        j %= N;
        if (j < 0) j += N;
        const float v = axis == 0 ? src[(size_t)ix * Nz + j] : src[(size_t)j * Nz + iz];
        acc = fmaf(k[a + K], v, acc);
    }
    out[idx] = acc;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int BC>
static void scatter_bc(Plan *p, const float *img, int B, hipStream_t s, int chunk)
{
    dim3 grid(p->D, p->S, (B + BC - 1) / BC);
    const size_t lds = sizeof(float) * BC * p->geom.NB;
    if (p->use_sir)
        hipLaunchKernelGGL((k_scatter<BC, true>), grid, dim3(256), lds, s, p->d_tq, p->d_pix, img,
                           p->d_bins, p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd,
                           p->P, p->D, B, p->S, chunk);
    else
        hipLaunchKernelGGL((k_scatter<BC, false>), grid, dim3(256), lds, s, p->d_tq, p->d_pix,
                           img, p->d_bins, p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom,
                           p->Pd, p->P, p->D, B, p->S, chunk);
}

int launch_scatter(Plan *p, const float *img, int B, hipStream_t s)
{
    int chunk = (p->Pd + p->S - 1) / p->S;
    chunk = (chunk + 63) / 64 * 64;
    // LDS budget 64 KiB per workgroup: BC * NB floats
    const size_t row = sizeof(float) * p->geom.NB;
    if (row * 2 <= 65536 && B >= 2)
        scatter_bc<2>(p, img, B, s, chunk);
    else if (row <= 65536)
        scatter_bc<1>(p, img, B, s, chunk);
    else
        return set_error(MBREC_ERR_ARG, "mbrec:size", "sinogram too long for one LDS bin row (nT=%d)", p->nT);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

int launch_bins_to_sino(Plan *p, float *sino, int B, bool model, hipStream_t s)
{
    const bool use_ir = model && p->d_ir;
    const int Lir = use_ir ? p->Lir : 0;
    const int tiles = (p->nT + kTimeTile - 1) / kTimeTile;
    const int span_r = kTimeTile + (use_ir ? Lir : 0) + 2;
    const int nj = 3 * (span_r + p->taps.nw + 2) + 3;
    const size_t lds = sizeof(float) * (nj + span_r + Lir + 8);
    dim3 grid(tiles, p->D, B);
    const float scale = model ? (float)(1.0 / p->sigma) : 1.0f;
    hipLaunchKernelGGL(k_bins_to_sino, grid, dim3(256), lds, s, p->d_bins, p->S, sino, p->taps,
                       p->geom, use_ir ? p->d_ir : nullptr, Lir, p->ir_per_det, p->D, B, scale);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

int launch_sino_to_bins(Plan *p, const float *sino, int B, bool model, hipStream_t s)
{
    const bool use_ir = model && p->d_ir;
    const int Lir = use_ir ? p->Lir : 0;
    const int tiles = (p->geom.NB + kBinTile - 1) / kBinTile;
    const int nz = kBinTile / 3 + p->taps.nw + 4;
    const size_t lds = sizeof(float) * (nz + (use_ir ? (nz + 2 * Lir + 4) : 0) + 8);
    dim3 grid(tiles, p->D, B);
    hipLaunchKernelGGL(k_sino_to_bins, grid, dim3(256), lds, s, sino, p->d_bins, p->taps, p->geom,
                       use_ir ? p->d_ir : nullptr, Lir, p->ir_per_det, p->D, B);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

template <int BC>
static void gather_bc(Plan *p, float *img, int B, float scale, hipStream_t s)
{
    dim3 grid(p->Pd / 64, (B + BC - 1) / BC);
    if (p->use_sir)
        hipLaunchKernelGGL((k_gather<BC, true>), grid, dim3(256), 0, s, p->d_tq, p->d_pix,
                           p->d_bins, img, p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd,
                           p->P, p->D, B, scale);
    else
        hipLaunchKernelGGL((k_gather<BC, false>), grid, dim3(256), 0, s, p->d_tq, p->d_pix,
                           p->d_bins, img, p->d_sidx, p->d_sir, p->d_swin, p->Ls, p->geom, p->Pd,
                           p->P, p->D, B, scale);
}

int launch_gather(Plan *p, float *img, int B, float scale, hipStream_t s)
{
    if (B >= 4)
        gather_bc<4>(p, img, B, scale, s);
    else if (B >= 2)
        gather_bc<2>(p, img, B, scale, s);
    else
        gather_bc<1>(p, img, B, scale, s);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

int launch_prefilter(Plan *p, const float *in, float *out, float *tmp, int B, bool transp,
                     hipStream_t s)
{
    const size_t n = (size_t)p->P * B;
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_filt_axis, dim3(blocks), dim3(256), 0, s, in, tmp, p->d_kz, p->Kz, p->Nz,
                       p->Nx, B, 0, transp ? 1 : 0);
    hipLaunchKernelGGL(k_filt_axis, dim3(blocks), dim3(256), 0, s, (const float *)tmp, out,
                       p->d_kx, p->Kx, p->Nz, p->Nx, B, 1, transp ? 1 : 0);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

// ---------------------------------------------------------------------------
// composite operators
// ---------------------------------------------------------------------------
int op_forward(Plan *p, const float *img, float *sino, int B, bool model, hipStream_t s)
{
    int rc = ensure_bins(p, B);
    if (rc) return rc;
    const float *src = img;
    if (model && p->Kz >= 0) {
        if ((rc = ensure_tmp(p, B))) return rc;
        float *t1 = p->d_img_tmp, *t2 = p->d_img_tmp + (size_t)p->tmp_bcap * p->P;
        if ((rc = launch_prefilter(p, img, t1, t2, B, false, s))) return rc;
        src = t1;
    }
    if ((rc = launch_scatter(p, src, B, s))) return rc;
    return launch_bins_to_sino(p, sino, B, model, s);
}

int op_transpose(Plan *p, const float *sino, float *img, int B, bool model, hipStream_t s)
{
    int rc = ensure_bins(p, B);
    if (rc) return rc;
    if ((rc = launch_sino_to_bins(p, sino, B, model, s))) return rc;
    const float scale = model ? (float)(1.0 / p->sigma) : 1.0f;
    if (model && p->Kz >= 0) {
        if ((rc = ensure_tmp(p, B))) return rc;
        float *t1 = p->d_img_tmp, *t2 = p->d_img_tmp + (size_t)p->tmp_bcap * p->P;
        if ((rc = launch_gather(p, t1, B, scale, s))) return rc;
        return launch_prefilter(p, t1, img, t2, B, true, s);
    }
    return launch_gather(p, img, B, scale, s);
}

// ---------------------------------------------------------------------------
// profiling: average duration of one pipeline stage, HIP events on stream s
// ---------------------------------------------------------------------------
int time_stage(Plan *p, int which, int B, int reps, hipStream_t s, float *ms)
{
    int rc = ensure_bins(p, B);
    if (rc) return rc;
    if ((rc = ensure_tmp(p, B))) return rc;
    float *img = p->d_img_tmp;
    float *sino = nullptr;
    MBREC_HIP(hipMalloc((void **)&sino, sizeof(float) * (size_t)B * p->D * p->nT));
    MBREC_HIP(hipMemsetAsync(sino, 0, sizeof(float) * (size_t)B * p->D * p->nT, s));
    MBREC_HIP(hipMemsetAsync(img, 0, sizeof(float) * (size_t)B * p->P, s));
    auto run = [&]() -> int {
        switch (which) {
        case 0: return launch_scatter(p, img, B, s);
        case 1: return launch_bins_to_sino(p, sino, B, true, s);
        case 2: return launch_sino_to_bins(p, sino, B, true, s);
        case 3: return launch_gather(p, img, B, 1.0f, s);
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

}  // namespace mbrec
