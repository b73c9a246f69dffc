// mbrec_unet.hip -- learned warm start: a 4-level convolutional U-Net that maps
// the back-projected image A^T y to an initial NNLS estimate.
//
// No counterpart in the reference (its NNLS loops always start from zero,
// nnls_L2_less.m:28-29); the estimate enters the solver through the initial
// image of the warm-start form of rec_nn_with_L2_reg (rec_nn_with_L2_reg.m:41).
//
// Architecture (base width c, all 3x3 convolutions with zero padding, ReLU):
//   enc1 1->c, c->c | pool | enc2 c->2c, 2c->2c | pool | enc3 2c->4c, 4c->4c | pool
//   bottleneck 4c->8c, 8c->8c | up+cat enc3 | 12c->4c, 4c->4c | up+cat enc2 | 6c->2c, 2c->2c
//   | up+cat enc1 | 3c->c, c->c | 1x1 head c->1 (linear)
// Pooling is a 2x2 average, upsampling nearest-neighbour.  Each image is scaled
// by s = rms(A^T y) before the network and the output is multiplied by s.
//
// Numerics: activations and activation gradients are stored in bfloat16
// (round to nearest even) in pixel-major / channel-minor order; every 3x3
// convolution with >= 32 input channels runs on the matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) as an implicit GEMM over
// (pixels x output channels x 9*input channels); its input gradient is the same
// kernel with the flipped, transposed filter, and its filter gradient an
// implicit GEMM over pixels.  The first convolution (one input channel), the
// 1x1 head, pooling, upsampling, the loss and Adam run on the vector ALUs in
// FP32.  Master weights and Adam moments are FP32.
#include "mbrec_internal.h"

#include <cmath>
#include <algorithm>

namespace mbrec {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
    union { uint32_t u; float f; } c;
    c.u = (uint32_t)h << 16;
    return c.f;
}

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    union { uint32_t u; float f; } c;
    c.f = f;
    uint32_t u = c.u;
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM on the matrix cores.
// y[p][yo + co] = act(bias[co] + sum_{t, ci} x[p + shift(t)][xo + ci] * w[co][t][ci])
// One wave computes 16 pixels x 16*NF output channels; a workgroup 64 pixels.
// mode 0: + bias, ReLU.  mode 1: no bias, multiply by (m[p][mo + co] > 0).
// Requires Cin % 32 == 0, Cout % (16*NF) == 0, strides / offsets % 8 == 0.
// ---------------------------------------------------------------------------
template <int NF>
__global__ __launch_bounds__(256) void k_conv3_mfma(
    const uint16_t *__restrict__ x, int xs, int xo, const uint16_t *__restrict__ w, const float *__restrict__ bias,
    uint16_t *__restrict__ y, int ys, int yo, const uint16_t *__restrict__ m, int ms, int mo, int N, int H, int W,
    int Cin, int Cout, int mode)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int HW = H * W;
    const int NP = N * HW;
    const int p0 = (blockIdx.x * 4 + wv) * 16;
    const int co0 = blockIdx.y * 16 * NF;
    const int r = lane & 15, kq = lane >> 4;
    const int p = p0 + r;
    const bool pv = p < NP;
    int n = 0, h = 0, ww = 0;
    if (pv) {
        n = p / HW;
        const int rem = p - n * HW;
        h = rem / W;
        ww = rem - h * W;
    }
    f32x4 acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = 0; t < 9; ++t) {
        const int hh = h + t / 3 - 1, wq = ww + t % 3 - 1;
        const bool ok = pv && hh >= 0 && hh < H && wq >= 0 && wq < W;
        const uint16_t *xp = x + ((size_t)(n * H + (ok ? hh : 0)) * W + (ok ? wq : 0)) * xs + xo + 8 * kq;
        const uint16_t *wp = w + ((size_t)(co0 + r) * 9 + t) * Cin + 8 * kq;
        for (int c0 = 0; c0 < Cin; c0 += 32) {
            u16x8 av = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if (ok) av = *reinterpret_cast<const u16x8 *>(xp + c0);
            const bf16x8 a = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const u16x8 bv = *reinterpret_cast<const u16x8 *>(wp + (size_t)f * 16 * 9 * Cin + c0);
                acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, bv), acc[f], 0, 0, 0);
            }
        }
    }
    // lane holds rows (pixels) 4*kq + j, column (output channel) r of each 16x16 block
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int co = co0 + 16 * f + r;
        const float b = (mode == 0 && bias) ? bias[co] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int pp = p0 + 4 * kq + j;
            if (pp >= NP) continue;
            float v = acc[f][j] + b;
            if (mode == 0) v = fmaxf(v, 0.0f);
            else if (m && !(bf2f(m[(size_t)pp * ms + mo + co]) > 0.0f)) v = 0.0f;
            y[(size_t)pp * ys + yo + co] = f2bf(v);
        }
    }
}

// ---------------------------------------------------------------------------
// filter gradient of a 3x3 convolution (implicit GEMM over pixels):
// dw[co][t][ci] += sum_p g[p][go + co] * x[p + shift(t)][xo + ci]
// grid (co/32 * ci/32, row groups, 9 taps); each wave walks image rows in
// steps of 32 pixels, 32 x 32 (co x ci) per wave, FP32 atomics at the end.
// ---------------------------------------------------------------------------
constexpr int kWgRowsPerWave = 8;

__global__ __launch_bounds__(256) void k_conv3_wgrad(
    const uint16_t *__restrict__ g, int gs, int go, const uint16_t *__restrict__ x, int xs, int xo,
    float *__restrict__ dw, int N, int H, int W, int Cin, int Cout)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int nci = Cin / 32;
    const int co0 = (blockIdx.x / nci) * 32, ci0 = (blockIdx.x % nci) * 32;
    const int t = blockIdx.z;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int rows = N * H;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int row_begin = (blockIdx.y * 4 + wv) * kWgRowsPerWave;
    for (int rr = 0; rr < kWgRowsPerWave; ++rr) {
        const int row = row_begin + rr;
        if (row >= rows) break;
        const int n = row / H, h = row - n * H;
        const int hh = h + dy;
        if (hh < 0 || hh >= H) continue;
        const uint16_t *grow = g + (size_t)row * W * gs + go;
        const uint16_t *xrow = x + (size_t)(n * H + hh) * W * xs + xo;
        for (int w0 = 0; w0 < W; w0 += 32) {
            u16x8 av[2], bv[2];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int wp = w0 + 8 * kq + j;
                const int wx = wp + dx;
                const bool gv = wp < W;
                const bool xv = gv && wx >= 0 && wx < W;
#pragma unroll
                for (int a = 0; a < 2; ++a) av[a][j] = gv ? grow[(size_t)wp * gs + co0 + 16 * a + r] : (uint16_t)0;
#pragma unroll
                for (int b = 0; b < 2; ++b) bv[b][j] = xv ? xrow[(size_t)wx * xs + ci0 + 16 * b + r] : (uint16_t)0;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[a]),
                                                                        __builtin_bit_cast(bf16x8, bv[b]), acc[a][b],
                                                                        0, 0, 0);
        }
    }
    // lane holds rows (co) 4*kq + j, column (ci) r
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = acc[a][b][j];
                if (v != 0.0f)
                    atomicAdd(&dw[((size_t)(co0 + 16 * a + 4 * kq + j) * 9 + t) * Cin + ci0 + 16 * b + r], v);
            }
}

// bias gradient: db[co] += sum_p g[p][go + co]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ g, int gs, int go, float *__restrict__ db,
                                                   int NP, int Cout, int chunk)
{
    const int p0 = blockIdx.x * chunk, p1 = min(NP, p0 + chunk);
    if (Cout >= 256) {   // one thread per channel (Cout is a multiple of 32)
        for (int co = threadIdx.x; co < Cout; co += 256) {
            float s = 0.0f;
            for (int p = p0; p < p1; ++p) s += bf2f(g[(size_t)p * gs + go + co]);
            atomicAdd(&db[co], s);
        }
        return;
    }
    const int lanes_per_co = 256 / Cout;   // Cout in {32, 64, 128}
    const int co = threadIdx.x % Cout, pg = threadIdx.x / Cout;
    float s = 0.0f;
    for (int p = p0 + pg; p < p1; p += lanes_per_co) s += bf2f(g[(size_t)p * gs + go + co]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    if (pg == 0) {
        for (int k = 1; k < lanes_per_co; ++k) s += red[k * Cout + co];
        atomicAdd(&db[co], s);
    }
}

// ---------------------------------------------------------------------------
// per-image scale s_b = sqrt(mean(in_b^2)) (tiny floor), normalised input in_n = in / s
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_image_scale(const float *__restrict__ in, float *__restrict__ s, int HW)
{
    const float *src = in + (size_t)blockIdx.x * HW;
    double acc = 0.0;
    for (int i = threadIdx.x; i < HW; i += blockDim.x) acc += (double)src[i] * src[i];
    __shared__ double red[1024];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) s[blockIdx.x] = (float)fmax(sqrt(red[0] / HW), 1e-30);
}

__global__ void k_normalize(const float *__restrict__ in, const float *__restrict__ s, float *__restrict__ out, int HW,
                            size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] / s[i / HW];
}

// ---------------------------------------------------------------------------
// first convolution (1 input channel, C outputs), FP32: y[p][co] = relu(b + sum_t w[co][t] x[p+shift(t)])
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv_first(const float *__restrict__ x, const float *__restrict__ w,
                                                    const float *__restrict__ b, uint16_t *__restrict__ y, int N,
                                                    int H, int W, int C)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int HW = H * W;
    if (p >= N * HW) return;
    const int n = p / HW, rem = p - n * HW, h = rem / W, ww = rem - h * W;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int hh = h + t / 3 - 1, wq = ww + t % 3 - 1;
        v[t] = (hh >= 0 && hh < H && wq >= 0 && wq < W) ? x[(size_t)n * HW + hh * W + wq] : 0.0f;
    }
    for (int c0 = 0; c0 < C; c0 += 8) {
        u16x8 o;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int co = c0 + k;
            float a = b[co];
#pragma unroll
            for (int t = 0; t < 9; ++t) a = fmaf(w[co * 9 + t], v[t], a);
            o[k] = f2bf(fmaxf(a, 0.0f));
        }
        *reinterpret_cast<u16x8 *>(y + (size_t)p * C + c0) = o;
    }
}

// filter / bias gradient of the first convolution: dw[co][t] += sum_p g[p][co] x[p+shift(t)], db[co] += sum_p g[p][co]
__global__ __launch_bounds__(256) void k_conv_first_wgrad(const uint16_t *__restrict__ g, const float *__restrict__ x,
                                                          float *__restrict__ dw, float *__restrict__ db, int N, int H,
                                                          int W, int C, int chunk)
{
    const int HW = H * W, NP = N * HW;
    const int co = threadIdx.x % C, pg = threadIdx.x / C, npg = 256 / C;
    const int p0 = blockIdx.x * chunk, p1 = min(NP, p0 + chunk);
    float a[10];
#pragma unroll
    for (int t = 0; t < 10; ++t) a[t] = 0.0f;
    for (int p = p0 + pg; p < p1; p += npg) {
        const float gv = bf2f(g[(size_t)p * C + co]);
        if (gv == 0.0f) continue;
        const int n = p / HW, rem = p - n * HW, h = rem / W, ww = rem - h * W;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int hh = h + t / 3 - 1, wq = ww + t % 3 - 1;
            if (hh >= 0 && hh < H && wq >= 0 && wq < W) a[t] = fmaf(gv, x[(size_t)n * HW + hh * W + wq], a[t]);
        }
        a[9] += gv;
    }
    __shared__ float red[10][256];
#pragma unroll
    for (int t = 0; t < 10; ++t) red[t][threadIdx.x] = a[t];
    __syncthreads();
    if (pg == 0) {
        for (int t = 0; t < 10; ++t) {
            float s = a[t];
            for (int k = 1; k < npg; ++k) s += red[t][k * C + co];
            if (t < 9) atomicAdd(&dw[co * 9 + t], s);
            else atomicAdd(&db[co], s);
        }
    }
}

// ---------------------------------------------------------------------------
// 1x1 head (C -> 1), FP32: on[p] = b + sum_ci w[ci] a[p][ci];  out[p] = on[p] * s[image]
// ---------------------------------------------------------------------------
__global__ void k_head(const uint16_t *__restrict__ a, const float *__restrict__ w, const float *__restrict__ b,
                       const float *__restrict__ s, float *__restrict__ on, float *__restrict__ out, int HW, int NP,
                       int C)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    float acc = b[0];
    for (int c0 = 0; c0 < C; c0 += 8) {
        const u16x8 v = *reinterpret_cast<const u16x8 *>(a + (size_t)p * C + c0);
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = fmaf(w[c0 + k], bf2f(v[k]), acc);
    }
    if (on) on[p] = acc;
    if (out) out[p] = acc * s[p / HW];
}

// relative loss: pw[b] = mean over item b of (target / s_b)^2, then the weights
// wt[b] = mean_b' pw[b'] / pw[b] (1 for an item without target power)
__global__ __launch_bounds__(1024) void k_target_power(const float *__restrict__ target, const float *__restrict__ s,
                                                      float *__restrict__ pw, int HW)
{
    const float *src = target + (size_t)blockIdx.x * HW;
    const double sb = s[blockIdx.x];
    double acc = 0.0;
    for (int i = threadIdx.x; i < HW; i += blockDim.x) {
        const double v = src[i] / sb;
        acc += v * v;
    }
    __shared__ double red[1024];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) pw[blockIdx.x] = (float)(red[0] / HW);
}

__global__ void k_loss_weights(float *__restrict__ wt, int B)
{
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int b = 0; b < B; ++b) m += wt[b];
    m /= B;
    for (int b = 0; b < B; ++b) wt[b] = wt[b] > 0.0f ? (float)(m / wt[b]) : 1.0f;
}

// loss (mean squared error of the normalised output, item b weighted by wt[b] if wt) and
// head backward: d = on - target/s;  loss += wt d^2;  g_on = 2 wt d / NP;
// ga[p][ci] = g_on w[ci] (a > 0);  dw[ci] += g_on a[p][ci];  db += g_on
__global__ __launch_bounds__(256) void k_head_loss_bwd(const uint16_t *__restrict__ a, const float *__restrict__ w,
                                                       const float *__restrict__ on, const float *__restrict__ target,
                                                       const float *__restrict__ s, const float *__restrict__ wt,
                                                       uint16_t *__restrict__ ga, float *__restrict__ dw,
                                                       float *__restrict__ db, double *__restrict__ loss, int HW, int NP,
                                                       int C)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float red[65][256];
    float dwl[64];
    for (int c = 0; c < C; ++c) dwl[c] = 0.0f;
    float gon = 0.0f, l2 = 0.0f;
    if (p < NP) {
        const float d = on[p] - target[p] / s[p / HW];
        const float wi = wt ? wt[p / HW] : 1.0f;
        l2 = wi * d * d;
        gon = 2.0f * wi * d / (float)NP;
        for (int c0 = 0; c0 < C; c0 += 8) {
            const u16x8 v = *reinterpret_cast<const u16x8 *>(a + (size_t)p * C + c0);
            u16x8 o;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float av = bf2f(v[k]);
                o[k] = f2bf(av > 0.0f ? gon * w[c0 + k] : 0.0f);
                dwl[c0 + k] = gon * av;
            }
            *reinterpret_cast<u16x8 *>(ga + (size_t)p * C + c0) = o;
        }
    }
    for (int c = 0; c < C; ++c) red[c][threadIdx.x] = dwl[c];
    red[64][threadIdx.x] = gon;
    __syncthreads();
    // column sums: thread c < C sums channel c, thread 64 the bias
    const int tc = threadIdx.x;
    if (tc < C || tc == 64) {
        float acc = 0.0f;
        for (int k = 0; k < 256; ++k) acc += red[tc][k];
        if (tc == 64) atomicAdd(db, acc);
        else atomicAdd(&dw[tc], acc);
    }
    __syncthreads();
    red[0][threadIdx.x] = l2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int k = 0; k < 256; ++k) acc += red[0][k];
        atomicAdd(loss, acc);
    }
}

// ---------------------------------------------------------------------------
// 2x2 average pooling: y[n][h][w][c] = mean of x[n][2h+i][2w+j][xo + c]
// ---------------------------------------------------------------------------
__global__ void k_pool(const uint16_t *__restrict__ x, int xs, int xo, uint16_t *__restrict__ y, int N, int H, int W,
                       int C)
{
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Ho * Wo * C8) return;
    const int c8 = (int)(i % C8);
    const size_t q = i / C8;
    const int n = (int)(q / ((size_t)Ho * Wo)), rem = (int)(q - (size_t)n * Ho * Wo), h = rem / Wo, w = rem - h * Wo;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const u16x8 v = *reinterpret_cast<const u16x8 *>(
                x + ((size_t)(n * H + 2 * h + a) * W + 2 * w + b) * xs + xo + 8 * c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] += bf2f(v[k]);
        }
    u16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f2bf(0.25f * acc[k]);
    *reinterpret_cast<u16x8 *>(y + q * C + 8 * c8) = o;
}

// nearest 2x upsampling into the first C channels of a concatenation buffer
__global__ void k_upsample(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, int ys, int N, int H, int W,
                           int C)
{
    // x [N][H/2][W/2][C] -> y[n][h][w][0..C)
    const int C8 = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C8) return;
    const int c8 = (int)(i % C8);
    const size_t q = i / C8;
    const int n = (int)(q / ((size_t)H * W)), rem = (int)(q - (size_t)n * H * W), h = rem / W, w = rem - h * W;
    const u16x8 v =
        *reinterpret_cast<const u16x8 *>(x + ((size_t)(n * (H / 2) + h / 2) * (W / 2) + w / 2) * C + 8 * c8);
    *reinterpret_cast<u16x8 *>(y + q * ys + 8 * c8) = v;
}

// upsampling backward: g[n][h][w][c] = (sum of the 4 children of gy[.][0..C)) * (act[n][h][w][c] > 0)
__global__ void k_upsample_bwd(const uint16_t *__restrict__ gy, int gys, const uint16_t *__restrict__ act,
                               uint16_t *__restrict__ g, int N, int Hs, int Ws, int C)
{
    // Hs, Ws: the small (input) resolution
    const int C8 = C / 8, H = 2 * Hs, W = 2 * Ws;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * Hs * Ws * C8) return;
    const int c8 = (int)(i % C8);
    const size_t q = i / C8;
    const int n = (int)(q / ((size_t)Hs * Ws)), rem = (int)(q - (size_t)n * Hs * Ws), h = rem / Ws, w = rem - h * Ws;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const u16x8 v = *reinterpret_cast<const u16x8 *>(
                gy + ((size_t)(n * H + 2 * h + a) * W + 2 * w + b) * gys + 8 * c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] += bf2f(v[k]);
        }
    const u16x8 av = *reinterpret_cast<const u16x8 *>(act + q * C + 8 * c8);
    u16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f2bf(bf2f(av[k]) > 0.0f ? acc[k] : 0.0f);
    *reinterpret_cast<u16x8 *>(g + q * C + 8 * c8) = o;
}

// skip-connection gradient, in place in the concatenation gradient buffer:
// gc[p][so + c] = (gc[p][so + c] + gp[parent(p)][c] / 4) * (cat[p][so + c] > 0)
__global__ void k_skip_grad(uint16_t *__restrict__ gc, const uint16_t *__restrict__ cat, int cs, int so,
                            const uint16_t *__restrict__ gp, int N, int H, int W, int C)
{
    const int C8 = C / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C8) return;
    const int c8 = (int)(i % C8);
    const size_t q = i / C8;
    const int n = (int)(q / ((size_t)H * W)), rem = (int)(q - (size_t)n * H * W), h = rem / W, w = rem - h * W;
    const size_t par = ((size_t)(n * (H / 2) + h / 2) * (W / 2) + w / 2) * C + 8 * c8;
    u16x8 gv = *reinterpret_cast<const u16x8 *>(gc + q * cs + so + 8 * c8);
    const u16x8 av = *reinterpret_cast<const u16x8 *>(cat + q * cs + so + 8 * c8);
    const u16x8 pv = *reinterpret_cast<const u16x8 *>(gp + par);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float v = bf2f(gv[k]) + 0.25f * bf2f(pv[k]);
        gv[k] = f2bf(bf2f(av[k]) > 0.0f ? v : 0.0f);
    }
    *reinterpret_cast<u16x8 *>(gc + q * cs + so + 8 * c8) = gv;
}

// ---------------------------------------------------------------------------
// parameters: Adam, bf16 copies of the filters, flipped / transposed filters
// ---------------------------------------------------------------------------
__global__ void k_fill(float *__restrict__ p, float v, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_adam(float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// wb[co][t][ci] = bf16(w[co][t][ci]);  wt[ci][t][co] = bf16(w[co][8 - t][ci])
__global__ void k_pack_filters(const float *__restrict__ w, uint16_t *__restrict__ wb, uint16_t *__restrict__ wt,
                               int Cin, int Cout)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)Cout * 9 * Cin) return;
    const int ci = (int)(i % Cin);
    const int t = (int)((i / Cin) % 9);
    const int co = (int)(i / ((size_t)9 * Cin));
    const uint16_t b = f2bf(w[i]);
    wb[i] = b;
    wt[((size_t)ci * 9 + (8 - t)) * Cout + co] = b;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Layer {
    int cin, cout, k;         // k = 3 or 1
    size_t w_off, b_off;      // into the parameter vector
};

struct Feat {                 // one bf16 activation tensor [pixels][ch]
    uint16_t *p = nullptr;
    int ch = 0;
};

struct UNet {
    int c = 32;
    std::vector<Layer> L;     // 15 layers (see the architecture above)
    size_t nparam = 0;
    float *th = nullptr, *gr = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wb = nullptr, *wt = nullptr;
    long step = 0;
    // workspace for (B, H, W)
    int B = 0, H = 0, W = 0;
    bool grads = false;
    std::vector<void *> bufs;
    float *s = nullptr, *in_n = nullptr, *on = nullptr;
    float *lw = nullptr;   // per-item loss weights (relative loss)
    int rel_loss = 0;      // 1: every item's squared error divided by its own normalised target power
    double *loss = nullptr;
    Feat a1, cat1, d1a, d1b, p1, a2, cat2, d2a, d2b, p2, a3, cat3, d3a, d3b, p3, a4, a4b;
    Feat g_a1, g_cat1, g_d1a, g_d1b, g_p1, g_a2, g_cat2, g_d2a, g_d2b, g_p2, g_a3, g_cat3, g_d3a, g_d3b, g_p3, g_a4,
        g_a4b;
};

enum { E1A, E1B, E2A, E2B, E3A, E3B, BA, BB, D3A, D3B, D2A, D2B, D1A, D1B, HEAD, NLAYER };

static void free_ws(UNet *u)
{
    for (void *q : u->bufs) (void)hipFree(q);
    u->bufs.clear();
    u->B = u->H = u->W = 0;
    u->grads = false;
}

static int walloc(UNet *u, void **p, size_t bytes)
{
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(MBREC_ERR_HIP, "mbrec:oom", "unet: device allocation of %zu bytes failed", bytes);
    }
    u->bufs.push_back(*p);
    return MBREC_OK;
}

static int feat(UNet *u, Feat &f, size_t px, int ch)
{
    f.ch = ch;
    return walloc(u, (void **)&f.p, px * ch * sizeof(uint16_t));
}

static int ensure_ws(UNet *u, int B, int H, int W, bool grads)
{
    if (u->B == B && u->H == H && u->W == W && (u->grads || !grads)) return MBREC_OK;
    MBREC_HIP(hipDeviceSynchronize());
    free_ws(u);
    const int c = u->c;
    const size_t P1 = (size_t)B * H * W, P2 = P1 / 4, P3 = P1 / 16, P4 = P1 / 64;
    int rc;
    if ((rc = walloc(u, (void **)&u->s, sizeof(float) * B)) || (rc = walloc(u, (void **)&u->lw, sizeof(float) * B)) ||
        (rc = walloc(u, (void **)&u->in_n, sizeof(float) * P1)) ||
        (rc = walloc(u, (void **)&u->on, sizeof(float) * P1)) || (rc = walloc(u, (void **)&u->loss, sizeof(double))) ||
        (rc = feat(u, u->a1, P1, c)) || (rc = feat(u, u->cat1, P1, 3 * c)) || (rc = feat(u, u->d1a, P1, c)) ||
        (rc = feat(u, u->d1b, P1, c)) || (rc = feat(u, u->p1, P2, c)) || (rc = feat(u, u->a2, P2, 2 * c)) ||
        (rc = feat(u, u->cat2, P2, 6 * c)) || (rc = feat(u, u->d2a, P2, 2 * c)) || (rc = feat(u, u->d2b, P2, 2 * c)) ||
        (rc = feat(u, u->p2, P3, 2 * c)) || (rc = feat(u, u->a3, P3, 4 * c)) || (rc = feat(u, u->cat3, P3, 12 * c)) ||
        (rc = feat(u, u->d3a, P3, 4 * c)) || (rc = feat(u, u->d3b, P3, 4 * c)) || (rc = feat(u, u->p3, P4, 4 * c)) ||
        (rc = feat(u, u->a4, P4, 8 * c)) || (rc = feat(u, u->a4b, P4, 8 * c))) {
        free_ws(u);
        return rc;
    }
    if (grads) {
        if ((rc = feat(u, u->g_a1, P1, c)) || (rc = feat(u, u->g_cat1, P1, 3 * c)) || (rc = feat(u, u->g_d1a, P1, c)) ||
            (rc = feat(u, u->g_d1b, P1, c)) || (rc = feat(u, u->g_p1, P2, c)) || (rc = feat(u, u->g_a2, P2, 2 * c)) ||
            (rc = feat(u, u->g_cat2, P2, 6 * c)) || (rc = feat(u, u->g_d2a, P2, 2 * c)) ||
            (rc = feat(u, u->g_d2b, P2, 2 * c)) || (rc = feat(u, u->g_p2, P3, 2 * c)) ||
            (rc = feat(u, u->g_a3, P3, 4 * c)) || (rc = feat(u, u->g_cat3, P3, 12 * c)) ||
            (rc = feat(u, u->g_d3a, P3, 4 * c)) || (rc = feat(u, u->g_d3b, P3, 4 * c)) ||
            (rc = feat(u, u->g_p3, P4, 4 * c)) || (rc = feat(u, u->g_a4, P4, 8 * c)) ||
            (rc = feat(u, u->g_a4b, P4, 8 * c))) {
            free_ws(u);
            return rc;
        }
    }
    u->B = B;
    u->H = H;
    u->W = W;
    u->grads = grads;
    return MBREC_OK;
}

static inline unsigned nb(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// y = conv3x3(x) (+bias, ReLU) or the input gradient form (mode 1, masked)
static void conv3(const Feat &x, int xo, int cin, const uint16_t *w, const float *bias, const Feat &y, int yo, int cout,
                  int N, int H, int W, int mode, const Feat *mask, int mo, hipStream_t s)
{
    const size_t NP = (size_t)N * H * W;
    const uint16_t *mp = mask ? mask->p : nullptr;
    const int ms = mask ? mask->ch : 0;
    if (cout % 64 == 0) {
        dim3 grid(nb(NP, 64), cout / 64);
        hipLaunchKernelGGL(k_conv3_mfma<4>, grid, dim3(256), 0, s, x.p, x.ch, xo, w, bias, y.p, y.ch, yo, mp, ms, mo, N,
                           H, W, cin, cout, mode);
    } else {
        dim3 grid(nb(NP, 64), cout / 32);
        hipLaunchKernelGGL(k_conv3_mfma<2>, grid, dim3(256), 0, s, x.p, x.ch, xo, w, bias, y.p, y.ch, yo, mp, ms, mo, N,
                           H, W, cin, cout, mode);
    }
}

static void wgrad3(const Feat &g, int go, const Feat &x, int xo, int cin, int cout, float *dw, float *db, int N, int H,
                   int W, hipStream_t s)
{
    const int rows = N * H;
    dim3 grid((cout / 32) * (cin / 32), (rows + 4 * kWgRowsPerWave - 1) / (4 * kWgRowsPerWave), 9);
    hipLaunchKernelGGL(k_conv3_wgrad, grid, dim3(256), 0, s, g.p, g.ch, go, x.p, x.ch, xo, dw, N, H, W, cin, cout);
    const size_t NP = (size_t)N * H * W;
    const int chunk = 4096;
    hipLaunchKernelGGL(k_bias_grad, dim3(nb(NP, chunk)), dim3(256), 0, s, g.p, g.ch, go, db, (int)NP, cout, chunk);
}

static void pack_all(UNet *u, hipStream_t s)
{
    for (int l = E1B; l <= D1B; ++l) {
        const Layer &ly = u->L[l];
        const size_t n = (size_t)ly.cout * 9 * ly.cin;
        hipLaunchKernelGGL(k_pack_filters, dim3(nb(n)), dim3(256), 0, s, u->th + ly.w_off, u->wb + ly.w_off,
                           u->wt + ly.w_off, ly.cin, ly.cout);
    }
}

static const uint16_t *WB(UNet *u, int l) { return u->wb + u->L[l].w_off; }
static const uint16_t *WT(UNet *u, int l) { return u->wt + u->L[l].w_off; }
static const float *BIAS(UNet *u, int l) { return u->th + u->L[l].b_off; }

static int forward(UNet *u, const float *in, float *out, int B, int H, int W, hipStream_t s)
{
    const int c = u->c;
    const int HW = H * W;
    const size_t P1 = (size_t)B * HW;
    hipLaunchKernelGGL(k_image_scale, dim3(B), dim3(1024), 0, s, in, u->s, HW);
    hipLaunchKernelGGL(k_normalize, dim3(nb(P1)), dim3(256), 0, s, in, u->s, u->in_n, HW, P1);
    hipLaunchKernelGGL(k_conv_first, dim3(nb(P1)), dim3(256), 0, s, u->in_n, u->th + u->L[E1A].w_off, BIAS(u, E1A),
                       u->a1.p, B, H, W, c);
    const int H2 = H / 2, W2 = W / 2, H3 = H / 4, W3 = W / 4, H4 = H / 8, W4 = W / 8;
    // encoder
    conv3(u->a1, 0, c, WB(u, E1B), BIAS(u, E1B), u->cat1, 2 * c, c, B, H, W, 0, nullptr, 0, s);
    hipLaunchKernelGGL(k_pool, dim3(nb((size_t)B * H2 * W2 * c / 8)), dim3(256), 0, s, u->cat1.p, u->cat1.ch, 2 * c,
                       u->p1.p, B, H, W, c);
    conv3(u->p1, 0, c, WB(u, E2A), BIAS(u, E2A), u->a2, 0, 2 * c, B, H2, W2, 0, nullptr, 0, s);
    conv3(u->a2, 0, 2 * c, WB(u, E2B), BIAS(u, E2B), u->cat2, 4 * c, 2 * c, B, H2, W2, 0, nullptr, 0, s);
    hipLaunchKernelGGL(k_pool, dim3(nb((size_t)B * H3 * W3 * 2 * c / 8)), dim3(256), 0, s, u->cat2.p, u->cat2.ch,
                       4 * c, u->p2.p, B, H2, W2, 2 * c);
    conv3(u->p2, 0, 2 * c, WB(u, E3A), BIAS(u, E3A), u->a3, 0, 4 * c, B, H3, W3, 0, nullptr, 0, s);
    conv3(u->a3, 0, 4 * c, WB(u, E3B), BIAS(u, E3B), u->cat3, 8 * c, 4 * c, B, H3, W3, 0, nullptr, 0, s);
    hipLaunchKernelGGL(k_pool, dim3(nb((size_t)B * H4 * W4 * 4 * c / 8)), dim3(256), 0, s, u->cat3.p, u->cat3.ch,
                       8 * c, u->p3.p, B, H3, W3, 4 * c);
    conv3(u->p3, 0, 4 * c, WB(u, BA), BIAS(u, BA), u->a4, 0, 8 * c, B, H4, W4, 0, nullptr, 0, s);
    conv3(u->a4, 0, 8 * c, WB(u, BB), BIAS(u, BB), u->a4b, 0, 8 * c, B, H4, W4, 0, nullptr, 0, s);
    // decoder
    hipLaunchKernelGGL(k_upsample, dim3(nb((size_t)B * H3 * W3 * 8 * c / 8)), dim3(256), 0, s, u->a4b.p, u->cat3.p,
                       u->cat3.ch, B, H3, W3, 8 * c);
    conv3(u->cat3, 0, 12 * c, WB(u, D3A), BIAS(u, D3A), u->d3a, 0, 4 * c, B, H3, W3, 0, nullptr, 0, s);
    conv3(u->d3a, 0, 4 * c, WB(u, D3B), BIAS(u, D3B), u->d3b, 0, 4 * c, B, H3, W3, 0, nullptr, 0, s);
    hipLaunchKernelGGL(k_upsample, dim3(nb((size_t)B * H2 * W2 * 4 * c / 8)), dim3(256), 0, s, u->d3b.p, u->cat2.p,
                       u->cat2.ch, B, H2, W2, 4 * c);
    conv3(u->cat2, 0, 6 * c, WB(u, D2A), BIAS(u, D2A), u->d2a, 0, 2 * c, B, H2, W2, 0, nullptr, 0, s);
    conv3(u->d2a, 0, 2 * c, WB(u, D2B), BIAS(u, D2B), u->d2b, 0, 2 * c, B, H2, W2, 0, nullptr, 0, s);
    hipLaunchKernelGGL(k_upsample, dim3(nb((size_t)B * H * W * 2 * c / 8)), dim3(256), 0, s, u->d2b.p, u->cat1.p,
                       u->cat1.ch, B, H, W, 2 * c);
    conv3(u->cat1, 0, 3 * c, WB(u, D1A), BIAS(u, D1A), u->d1a, 0, c, B, H, W, 0, nullptr, 0, s);
    conv3(u->d1a, 0, c, WB(u, D1B), BIAS(u, D1B), u->d1b, 0, c, B, H, W, 0, nullptr, 0, s);
    hipLaunchKernelGGL(k_head, dim3(nb(P1)), dim3(256), 0, s, u->d1b.p, u->th + u->L[HEAD].w_off, BIAS(u, HEAD), u->s,
                       u->on, out, HW, (int)P1, c);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

static float *GW(UNet *u, int l) { return u->gr + u->L[l].w_off; }
static float *GB(UNet *u, int l) { return u->gr + u->L[l].b_off; }

// gradients of the loss with respect to all parameters (u->gr, zeroed here)
static int backward(UNet *u, const float *target, int B, int H, int W, hipStream_t s)
{
    const int c = u->c;
    const int HW = H * W;
    const size_t P1 = (size_t)B * HW;
    const int H2 = H / 2, W2 = W / 2, H3 = H / 4, W3 = W / 4, H4 = H / 8, W4 = W / 8;
    hipLaunchKernelGGL(k_fill, dim3(nb(u->nparam)), dim3(256), 0, s, u->gr, 0.0f, u->nparam);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, (float *)u->loss, 0.0f, (size_t)2);
    if (u->rel_loss) {
        hipLaunchKernelGGL(k_target_power, dim3(B), dim3(1024), 0, s, target, (const float *)u->s, u->lw, HW);
        hipLaunchKernelGGL(k_loss_weights, dim3(1), dim3(64), 0, s, u->lw, B);
    }
    hipLaunchKernelGGL(k_head_loss_bwd, dim3(nb(P1)), dim3(256), 0, s, u->d1b.p, u->th + u->L[HEAD].w_off, u->on,
                       target, u->s, u->rel_loss ? (const float *)u->lw : nullptr, u->g_d1b.p, GW(u, HEAD),
                       GB(u, HEAD), u->loss, HW, (int)P1, c);
    // level 1 decoder
    wgrad3(u->g_d1b, 0, u->d1a, 0, c, c, GW(u, D1B), GB(u, D1B), B, H, W, s);
    conv3(u->g_d1b, 0, c, WT(u, D1B), nullptr, u->g_d1a, 0, c, B, H, W, 1, &u->d1a, 0, s);
    wgrad3(u->g_d1a, 0, u->cat1, 0, 3 * c, c, GW(u, D1A), GB(u, D1A), B, H, W, s);
    conv3(u->g_d1a, 0, c, WT(u, D1A), nullptr, u->g_cat1, 0, 3 * c, B, H, W, 1, nullptr, 0, s);
    hipLaunchKernelGGL(k_upsample_bwd, dim3(nb((size_t)B * H2 * W2 * 2 * c / 8)), dim3(256), 0, s, u->g_cat1.p,
                       u->g_cat1.ch, u->d2b.p, u->g_d2b.p, B, H2, W2, 2 * c);
    // level 2 decoder
    wgrad3(u->g_d2b, 0, u->d2a, 0, 2 * c, 2 * c, GW(u, D2B), GB(u, D2B), B, H2, W2, s);
    conv3(u->g_d2b, 0, 2 * c, WT(u, D2B), nullptr, u->g_d2a, 0, 2 * c, B, H2, W2, 1, &u->d2a, 0, s);
    wgrad3(u->g_d2a, 0, u->cat2, 0, 6 * c, 2 * c, GW(u, D2A), GB(u, D2A), B, H2, W2, s);
    conv3(u->g_d2a, 0, 2 * c, WT(u, D2A), nullptr, u->g_cat2, 0, 6 * c, B, H2, W2, 1, nullptr, 0, s);
    hipLaunchKernelGGL(k_upsample_bwd, dim3(nb((size_t)B * H3 * W3 * 4 * c / 8)), dim3(256), 0, s, u->g_cat2.p,
                       u->g_cat2.ch, u->d3b.p, u->g_d3b.p, B, H3, W3, 4 * c);
    // level 3 decoder
    wgrad3(u->g_d3b, 0, u->d3a, 0, 4 * c, 4 * c, GW(u, D3B), GB(u, D3B), B, H3, W3, s);
    conv3(u->g_d3b, 0, 4 * c, WT(u, D3B), nullptr, u->g_d3a, 0, 4 * c, B, H3, W3, 1, &u->d3a, 0, s);
    wgrad3(u->g_d3a, 0, u->cat3, 0, 12 * c, 4 * c, GW(u, D3A), GB(u, D3A), B, H3, W3, s);
    conv3(u->g_d3a, 0, 4 * c, WT(u, D3A), nullptr, u->g_cat3, 0, 12 * c, B, H3, W3, 1, nullptr, 0, s);
    hipLaunchKernelGGL(k_upsample_bwd, dim3(nb((size_t)B * H4 * W4 * 8 * c / 8)), dim3(256), 0, s, u->g_cat3.p,
                       u->g_cat3.ch, u->a4b.p, u->g_a4b.p, B, H4, W4, 8 * c);
    // bottleneck
    wgrad3(u->g_a4b, 0, u->a4, 0, 8 * c, 8 * c, GW(u, BB), GB(u, BB), B, H4, W4, s);
    conv3(u->g_a4b, 0, 8 * c, WT(u, BB), nullptr, u->g_a4, 0, 8 * c, B, H4, W4, 1, &u->a4, 0, s);
    wgrad3(u->g_a4, 0, u->p3, 0, 4 * c, 8 * c, GW(u, BA), GB(u, BA), B, H4, W4, s);
    conv3(u->g_a4, 0, 8 * c, WT(u, BA), nullptr, u->g_p3, 0, 4 * c, B, H4, W4, 1, nullptr, 0, s);
    // level 3 encoder (skip = cat3[8c:12c))
    hipLaunchKernelGGL(k_skip_grad, dim3(nb((size_t)B * H3 * W3 * 4 * c / 8)), dim3(256), 0, s, u->g_cat3.p, u->cat3.p,
                       12 * c, 8 * c, u->g_p3.p, B, H3, W3, 4 * c);
    wgrad3(u->g_cat3, 8 * c, u->a3, 0, 4 * c, 4 * c, GW(u, E3B), GB(u, E3B), B, H3, W3, s);
    conv3(u->g_cat3, 8 * c, 4 * c, WT(u, E3B), nullptr, u->g_a3, 0, 4 * c, B, H3, W3, 1, &u->a3, 0, s);
    wgrad3(u->g_a3, 0, u->p2, 0, 2 * c, 4 * c, GW(u, E3A), GB(u, E3A), B, H3, W3, s);
    conv3(u->g_a3, 0, 4 * c, WT(u, E3A), nullptr, u->g_p2, 0, 2 * c, B, H3, W3, 1, nullptr, 0, s);
    // level 2 encoder (skip = cat2[4c:6c))
    hipLaunchKernelGGL(k_skip_grad, dim3(nb((size_t)B * H2 * W2 * 2 * c / 8)), dim3(256), 0, s, u->g_cat2.p, u->cat2.p,
                       6 * c, 4 * c, u->g_p2.p, B, H2, W2, 2 * c);
    wgrad3(u->g_cat2, 4 * c, u->a2, 0, 2 * c, 2 * c, GW(u, E2B), GB(u, E2B), B, H2, W2, s);
    conv3(u->g_cat2, 4 * c, 2 * c, WT(u, E2B), nullptr, u->g_a2, 0, 2 * c, B, H2, W2, 1, &u->a2, 0, s);
    wgrad3(u->g_a2, 0, u->p1, 0, c, 2 * c, GW(u, E2A), GB(u, E2A), B, H2, W2, s);
    conv3(u->g_a2, 0, 2 * c, WT(u, E2A), nullptr, u->g_p1, 0, c, B, H2, W2, 1, nullptr, 0, s);
    // level 1 encoder (skip = cat1[2c:3c))
    hipLaunchKernelGGL(k_skip_grad, dim3(nb((size_t)B * H * W * c / 8)), dim3(256), 0, s, u->g_cat1.p, u->cat1.p, 3 * c,
                       2 * c, u->g_p1.p, B, H, W, c);
    wgrad3(u->g_cat1, 2 * c, u->a1, 0, c, c, GW(u, E1B), GB(u, E1B), B, H, W, s);
    conv3(u->g_cat1, 2 * c, c, WT(u, E1B), nullptr, u->g_a1, 0, c, B, H, W, 1, &u->a1, 0, s);
    const int chunk = 4096;
    hipLaunchKernelGGL(k_conv_first_wgrad, dim3(nb(P1, chunk)), dim3(256), 0, s, u->g_a1.p, u->in_n, GW(u, E1A),
                       GB(u, E1A), B, H, W, c, chunk);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

static int check_shape(UNet *u, int B, int H, int W)
{
    if (B <= 0 || H <= 0 || W <= 0 || H % 8 || W % 8)
        return set_error(MBREC_ERR_ARG, "mbrec:unetShape", "unet: B, H, W must be positive and H, W multiples of 8 "
                                                           "(got B=%d H=%d W=%d)", B, H, W);
    if ((size_t)B * H * W * 12 * u->c > (size_t)INT32_MAX * 4)
        return set_error(MBREC_ERR_ARG, "mbrec:unetShape", "unet: batch too large");
    return MBREC_OK;
}

}  // namespace mbrec

using namespace mbrec;

struct mbrec_unet_s {
    UNet u;
};

extern "C" {

int mbrec_unet_create(mbrec_unet *out, int base_channels, unsigned long long seed)
{
    if (!out) return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_create: null output");
    *out = nullptr;
    if (base_channels <= 0 || base_channels % 32 || base_channels > 64)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_create: base_channels must be 32 or 64 (got %d)",
                         base_channels);
    mbrec_unet h = new mbrec_unet_s;
    UNet *u = &h->u;
    const int c = base_channels;
    u->c = c;
    const int io[NLAYER][3] = {{1, c, 3},          {c, c, 3},          {c, 2 * c, 3},      {2 * c, 2 * c, 3},
                               {2 * c, 4 * c, 3},  {4 * c, 4 * c, 3},  {4 * c, 8 * c, 3},  {8 * c, 8 * c, 3},
                               {12 * c, 4 * c, 3}, {4 * c, 4 * c, 3},  {6 * c, 2 * c, 3},  {2 * c, 2 * c, 3},
                               {3 * c, c, 3},      {c, c, 3},          {c, 1, 1}};
    size_t off = 0;
    for (int l = 0; l < NLAYER; ++l) {
        Layer ly{io[l][0], io[l][1], io[l][2], 0, 0};
        ly.w_off = off;
        off += (size_t)ly.cout * ly.k * ly.k * ly.cin;
        ly.b_off = off;
        off += (size_t)ly.cout;
        off = (off + 7) / 8 * 8;   // keep every filter 16-byte aligned in the bf16 copies
        u->L.push_back(ly);
    }
    u->nparam = off;
    // He-normal filters, zero biases (seeded splitmix64 + Box-Muller)
    std::vector<float> th(off, 0.0f);
    uint64_t st = seed ^ 0x9e3779b97f4a7c15ull;
    auto next = [&]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < NLAYER; ++l) {
        const Layer &ly = u->L[l];
        const size_t n = (size_t)ly.cout * ly.k * ly.k * ly.cin;
        const double sd = std::sqrt(2.0 / (ly.k * ly.k * ly.cin)) * (l == HEAD ? 0.5 : 1.0);
        for (size_t i = 0; i < n; ++i) {
            const double u1 = next(), u2 = next();
            th[ly.w_off + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
    auto fail = [&](int rc) {
        mbrec_unet_destroy(h);
        return rc;
    };
    const size_t bytes = sizeof(float) * off;
    if (hipMalloc((void **)&u->th, bytes) != hipSuccess || hipMalloc((void **)&u->gr, bytes) != hipSuccess ||
        hipMalloc((void **)&u->m1, bytes) != hipSuccess || hipMalloc((void **)&u->m2, bytes) != hipSuccess ||
        hipMalloc((void **)&u->wb, sizeof(uint16_t) * off) != hipSuccess ||
        hipMalloc((void **)&u->wt, sizeof(uint16_t) * off) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(MBREC_ERR_HIP, "mbrec:oom", "unet_create: device allocation failed"));
    }
    if (hipMemcpy(u->th, th.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(u->m1, 0, bytes) != hipSuccess || hipMemset(u->m2, 0, bytes) != hipSuccess ||
        hipMemset(u->gr, 0, bytes) != hipSuccess || hipMemset(u->wb, 0, sizeof(uint16_t) * off) != hipSuccess ||
        hipMemset(u->wt, 0, sizeof(uint16_t) * off) != hipSuccess)
        return fail(set_error(MBREC_ERR_HIP, "mbrec:hip", "unet_create: upload failed"));
    pack_all(u, 0);
    if (hipDeviceSynchronize() != hipSuccess) return fail(set_error(MBREC_ERR_HIP, "mbrec:hip", "unet_create: pack"));
    *out = h;
    return MBREC_OK;
}

int mbrec_unet_destroy(mbrec_unet h)
{
    if (!h) return MBREC_OK;
    (void)hipDeviceSynchronize();
    UNet *u = &h->u;
    free_ws(u);
    void *p[] = {u->th, u->gr, u->m1, u->m2, u->wb, u->wt};
    for (void *q : p)
        if (q) (void)hipFree(q);
    delete h;
    return MBREC_OK;
}

size_t mbrec_unet_num_params(mbrec_unet h) { return h ? h->u.nparam : 0; }

int mbrec_unet_layout(mbrec_unet h, int layer, int *cin, int *cout, int *ksize, size_t *w_off, size_t *b_off)
{
    if (!h || layer < 0 || layer >= NLAYER) return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_layout: bad layer");
    const Layer &ly = h->u.L[layer];
    *cin = ly.cin;
    *cout = ly.cout;
    *ksize = ly.k;
    *w_off = ly.w_off;
    *b_off = ly.b_off;
    return MBREC_OK;
}

int mbrec_unet_get_params(mbrec_unet h, float *host, size_t n)
{
    if (!h || !host || n != h->u.nparam)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_get_params: expected %zu values", h ? h->u.nparam : 0);
    MBREC_HIP(hipDeviceSynchronize());
    MBREC_HIP(hipMemcpy(host, h->u.th, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MBREC_OK;
}

int mbrec_unet_set_params(mbrec_unet h, const float *host, size_t n)
{
    if (!h || !host || n != h->u.nparam)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_set_params: expected %zu values", h ? h->u.nparam : 0);
    UNet *u = &h->u;
    MBREC_HIP(hipDeviceSynchronize());
    MBREC_HIP(hipMemcpy(u->th, host, sizeof(float) * n, hipMemcpyHostToDevice));
    MBREC_HIP(hipMemset(u->m1, 0, sizeof(float) * n));
    MBREC_HIP(hipMemset(u->m2, 0, sizeof(float) * n));
    u->step = 0;
    pack_all(u, 0);
    MBREC_HIP(hipDeviceSynchronize());
    return MBREC_OK;
}

int mbrec_unet_forward(mbrec_unet h, const float *in, float *out, int B, int H, int W, mbrec_stream stream)
{
    if (!h || !in || !out) return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_forward: null argument");
    UNet *u = &h->u;
    int rc;
    if ((rc = check_shape(u, B, H, W))) return rc;
    if ((rc = ensure_ws(u, B, H, W, false))) return rc;
    return forward(u, in, out, B, H, W, (hipStream_t)stream);
}

int mbrec_unet_gradients(mbrec_unet h, const float *in, const float *target, int B, int H, int W, float *grad_host,
                         double *loss_host, mbrec_stream stream)
{
    if (!h || !in || !target) return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_gradients: null argument");
    UNet *u = &h->u;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = check_shape(u, B, H, W))) return rc;
    if ((rc = ensure_ws(u, B, H, W, true))) return rc;
    if ((rc = forward(u, in, nullptr, B, H, W, s))) return rc;
    if ((rc = backward(u, target, B, H, W, s))) return rc;
    MBREC_HIP(hipStreamSynchronize(s));
    if (grad_host) MBREC_HIP(hipMemcpy(grad_host, u->gr, sizeof(float) * u->nparam, hipMemcpyDeviceToHost));
    if (loss_host) {
        double l = 0.0;
        MBREC_HIP(hipMemcpy(&l, u->loss, sizeof(double), hipMemcpyDeviceToHost));
        *loss_host = l / ((double)B * H * W);
    }
    return MBREC_OK;
}

int mbrec_unet_set_relative_loss(mbrec_unet h, int on)
{
    if (!h) return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_set_relative_loss: null argument");
    h->u.rel_loss = on ? 1 : 0;
    return MBREC_OK;
}

int mbrec_unet_train_step(mbrec_unet h, const float *in, const float *target, int B, int H, int W, float lr,
                          float beta1, float beta2, float eps, double *loss_host, mbrec_stream stream)
{
    if (!h || !in || !target) return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_train_step: null argument");
    if (!(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "unet_train_step: invalid Adam settings");
    UNet *u = &h->u;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if ((rc = check_shape(u, B, H, W))) return rc;
    if ((rc = ensure_ws(u, B, H, W, true))) return rc;
    if ((rc = forward(u, in, nullptr, B, H, W, s))) return rc;
    if ((rc = backward(u, target, B, H, W, s))) return rc;
    ++u->step;
    const float c1 = 1.0f - std::pow(beta1, (float)u->step), c2 = 1.0f - std::pow(beta2, (float)u->step);
    hipLaunchKernelGGL(k_adam, dim3(nb(u->nparam)), dim3(256), 0, s, u->th, u->gr, u->m1, u->m2, u->nparam, lr, beta1,
                       beta2, eps, c1, c2);
    pack_all(u, s);
    MBREC_HIP(hipGetLastError());
    if (loss_host) {
        double l = 0.0;
        MBREC_HIP(hipMemcpyAsync(&l, u->loss, sizeof(double), hipMemcpyDeviceToHost, s));
        MBREC_HIP(hipStreamSynchronize(s));
        *loss_host = l / ((double)B * H * W);
    }
    return MBREC_OK;
}

// One 3x3 convolution layer (bf16 NHWC input [N][H][W][Cin], bf16 filters
// [Cout][3][3][Cin], FP32 bias, + ReLU) through the matrix-core kernel.
int mbrec_conv3x3_bf16(const uint16_t *x, const uint16_t *w, const float *bias, uint16_t *y, int N, int H, int W,
                       int Cin, int Cout, int relu, mbrec_stream stream)
{
    if (!x || !w || !y || N <= 0 || H <= 0 || W <= 0 || Cin % 32 || Cout % 32 || Cin <= 0 || Cout <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "conv3x3_bf16: Cin and Cout must be positive multiples of 32");
    Feat fx{const_cast<uint16_t *>(x), Cin}, fy{y, Cout};
    if (relu) {
        conv3(fx, 0, Cin, w, bias, fy, 0, Cout, N, H, W, 0, nullptr, 0, (hipStream_t)stream);
    } else {
        // mode 1 without a mask: plain convolution (bias ignored)
        if (bias) return set_error(MBREC_ERR_ARG, "mbrec:args", "conv3x3_bf16: bias requires relu=1");
        conv3(fx, 0, Cin, w, nullptr, fy, 0, Cout, N, H, W, 1, nullptr, 0, (hipStream_t)stream);
    }
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

}  // extern "C"
