// mbrec_solver.hip -- device-resident non-negative least-squares loops.
//
// nnls_L2_less (reconstruction/nnls_L2_less.m:28-72) around MATLAB-semantics
// LSQR (Paige & Saunders 1982, in the d-vector form of MATLAB's lsqr: initial
// residual b - A x0, one transpose product at the end of every iteration,
// residual estimate normr = |s| normr, stopping tests normr <= tol ||b|| and
// ||A'r|| / (||A|| ||r||) <= tol, stagnation counter).  The augmented operator
// is A_aug = [sqrt(l) A_reg; sqrt(b) I] (add_nn_constraint_term_to_model_matrix,
// nnls_L2_less.m:79-101) with A_reg = [A; sqrt(lambdaL2) R1; sqrt(lambdaL2_2) R2]
// (combine_forward_and_transpose_with_reg.m:6-14, _with_two_regs.m:8-14): each R is the
// identity (fused into the vector kernels), MATLAB del2 applied per image with del2 also
// as its "transpose" (laplacian_per_wavelength.m, reconstruction_overview_script.m:50-51),
// or a caller-supplied operator pair.
//
// Vectors are FP32 on the device; every inner product, norm and recurrence
// scalar is FP64 and lives on the device, one scalar set per group (all B
// sinograms in one group = the reference's stacked multi-wavelength call;
// one group per sinogram = independent batched solves).  No host round trip
// happens inside a solve, so the launch sequence is graph-capturable.
#include "mbrec_reduce.h"

#include <cmath>
#include <vector>

namespace mbrec {

struct Scal {
    double alpha, beta, c, s, phibar, phi, thet, rho;
    double normr, norma, normar, tolb, n2b;
    double inv_u, inv_v;           // pending normalisations of u and v
    int stag, flag, active, iter;
};

struct Consts {
    float sl, sb, sr;   // sqrt(lambda), sqrt(beta), sqrt(lambda)*sqrt(lambdaL2)
    int reg;            // lambdaL2 > 0
    double tol;
    int coupled;
    const float *srv;   // per-sinogram sqrt(lambdaL2) (independent solves), or null: sr for all
    float sr2;          // sqrt(lambda)*sqrt(lambdaL2_2)
    int reg2;           // lambdaL2_2 > 0
};

// One L2 term's operator: kind 0 identity, 1 MATLAB del2 per image (its own transpose,
// as the reference uses it), 2 the caller's pair (fn enqueues out = R in / R' in on s).
struct RegOp {
    int kind = 0;
    mbrec_reg_apply fn = nullptr;
    void *ctx = nullptr;
};

__device__ __forceinline__ float sr_of(const Consts &k, int b) { return k.srv ? k.srv[b] : k.sr; }

// ---- outer-iteration setup: u = b - A_aug x0 with x0 = w --------------------
// sinogram block: u_s = sl*(y - A w); ||u||^2 -> Q_A, ||b||^2 -> Q_B
__global__ __launch_bounds__(256) void k_init_sino(const float *__restrict__ y, const float *__restrict__ aw,
                                                   float *__restrict__ us, size_t n, int B, Consts k,
                                                   double *part)
{
    const int b = blockIdx.y;
    const size_t off = (size_t)b * n;
    double uu = 0.0, bb = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float yy = y[off + i];
        const float u = k.sl * (yy - aw[off + i]);
        us[off + i] = u;
        uu += (double)u * u;
        const double bv = (double)k.sl * yy;
        bb += bv * bv;
    }
    block_sum_store(uu, part, Q_A, b, B);
    block_sum_store(bb, part, Q_B, b, B);
}

// pixel blocks: u_r = -sr*R1 w, u_r2 = -sr2*R2 w, u_c = sb*vm, x = w, d = 0;
// ||u||^2 -> Q_C, ||b||^2 -> Q_D.  r1w / r2w: R1 w and R2 w (w itself for an identity term)
__global__ __launch_bounds__(256) void k_init_pix(const float *__restrict__ w, const float *__restrict__ vm,
                                                  const float *__restrict__ r1w, const float *__restrict__ r2w,
                                                  float *__restrict__ ur, float *__restrict__ ur2,
                                                  float *__restrict__ uc, float *__restrict__ x,
                                                  float *__restrict__ dv, size_t n, int B, Consts k, double *part)
{
    const int b = blockIdx.y;
    const size_t off = (size_t)b * n;
    double uu = 0.0, bb = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float wi = w[off + i], vi = vm[off + i];
        if (k.reg) {
            const float r = -sr_of(k, b) * r1w[off + i];
            ur[off + i] = r;
            uu += (double)r * r;
        }
        if (k.reg2) {
            const float r = -k.sr2 * r2w[off + i];
            ur2[off + i] = r;
            uu += (double)r * r;
        }
        const float bc = k.sb * (wi + vi);
        const float c = bc - k.sb * wi;
        uc[off + i] = c;
        uu += (double)c * c;
        bb += (double)bc * bc;
        x[off + i] = wi;
        dv[off + i] = 0.0f;
    }
    block_sum_store(uu, part, Q_C, b, B);
    block_sum_store(bb, part, Q_D, b, B);
}

__global__ void k_scal_init(Scal *sc, const double *part, int G, int B, Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        Scal &S = sc[g];
        const double bb = group_sum(part, Q_B, g, B, k.coupled) + group_sum(part, Q_D, g, B, k.coupled);
        const double uu = group_sum(part, Q_A, g, B, k.coupled) + group_sum(part, Q_C, g, B, k.coupled);
        S.n2b = sqrt(bb);
        S.tolb = k.tol * S.n2b;
        S.beta = sqrt(uu);
        S.normr = S.beta;
        S.c = 1.0;
        S.s = 0.0;
        S.phibar = S.beta;
        S.inv_u = S.beta != 0.0 ? 1.0 / S.beta : 1.0;
        S.norma = 0.0;
        S.stag = 0;
        S.iter = 0;
        S.flag = 1;
        S.active = 1;
        if (S.n2b == 0.0) { S.flag = 0; S.active = 0; S.iter = -1; }  // x = 0 (b == 0)
    }
}

// u and v are kept unnormalised: the stored vectors times the pending factors
// inv_u / inv_v of their group are LSQR's unit vectors, and every consumer
// applies the factor as it reads (no separate scaling pass over the vectors).
//
// v = inv_u*(sl*A'u_s + sr*R1'u_r + sr2*R2'u_r2 + sb*u_c) - beta*inv_v*v (first=1: without
// the beta term); ||v||^2 -> Q_E.  atu = A' applied to the stored u_s; rt1 / rt2 = R1'u_r /
// R2'u_r2 (u_r / u_r2 themselves for an identity term).
__global__ __launch_bounds__(256) void k_v_update(const float *__restrict__ atu, const float *__restrict__ ur,
                                                  const float *__restrict__ ur2, const float *__restrict__ uc,
                                                  float *__restrict__ v, size_t n, int B, Consts k, const Scal *sc,
                                                  int first, double *part)
{
    const int b = blockIdx.y;
    const Scal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    double vv = 0.0;
    if (S.active) {
        const float bv = first ? 0.0f : (float)(S.beta * S.inv_v), fu = (float)S.inv_u;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            float t = k.sl * atu[off + i] + k.sb * uc[off + i];
            if (k.reg) t += sr_of(k, b) * ur[off + i];
            if (k.reg2) t += k.sr2 * ur2[off + i];
            t *= fu;
            const float vn = first ? t : t - bv * v[off + i];   // v is not yet defined when first
            v[off + i] = vn;
            vv += (double)vn * vn;
        }
    }
    block_sum_store(vv, part, Q_E, b, B);
}

__global__ void k_scal_v0(Scal *sc, const double *part, int G, int B, Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        Scal &S = sc[g];
        if (!S.active) continue;
        const double vv = group_sum(part, Q_E, g, B, k.coupled);
        S.alpha = sqrt(vv);
        S.inv_v = S.alpha != 0.0 ? 1.0 / S.alpha : 1.0;
        S.normar = S.alpha * S.beta;
        if (S.normar == 0.0) { S.flag = 0; S.active = 0; }
    }
}

// u_s = sl*inv_v*(A v) - alpha*inv_u*u_s (stored vectors); ||.||^2 -> Q_A
__global__ __launch_bounds__(256) void k_u_sino(const float *__restrict__ av, float *__restrict__ us, size_t n,
                                                int B, Consts k, const Scal *sc, double *part)
{
    const int b = blockIdx.y;
    const Scal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    double uu = 0.0;
    if (S.active) {
        const float a = (float)(S.alpha * S.inv_u), fv = (float)(k.sl * S.inv_v);
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float u = fv * av[off + i] - a * us[off + i];
            us[off + i] = u;
            uu += (double)u * u;
        }
    }
    block_sum_store(uu, part, Q_A, b, B);
}

// u_r = sr*inv_v*R1 v - alpha*inv_u*u_r, u_r2 = sr2*inv_v*R2 v - alpha*inv_u*u_r2,
// u_c = sb*inv_v*v - alpha*inv_u*u_c; ||.||^2 -> Q_C.  r1v / r2v: R1 v and R2 v (v itself for
// an identity term)
__global__ __launch_bounds__(256) void k_u_pix(const float *__restrict__ v, const float *__restrict__ r1v,
                                               const float *__restrict__ r2v, float *__restrict__ ur,
                                               float *__restrict__ ur2, float *__restrict__ uc, size_t n, int B,
                                               Consts k, const Scal *sc, double *part)
{
    const int b = blockIdx.y;
    const Scal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    double uu = 0.0;
    if (S.active) {
        const float a = (float)(S.alpha * S.inv_u), fv = (float)S.inv_v;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float vi = fv * v[off + i];
            if (k.reg) {
                const float r = sr_of(k, b) * (fv * r1v[off + i]) - a * ur[off + i];
                ur[off + i] = r;
                uu += (double)r * r;
            }
            if (k.reg2) {
                const float r = k.sr2 * (fv * r2v[off + i]) - a * ur2[off + i];
                ur2[off + i] = r;
                uu += (double)r * r;
            }
            const float c = k.sb * vi - a * uc[off + i];
            uc[off + i] = c;
            uu += (double)c * c;
        }
    }
    block_sum_store(uu, part, Q_C, b, B);
}

__global__ void k_scal_u(Scal *sc, const double *part, int G, int B, Consts k)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        Scal &S = sc[g];
        if (!S.active) continue;
        const double uu = group_sum(part, Q_A, g, B, k.coupled) + group_sum(part, Q_C, g, B, k.coupled);
        S.beta = sqrt(uu);
        S.inv_u = 1.0 / S.beta;
        S.norma = sqrt(S.norma * S.norma + S.alpha * S.alpha + S.beta * S.beta);
        S.thet = -S.s * S.alpha;
        const double rhot = S.c * S.alpha;
        S.rho = sqrt(rhot * rhot + S.beta * S.beta);
        S.c = rhot / S.rho;
        S.s = -S.beta / S.rho;
        S.phi = S.c * S.phibar;
        if (S.phi == 0.0) S.stag = 1;
        S.phibar = S.s * S.phibar;
    }
}

// ||x_old||^2 -> Q_B, d = (inv_v*v - thet*d)/rho, ||d||^2 -> Q_D, x += phi*d
__global__ __launch_bounds__(256) void k_dx(const float *__restrict__ v, float *__restrict__ dv,
                                            float *__restrict__ x, size_t n, int B, Consts k,
                                            const Scal *sc, double *part)
{
    const int b = blockIdx.y;
    const Scal &S = sc[k.coupled ? 0 : b];
    const size_t off = (size_t)b * n;
    double xx = 0.0, dd = 0.0;
    if (S.active) {
        const float thet = (float)S.thet, irho = (float)(1.0 / S.rho), phi = (float)S.phi, fv = (float)S.inv_v;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float xo = x[off + i];
            xx += (double)xo * xo;
            const float dn = (fv * v[off + i] - thet * dv[off + i]) * irho;
            dv[off + i] = dn;
            dd += (double)dn * dn;
            x[off + i] = xo + phi * dn;
        }
    }
    block_sum_store(xx, part, Q_B, b, B);
    block_sum_store(dd, part, Q_D, b, B);
}

__global__ void k_scal_v(Scal *sc, const double *part, int G, int B, Consts k, int ii)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        Scal &S = sc[g];
        if (!S.active) continue;
        const double xx = group_sum(part, Q_B, g, B, k.coupled);
        const double dd = group_sum(part, Q_D, g, B, k.coupled);
        if (fabs(S.phi) * sqrt(dd) < 2.220446049250313e-16 * sqrt(xx)) S.stag++;
        else S.stag = 0;
        const double vv = group_sum(part, Q_E, g, B, k.coupled);
        S.alpha = sqrt(vv);
        S.inv_v = 1.0 / S.alpha;
        S.normr = fabs(S.s) * S.normr;
        S.normar = S.alpha * fabs(S.s * S.phi);
        S.iter = ii;
        if (S.normr <= S.tolb) { S.flag = 0; S.active = 0; }
        else if (S.normar / (S.norma * S.normr) <= k.tol) { S.flag = 0; S.active = 0; }
        else if (S.stag >= 3) { S.flag = 3; S.active = 0; }
    }
}

// w = max(x - vm, 0); vm += w - x  (nnls_L2_less.m:60,64).  Groups whose
// right-hand side vanished (iter == -1) take x = 0 (MATLAB lsqr with b == 0).
__global__ __launch_bounds__(256) void k_nnls_update(const float *__restrict__ x, float *__restrict__ w,
                                                     float *__restrict__ vm, size_t n, int B, Consts k,
                                                     const Scal *sc)
{
    const int b = blockIdx.y;
    const Scal &S = sc[k.coupled ? 0 : b];
    const bool zero = S.iter == -1;
    const size_t off = (size_t)b * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float xi = zero ? 0.0f : x[off + i];
        const float vi = vm[off + i];
        const float wn = fmaxf(xi - vi, 0.0f);
        w[off + i] = wn;
        vm[off + i] = vi + wn - xi;
    }
}

__global__ __launch_bounds__(256) void k_copy_or_zero(const float *__restrict__ src, float *__restrict__ dst,
                                                      size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src ? src[i] : 0.0f;
}

// residual / L-curve terms: ||f - y||^2 -> Q_A ; ||w||^2 -> Q_B
__global__ __launch_bounds__(256) void k_resid(const float *__restrict__ f, const float *__restrict__ y,
                                               size_t n, int B, double *part)
{
    const int b = blockIdx.y;
    double e = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double r = (double)f[(size_t)b * n + i] - (double)y[(size_t)b * n + i];
        e += r * r;
    }
    block_sum_store(e, part, Q_A, b, B);
}

__global__ __launch_bounds__(256) void k_sumsq(const float *__restrict__ x, size_t n, int B, int q, double *part)
{
    const int b = blockIdx.y;
    double e = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = x[(size_t)b * n + i];
        e += v * v;
    }
    block_sum_store(e, part, q, b, B);
}

__global__ void k_group_sums(const double *part, int q, int G, int B, int coupled, double *out)
{
    for (int g = threadIdx.x; g < G; g += blockDim.x) out[g] = group_sum(part, q, g, B, coupled);
}

// MATLAB del2(U, 1) of images [B][Nx][Nz] (z fastest = MATLAB [Nz x Nx] columns): per
// dimension the halved centred second difference inside, linear extrapolation of it at
// both ends (n > 3), the middle value (n = 3) or zero (n < 3); (g_z + g_x) / 2.
__device__ __forceinline__ float del2_axis(const float *__restrict__ f, int i, int n, size_t st)
{
    if (n < 3) return 0.0f;
    auto g = [&](int j) { return 0.5f * (f[(size_t)(j + 1) * st] - 2.0f * f[(size_t)j * st] + f[(size_t)(j - 1) * st]); };
    if (n == 3) return g(1);
    if (i == 0) return 2.0f * g(1) - g(2);
    if (i == n - 1) return 2.0f * g(n - 2) - g(n - 3);
    return g(i);
}

__global__ __launch_bounds__(256) void k_del2(const float *__restrict__ in, float *__restrict__ out, int Nz, int Nx,
                                              int B)
{
    const size_t P = (size_t)Nz * Nx, n = P * B;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t b = t / P, r = t - b * P;
        const int ix = (int)(r / Nz), iz = (int)(r - (size_t)ix * Nz);
        const float *img = in + b * P;
        const float gz = del2_axis(img + (size_t)ix * Nz, iz, Nz, 1);
        const float gx = del2_axis(img + iz, ix, Nx, (size_t)Nz);
        out[t] = 0.5f * (gz + gx);
    }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
struct Workspace {
    float *x = nullptr, *w = nullptr, *vm = nullptr, *d = nullptr, *v = nullptr;
    float *ur = nullptr, *uc = nullptr, *us = nullptr, *ts = nullptr, *tp = nullptr;
    float *ur2 = nullptr;                  // second L2 term's block of u
    float *r1 = nullptr, *r2 = nullptr;    // R1 / R2 (or their transposes) applied to a vector
    RegOp op1, op2;
    double *part = nullptr;
    Scal *sc = nullptr;
    double *gs = nullptr;
    float *srv = nullptr;
    std::vector<void *> all;
    ~Workspace()
    {
        for (void *p : all) (void)hipFree(p);
    }
    template <typename T>
    int alloc(T **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(MBREC_ERR_HIP, "mbrec:alloc", "solver workspace allocation failed: %s",
                             hipGetErrorString(e));
        }
        all.push_back(*p);
        return MBREC_OK;
    }
};

// out = R in (transp = 0) or R' in (transp = 1) for a non-identity term; returns the
// pointer the fused kernels read (in itself for an identity term)
static int reg_apply(Plan *p, const RegOp &op, const float *in, float *out, int B, int transp, hipStream_t s,
                     const float **res)
{
    if (op.kind == 0) {
        *res = in;
        return MBREC_OK;
    }
    if (op.kind == 1) {
        const size_t n = (size_t)p->P * B;
        hipLaunchKernelGGL(k_del2, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, s, in, out,
                           p->Nz, p->Nx, B);
        MBREC_HIP(hipGetLastError());
    } else {
        const int rc = op.fn(op.ctx, in, out, B, p->P, transp, (mbrec_stream)s);
        if (rc) return set_error(MBREC_ERR_ARG, "mbrec:regOperator", "L2 term operator callback failed (%d)", rc);
    }
    *res = out;
    return MBREC_OK;
}

static int lsqr_solve(Plan *p, Workspace &ws, const float *y, int B, int G, int maxit, Consts k,
                      hipStream_t s, const int2 *win)
{
    const size_t np = (size_t)p->P, ns = (size_t)p->D * p->nT;
    int rc;
    const float *q1 = nullptr, *q2 = nullptr;
    auto apply_regs = [&](const float *in1, const float *in2, int transp) -> int {
        int e = 0;
        if (k.reg && (e = reg_apply(p, ws.op1, in1, ws.r1, B, transp, s, &q1))) return e;
        if (k.reg2 && (e = reg_apply(p, ws.op2, in2, ws.r2, B, transp, s, &q2))) return e;
        return MBREC_OK;
    };
    // u = b - A_aug w
    if ((rc = op_forward(p, ws.w, ws.ts, B, true, s, win))) return rc;
    hipLaunchKernelGGL(k_init_sino, vgrid(B), dim3(256), 0, s, y, ws.ts, ws.us, ns, B, k, ws.part);
    if ((rc = apply_regs(ws.w, ws.w, 0))) return rc;
    hipLaunchKernelGGL(k_init_pix, vgrid(B), dim3(256), 0, s, ws.w, ws.vm, q1, q2, ws.ur, ws.ur2, ws.uc, ws.x,
                       ws.d, np, B, k, ws.part);
    hipLaunchKernelGGL(k_scal_init, dim3(1), dim3(256), 0, s, ws.sc, ws.part, G, B, k);
    // v = A_aug' u
    if ((rc = op_transpose(p, ws.us, ws.tp, B, true, s, win))) return rc;
    if ((rc = apply_regs(ws.ur, ws.ur2, 1))) return rc;
    hipLaunchKernelGGL(k_v_update, vgrid(B), dim3(256), 0, s, ws.tp, q1, q2, ws.uc, ws.v, np, B, k, ws.sc, 1,
                       ws.part);
    hipLaunchKernelGGL(k_scal_v0, dim3(1), dim3(256), 0, s, ws.sc, ws.part, G, B, k);
    for (int ii = 1; ii <= maxit; ++ii) {
        if ((rc = op_forward(p, ws.v, ws.ts, B, true, s, win))) return rc;
        hipLaunchKernelGGL(k_u_sino, vgrid(B), dim3(256), 0, s, ws.ts, ws.us, ns, B, k, ws.sc, ws.part);
        if ((rc = apply_regs(ws.v, ws.v, 0))) return rc;
        hipLaunchKernelGGL(k_u_pix, vgrid(B), dim3(256), 0, s, ws.v, q1, q2, ws.ur, ws.ur2, ws.uc, np, B, k, ws.sc,
                           ws.part);
        hipLaunchKernelGGL(k_scal_u, dim3(1), dim3(256), 0, s, ws.sc, ws.part, G, B, k);
        hipLaunchKernelGGL(k_dx, vgrid(B), dim3(256), 0, s, ws.v, ws.d, ws.x, np, B, k, ws.sc, ws.part);
        // The transpose product that closes the last iteration feeds only the
        // stopping tests and the next direction; x is final here, and on
        // non-convergence the minimal-residual iterate is this last one (the
        // residual estimate |s| normr never increases), so that work is dead.
        if (ii == maxit) break;
        if ((rc = op_transpose(p, ws.us, ws.tp, B, true, s, win))) return rc;
        if ((rc = apply_regs(ws.ur, ws.ur2, 1))) return rc;
        hipLaunchKernelGGL(k_v_update, vgrid(B), dim3(256), 0, s, ws.tp, q1, q2, ws.uc, ws.v, np, B, k, ws.sc, 0,
                           ws.part);
        hipLaunchKernelGGL(k_scal_v, dim3(1), dim3(256), 0, s, ws.sc, ws.part, G, B, k, ii);
    }
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

// A prepared NNLS-L2 solve: workspace and constants for B sinograms on one plan.
struct L2Solve {
    Workspace ws;
    Consts k{};
    int B = 0, G = 0, num_iter = 0, reg = 0;
};

// Allocate the workspace, size the plan's operator workspace for B and upload the
// per-sinogram sqrt(lambdaL2) (lambda_b, host [B], may be null: one lambdaL2 for all).
// op1: the first L2 term's operator (identity when null); lambda2 / op2: the second term.
static int l2_prepare(Plan *p, L2Solve &c, int B, int coupled, int num_iter, double lambdaL2,
                      const double *lambda_b, hipStream_t s, const RegOp *op1 = nullptr, double lambda2 = 0.0,
                      const RegOp *op2 = nullptr)
{
    const size_t np = (size_t)p->P, ns = (size_t)p->D * p->nT;
    Workspace &ws = c.ws;
    int rc;
    int reg = lambdaL2 > 0.0;
    std::vector<float> srv;
    if (lambda_b) {
        reg = 0;
        srv.resize(B);
        for (int b = 0; b < B; ++b) {
            if (!(lambda_b[b] >= 0.0))
                return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_L2_reg: lambdaL2[%d] must be >= 0", b);
            srv[b] = (float)std::sqrt(lambda_b[b]);
            reg |= lambda_b[b] > 0.0;
        }
    }
    const int G = coupled ? 1 : B;
    if ((rc = ws.alloc(&ws.x, np * B)) || (rc = ws.alloc(&ws.vm, np * B)) ||
        (rc = ws.alloc(&ws.d, np * B)) || (rc = ws.alloc(&ws.v, np * B)) ||
        (rc = ws.alloc(&ws.uc, np * B)) || (rc = ws.alloc(&ws.tp, np * B)) ||
        (rc = ws.alloc(&ws.ur, reg ? np * B : 1)) || (rc = ws.alloc(&ws.us, ns * B)) ||
        (rc = ws.alloc(&ws.ts, ns * B)) || (rc = ws.alloc(&ws.part, (size_t)kMaxQ * B * kRedBlocks)) ||
        (rc = ws.alloc(&ws.sc, G)) || (rc = ws.alloc(&ws.gs, 3 * (size_t)G)) ||
        (rc = ws.alloc(&ws.srv, lambda_b ? (size_t)B : 1)))
        return rc;
    const int reg2 = lambda2 > 0.0;
    if (op1) ws.op1 = *op1;
    if (op2) ws.op2 = *op2;
    if ((rc = ws.alloc(&ws.ur2, reg2 ? np * B : 1)) || (rc = ws.alloc(&ws.r1, ws.op1.kind != 0 ? np * B : 1)) ||
        (rc = ws.alloc(&ws.r2, ws.op2.kind != 0 ? np * B : 1)))
        return rc;
    if (lambda_b) {
        MBREC_HIP(hipMemcpyAsync(ws.srv, srv.data(), sizeof(float) * B, hipMemcpyHostToDevice, s));
        MBREC_HIP(hipStreamSynchronize(s));   // srv is a host temporary
    }
    if ((rc = ensure_work(p, B))) return rc;
    Consts &k = c.k;
    k.sl = 1.0f;                                   // lambda = 1 (rec_nn_with_L2_reg.m:33)
    k.sb = (float)std::sqrt(1e-3);                 // beta = 1e-3 (:34)
    k.sr = (float)(reg && !lambda_b ? std::sqrt(lambdaL2) : 0.0);
    k.reg = reg;
    k.tol = 1e-8;                                  // inner_tol (nnls_L2_less.m:33)
    k.coupled = coupled;
    k.srv = lambda_b ? ws.srv : nullptr;
    k.sr2 = (float)(reg2 ? std::sqrt(lambda2) : 0.0);
    k.reg2 = reg2;
    c.B = B;
    c.G = G;
    c.num_iter = num_iter;
    c.reg = reg;
    return MBREC_OK;
}

// Launch the whole solve on stream s (kernels only: no allocation, copy or
// synchronisation, so the sequence can be captured into a graph).
static int l2_enqueue(Plan *p, L2Solve &c, const float *y, const float *init, float *rec, const int2 *win,
                      hipStream_t s, const float *init_dual = nullptr)
{
    const size_t np = (size_t)p->P;
    Workspace &ws = c.ws;
    const int B = c.B;
    int rc;
    ws.w = rec;   // the outer iterate lives in the output buffer
    const int nb = 1024;
    // w0 = init (warm start) or 0 (nnls_L2_less.m:28); v0 = init_dual (warm start) or 0 (:29)
    hipLaunchKernelGGL(k_copy_or_zero, dim3(nb), dim3(256), 0, s, init, ws.w, np * B);
    hipLaunchKernelGGL(k_copy_or_zero, dim3(nb), dim3(256), 0, s, init_dual, ws.vm, np * B);
    for (int it = 0; it < c.num_iter; ++it) {
        if ((rc = lsqr_solve(p, ws, y, B, c.G, 2, c.k, s, win))) return rc;
        hipLaunchKernelGGL(k_nnls_update, vgrid(B), dim3(256), 0, s, ws.x, ws.w, ws.vm, np, B, c.k, ws.sc);
    }
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

// lambda_b (host, [B], may be null): per-sinogram lambdaL2 of independent solves (coupled = 0).
// With terms (host, nterms <= 2): the L2 terms' weights and operators (lambdaL2 and
// lambda_b unused); lcurve then receives 3 values per group: ||A rec - y||^2, ||R1 rec||^2,
// ||R2 rec||^2.
int nnls_l2(Plan *p, const float *y, int B, int coupled, int num_iter, double lambdaL2,
            const float *init, float *rec, double *lcurve, hipStream_t s, const double *lambda_b = nullptr,
            const float *init_dual = nullptr, float *final_dual = nullptr, const mbrec_l2_term *terms = nullptr,
            int nterms = 0)
{
    const size_t np = (size_t)p->P, ns = (size_t)p->D * p->nT;
    L2Solve c;
    int rc;
    RegOp ops[2];
    double lam2 = 0.0;
    if (terms) {
        lambdaL2 = nterms > 0 ? terms[0].lambda : 0.0;
        lam2 = nterms > 1 && lambdaL2 > 0.0 ? terms[1].lambda : 0.0;
        for (int j = 0; j < nterms; ++j) {
            ops[j].kind = terms[j].kind;
            ops[j].fn = terms[j].apply;
            ops[j].ctx = terms[j].ctx;
        }
    }
    if ((rc = l2_prepare(p, c, B, coupled, num_iter, lambdaL2, lambda_b, s, &ops[0], lam2, &ops[1]))) return rc;
    if ((rc = l2_enqueue(p, c, y, init, rec, nullptr, s, init_dual))) return rc;
    Workspace &ws = c.ws;
    if (final_dual)   // (rec, final_dual) is the complete state: a later call continues from it
        MBREC_HIP(hipMemcpyAsync(final_dual, ws.vm, sizeof(float) * np * B, hipMemcpyDeviceToDevice, s));
    const int G = c.G;
    if (lcurve) {
        // rec_nn_with_L2_reg.m:51,54: ||A rec - y||^2 and ||RegL2 rec||^2 per group
        if ((rc = op_forward(p, ws.w, ws.ts, B, true, s))) return rc;
        // rec_nn_with_L2_reg.m:51-64 with the terms' operators: ||R1 rec||^2, ||R2 rec||^2
        const float *q1 = ws.w, *q2 = ws.w;
        if ((rc = reg_apply(p, ws.op1, ws.w, ws.r1, B, 0, s, &q1)) ||
            (rc = reg_apply(p, ws.op2, ws.w, ws.r2, B, 0, s, &q2)))
            return rc;
        hipLaunchKernelGGL(k_resid, vgrid(B), dim3(256), 0, s, ws.ts, y, ns, B, ws.part);
        hipLaunchKernelGGL(k_sumsq, vgrid(B), dim3(256), 0, s, q1, np, B, (int)Q_B, ws.part);
        hipLaunchKernelGGL(k_sumsq, vgrid(B), dim3(256), 0, s, q2, np, B, (int)Q_C, ws.part);
        hipLaunchKernelGGL(k_group_sums, dim3(1), dim3(256), 0, s, ws.part, (int)Q_A, G, B, coupled, ws.gs);
        hipLaunchKernelGGL(k_group_sums, dim3(1), dim3(256), 0, s, ws.part, (int)Q_B, G, B, coupled, ws.gs + G);
        hipLaunchKernelGGL(k_group_sums, dim3(1), dim3(256), 0, s, ws.part, (int)Q_C, G, B, coupled, ws.gs + 2 * G);
        std::vector<double> h(3 * (size_t)G);
        MBREC_HIP(hipMemcpyAsync(h.data(), ws.gs, sizeof(double) * 3 * G, hipMemcpyDeviceToHost, s));
        MBREC_HIP(hipStreamSynchronize(s));
        const int per = terms ? 3 : 2;
        for (int g = 0; g < G; ++g) {
            lcurve[per * g] = h[g];
            lcurve[per * g + 1] = (terms ? nterms > 0 : c.reg) ? h[G + g] : 0.0;
            if (terms) lcurve[per * g + 2] = nterms > 1 ? h[2 * G + g] : 0.0;
        }
    } else {
        MBREC_HIP(hipStreamSynchronize(s));   // workspace is freed on return
    }
    return MBREC_OK;
}

}  // namespace mbrec

using namespace mbrec;

extern "C" int mbrec_rec_nn_with_L2_reg_dev(mbrec_plan h, const float *data, int B, int coupled,
                                            int num_iter, double lambdaL2, const float *init,
                                            float *rec, double *lcurve, mbrec_stream s)
{
    if (!h || !data || !rec || B <= 0 || num_iter < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_L2_reg: invalid arguments");
    return nnls_l2(&h->p, data, B, coupled, num_iter, lambdaL2, init, rec, lcurve, (hipStream_t)s);
}

extern "C" int mbrec_del2_dev(const float *in, float *out, int Nz, int Nx, int B, mbrec_stream s)
{
    if (!in || !out || Nz <= 0 || Nx <= 0 || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "del2: invalid arguments");
    const size_t n = (size_t)Nz * Nx * B;
    hipLaunchKernelGGL(k_del2, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, (hipStream_t)s, in,
                       out, Nz, Nx, B);
    MBREC_HIP(hipGetLastError());
    return MBREC_OK;
}

extern "C" int mbrec_rec_nn_with_L2_reg_terms_dev(mbrec_plan h, const float *data, int B, int coupled,
                                                  int num_iter, const mbrec_l2_term *terms, int nterms,
                                                  const float *init, float *rec, double *lcurve, mbrec_stream s)
{
    if (!h || !data || !rec || B <= 0 || num_iter < 0 || nterms < 0 || nterms > 2 || (nterms && !terms))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_L2_reg: invalid arguments");
    for (int j = 0; j < nterms; ++j) {
        if (!(terms[j].lambda >= 0.0))
            return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_L2_reg: term %d weight must be >= 0", j + 1);
        if (terms[j].kind < MBREC_REG_IDENTITY || terms[j].kind > MBREC_REG_CALLBACK ||
            (terms[j].kind == MBREC_REG_CALLBACK && !terms[j].apply))
            return set_error(MBREC_ERR_ARG, "mbrec:regOperator", "rec_nn_with_L2_reg: term %d has no valid operator",
                             j + 1);
    }
    // incorporate_l2_reg_into_model_functions.m:4-21: the second term enters the operator
    // only together with the first (lambdaL2 > 0); the L-curve norms are reported for every
    // term given (rec_nn_with_L2_reg.m:55-64)
    return nnls_l2(&h->p, data, B, coupled, num_iter, 0.0, init, rec, lcurve, (hipStream_t)s, nullptr, nullptr,
                   nullptr, terms, nterms);
}

extern "C" int mbrec_rec_nn_with_L2_reg_batch_dev(mbrec_plan h, const float *data, int B, int num_iter,
                                                  const double *lambdaL2, const float *init, const float *init_dual,
                                                  float *rec, float *final_dual, double *lcurve, mbrec_stream s)
{
    if (!h || !data || !rec || !lambdaL2 || B <= 0 || num_iter < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "rec_nn_with_L2_reg_batch: invalid arguments");
    return nnls_l2(&h->p, data, B, 0, num_iter, 0.0, init, rec, lcurve, (hipStream_t)s, lambdaL2, init_dual,
                   final_dual);
}

// ---------------------------------------------------------------------------
// Batched solves over several geometries in one graph launch.
//
// A batch holds G groups; group g is a plan (grid + travel times on its own
// sample axis of nT_g samples) and B_g sinograms, each with a sample window
// [first, first + count) of that axis (a sinogram cropped at its start and/or
// end) and its own lambdaL2.  All groups' complete 50-iteration-style solves
// are captured once into a HIP graph -- one branch per group, so groups of
// different sizes run concurrently -- and every later run is a single graph
// launch.  Buffers are the caller's (their addresses are part of the graph).
// ---------------------------------------------------------------------------
namespace mbrec {

__global__ __launch_bounds__(256) void k_mask_sino(const float *__restrict__ in, float *__restrict__ out,
                                                   const int2 *__restrict__ win, int D, int nT, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(i % nT);
        const int b = (int)(i / ((size_t)D * nT));
        const int2 w = win[b];
        out[i] = (t >= w.x && t < w.y) ? in[i] : 0.0f;
    }
}

struct BatchGroup {
    Plan *p = nullptr;
    int B = 0;
    const float *data = nullptr;
    float *rec = nullptr;
    const float *init = nullptr;
    const float *init_dual = nullptr;
    float *ys = nullptr;       // masked copy of the data
    int2 *win = nullptr;       // [B] sample windows
    L2Solve solve;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    bool pinned = false;
};

struct Batch {
    std::vector<BatchGroup> g;
    int num_iter = 0;
    hipStream_t cap = nullptr;
    hipEvent_t fork = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

static void batch_free(Batch *b)
{
    for (int i = 0; i < 2; ++i) {
        if (b->exec[i]) (void)hipGraphExecDestroy(b->exec[i]);
        if (b->graph[i]) (void)hipGraphDestroy(b->graph[i]);
    }
    for (BatchGroup &g : b->g) {
        if (g.pinned) --g.p->pins;
        if (g.ys) (void)hipFree(g.ys);
        if (g.win) (void)hipFree(g.win);
        if (g.st) (void)hipStreamDestroy(g.st);
        if (g.done) (void)hipEventDestroy(g.done);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->cap) (void)hipStreamDestroy(b->cap);
}

// all groups' solves, group g on stream st_g, forked from / joined into stream s
static int batch_enqueue(Batch *b, int use_init, hipStream_t s)
{
    MBREC_HIP(hipEventRecord(b->fork, s));
    for (BatchGroup &g : b->g) {
        MBREC_HIP(hipStreamWaitEvent(g.st, b->fork, 0));
        const size_t n = (size_t)g.B * g.p->D * g.p->nT;
        hipLaunchKernelGGL(k_mask_sino, dim3(1024), dim3(256), 0, g.st, g.data, g.ys, g.win, g.p->D, g.p->nT, n);
        int rc = l2_enqueue(g.p, g.solve, g.ys, use_init ? g.init : nullptr, g.rec, g.win, g.st,
                            use_init ? g.init_dual : nullptr);
        if (rc) return rc;
        MBREC_HIP(hipEventRecord(g.done, g.st));
        MBREC_HIP(hipStreamWaitEvent(s, g.done, 0));
    }
    return MBREC_OK;
}

}  // namespace mbrec

struct mbrec_batch_s {
    mbrec::Batch b;
};

extern "C" int mbrec_batch_destroy(mbrec_batch h)
{
    if (!h) return MBREC_OK;
    (void)hipDeviceSynchronize();
    batch_free(&h->b);
    delete h;
    return MBREC_OK;
}

extern "C" int mbrec_batch_create(mbrec_batch *out, int ngroups, const mbrec_plan *plans, const int *group_B,
                                  const float *const *data, float *const *rec, const float *const *init,
                                  const float *const *init_dual, const int32_t *windows, const double *lambdaL2,
                                  int num_iter)
{
    if (!out || ngroups <= 0 || !plans || !group_B || !data || !rec || !windows || !lambdaL2 || num_iter < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "batch_create: invalid arguments");
    *out = nullptr;
    for (int i = 0; i < ngroups; ++i) {
        if (!plans[i] || group_B[i] <= 0 || !data[i] || !rec[i])
            return set_error(MBREC_ERR_ARG, "mbrec:args", "batch_create: group %d: invalid plan, size or buffer", i);
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i])
                return set_error(MBREC_ERR_ARG, "mbrec:args",
                                 "batch_create: groups %d and %d share a plan (merge them into one group)", j, i);
    }
    mbrec_batch h = new mbrec_batch_s;
    Batch *b = &h->b;
    b->num_iter = num_iter;
    b->g.resize(ngroups);
    auto fail = [&](int rc) {
        mbrec_batch_destroy(h);
        return rc;
    };
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return fail(set_error(MBREC_ERR_HIP, "mbrec:hip", "batch_create: stream / event creation failed"));
    size_t off = 0;
    for (int i = 0; i < ngroups; ++i) {
        BatchGroup &g = b->g[i];
        g.p = &plans[i]->p;
        g.B = group_B[i];
        g.data = data[i];
        g.rec = rec[i];
        g.init = init ? init[i] : nullptr;
        g.init_dual = init_dual ? init_dual[i] : nullptr;
        std::vector<int2> w(g.B);
        for (int k = 0; k < g.B; ++k) {
            const long long first = windows[2 * (off + k)], count = windows[2 * (off + k) + 1];
            if (first < 0 || count <= 0 || first + count > g.p->nT)
                return fail(set_error(MBREC_ERR_ARG, "mbrec:window",
                                      "batch_create: sinogram %zu window [%lld, %lld) outside [0, %d)", off + k,
                                      first, first + count, g.p->nT));
            w[k] = make_int2((int)first, (int)(first + count));
        }
        if (hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&g.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void **)&g.win, sizeof(int2) * g.B) != hipSuccess ||
            hipMalloc((void **)&g.ys, sizeof(float) * (size_t)g.B * g.p->D * g.p->nT) != hipSuccess ||
            hipMemcpy(g.win, w.data(), sizeof(int2) * g.B, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            return fail(set_error(MBREC_ERR_HIP, "mbrec:hip", "batch_create: group %d allocation failed", i));
        }
        int rc = l2_prepare(g.p, g.solve, g.B, 0, num_iter, 0.0, lambdaL2 + off, g.st);
        if (rc) return fail(rc);
        ++g.p->pins;   // its workspace and tables must stay where the graph saw them
        g.pinned = true;
        off += g.B;
    }
    *out = h;
    return MBREC_OK;
}

extern "C" int mbrec_batch_set_lambda(mbrec_batch h, const double *lambdaL2)
{
    if (!h || !lambdaL2) return set_error(MBREC_ERR_ARG, "mbrec:args", "batch_set_lambda: invalid arguments");
    Batch *b = &h->b;
    size_t off = 0;
    for (BatchGroup &g : b->g) {
        if (!g.solve.k.reg) {
            for (int k = 0; k < g.B; ++k)
                if (lambdaL2[off + k] > 0.0)
                    return set_error(MBREC_ERR_ARG, "mbrec:args",
                                     "batch_set_lambda: the batch was created without an L2 term in group %d",
                                     (int)(&g - b->g.data()));
        }
        std::vector<float> srv(g.B);
        for (int k = 0; k < g.B; ++k) {
            if (!(lambdaL2[off + k] >= 0.0))
                return set_error(MBREC_ERR_ARG, "mbrec:args", "batch_set_lambda: lambdaL2 must be >= 0");
            srv[k] = (float)std::sqrt(lambdaL2[off + k]);
        }
        MBREC_HIP(hipMemcpy(g.solve.ws.srv, srv.data(), sizeof(float) * g.B, hipMemcpyHostToDevice));
        off += g.B;
    }
    return MBREC_OK;
}

extern "C" int mbrec_batch_run(mbrec_batch h, int use_init, int use_graph, mbrec_stream stream)
{
    if (!h) return set_error(MBREC_ERR_ARG, "mbrec:args", "batch_run: null batch");
    Batch *b = &h->b;
    use_init = use_init ? 1 : 0;
    if (use_init)
        for (BatchGroup &g : b->g)
            if (!g.init) return set_error(MBREC_ERR_ARG, "mbrec:args", "batch_run: use_init without init buffers");
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) return batch_enqueue(b, use_init, s);
    if (!b->exec[use_init]) {
        MBREC_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        int rc = batch_enqueue(b, use_init, b->cap);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        MBREC_HIP(e);
        b->graph[use_init] = graph;
        MBREC_HIP(hipGraphInstantiate(&b->exec[use_init], graph, nullptr, nullptr, 0));
    }
    MBREC_HIP(hipGraphLaunch(b->exec[use_init], s));
    return MBREC_OK;
}
