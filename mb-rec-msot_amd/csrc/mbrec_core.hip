// mbrec_core.hip -- plans, device tables, error state and the device API.
#include "mbrec_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

namespace mbrec {

static thread_local char g_err[512] = "";
static thread_local char g_err_id[128] = "";

int set_error(int code, const char *id, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    snprintf(g_err_id, sizeof(g_err_id), "%s", id ? id : "mbrec:error");
    return code;
}

int hip_check(hipError_t e, const char *what)
{
    return set_error(MBREC_ERR_HIP, "mbrec:hip", "%s failed: %s", what, hipGetErrorString(e));
}

template <typename T>
static int dalloc(T **ptr, size_t n, size_t *acc)
{
    *ptr = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)ptr, n * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(MBREC_ERR_HIP, "mbrec:alloc", "hipMalloc of %zu bytes failed: %s",
                         n * sizeof(T), hipGetErrorString(e));
    }
    if (acc) *acc += n * sizeof(T);
    return MBREC_OK;
}

template <typename T>
static int upload(T *dst, const T *src, size_t n)
{
    MBREC_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return MBREC_OK;
}

int ensure_bins(Plan *p, int B)
{
    if (B <= p->bins_bcap) return MBREC_OK;
    if (p->d_bins) {
        MBREC_HIP(hipDeviceSynchronize());
        MBREC_HIP(hipFree(p->d_bins));
        p->d_bins = nullptr;
    }
    const size_t n = (size_t)p->S * B * p->D * p->geom.NB;
    int rc = dalloc(&p->d_bins, n, nullptr);
    if (rc) { p->bins_bcap = 0; return rc; }
    p->bins_bcap = B;
    return MBREC_OK;
}

int ensure_tmp(Plan *p, int B)
{
    if (B <= p->tmp_bcap) return MBREC_OK;
    if (p->d_img_tmp) {
        MBREC_HIP(hipDeviceSynchronize());
        MBREC_HIP(hipFree(p->d_img_tmp));
        p->d_img_tmp = nullptr;
    }
    int rc = dalloc(&p->d_img_tmp, (size_t)2 * B * p->P, nullptr);
    if (rc) { p->tmp_bcap = 0; return rc; }
    p->tmp_bcap = B;
    return MBREC_OK;
}

// Breakpoint geometry and per-phase hat taps for radius R.
static int make_geometry(Plan *p, double maxabs_tau)
{
    const double R = p->R;
    if (!(R > 0.0) || R > 30.0)
        return set_error(MBREC_ERR_ARG, "mbrec:radius", "interpolation radius R=%g outside (0, 30]", R);
    const int fR = (int)std::floor(R);
    double fr = R - fR;
    double a = fr, b = (fr == 0.0) ? 0.0 : 1.0 - fr;
    double phi1 = std::min(a, b), phi2 = std::max(a, b);
    BinGeom &g = p->geom;
    g.phi1 = (float)phi1;
    g.phi2 = (float)phi2;
    const double w0 = phi1, w1 = phi2 - phi1, w2 = 1.0 - phi2;
    g.iw0 = w0 > 0 ? (float)(1.0 / w0) : 0.0f;
    g.iw1 = w1 > 0 ? (float)(1.0 / w1) : 0.0f;
    g.iw2 = w2 > 0 ? (float)(1.0 / w2) : 0.0f;
    g.K = fR + 2;
    g.nT = p->nT;
    g.NB = 3 * (p->nT + 2 * g.K) + 1;
    // fixed-point format: |tau| + K + shifts must stay below 2^(31-F)
    const double need = maxabs_tau + g.K + 4.0;
    int ib = (int)std::ceil(std::log2(std::max(need, 2.0)));
    int F = 30 - ib;
    if (F > 20) F = 20;
    if (F < 8)
        return set_error(MBREC_ERR_RANGE, "mbrec:range",
                         "travel times up to %g samples exceed the fixed-point range", maxabs_tau);
    g.F = F;
    HatTaps &t = p->taps;
    t.kmin = -fR - 1;
    t.nw = 2 * fR + 3;
    if (t.nw > MBREC_MAX_TAPS)
        return set_error(MBREC_ERR_ARG, "mbrec:radius", "R=%g needs too many taps", R);
    const double ph[3] = {0.0, phi1, phi2};
    for (int m = 0; m < 3; ++m)
        for (int i = 0; i < MBREC_MAX_TAPS; ++i) {
            double w = 0.0;
            if (i < t.nw) {
                const double k = t.kmin + i;
                w = 1.0 - std::fabs(k - ph[m]) / R;
                if (w < 0.0) w = 0.0;
            }
            t.w[m][i] = (float)w;
        }
    return MBREC_OK;
}

// Device pixel order: TZ x TX tiles, tiles column-major, lz fastest in a tile.
static void make_pixel_order(Plan *p, std::vector<int32_t> &slot_pix)
{
    if (p->Nx == 1) { p->TZ = 64; p->TX = 1; } else { p->TZ = 8; p->TX = 8; }
    const int ntz = (p->Nz + p->TZ - 1) / p->TZ, ntx = (p->Nx + p->TX - 1) / p->TX;
    p->Pd = ntz * ntx * p->TZ * p->TX;
    slot_pix.assign(p->Pd, -1);
    int slot = 0;
    for (int tx = 0; tx < ntx; ++tx)
        for (int tz = 0; tz < ntz; ++tz)
            for (int lx = 0; lx < p->TX; ++lx)
                for (int lz = 0; lz < p->TZ; ++lz, ++slot) {
                    const int iz = tz * p->TZ + lz, ix = tx * p->TX + lx;
                    if (iz < p->Nz && ix < p->Nx) slot_pix[slot] = iz + p->Nz * ix;
                }
}

static void free_plan(Plan *p)
{
    void *ptrs[] = {p->d_pix, p->d_tq, p->d_sidx, p->d_sir, p->d_swin, p->d_ir,
                    p->d_kz,  p->d_kx, p->d_bins, p->d_img_tmp, p->d_sino_tmp};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
}

}  // namespace mbrec

using namespace mbrec;

extern "C" {

const char *mbrec_last_error(void) { return g_err; }
const char *mbrec_last_error_id(void) { return g_err_id; }
int mbrec_version(void) { return MBREC_VERSION; }

int mbrec_init(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return set_error(MBREC_ERR_NODEV, "mbrec:nodevice", "no HIP device available");
    }
    if (device < 0 || device >= n)
        return set_error(MBREC_ERR_ARG, "mbrec:device", "device %d out of range (%d devices)", device, n);
    MBREC_HIP(hipSetDevice(device));
    return MBREC_OK;
}

int mbrec_plan_create(mbrec_plan *out, int Nz, int Nx, int D, int nT, double R,
                      const double *tt)
{
    if (!out || !tt || Nz <= 0 || Nx <= 0 || D <= 0 || nT <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_create: invalid arguments");
    *out = nullptr;
    mbrec_plan h = new mbrec_plan_s();
    Plan *p = &h->p;
    p->Nz = Nz; p->Nx = Nx; p->P = Nz * Nx; p->D = D; p->nT = nT; p->R = R;
    double maxabs = 0.0;
    const size_t PD = (size_t)p->P * D;
    for (size_t i = 0; i < PD; ++i) {
        const double a = std::fabs(tt[i]);
        if (std::isfinite(a) && a > maxabs) maxabs = a;
    }
    int rc = make_geometry(p, maxabs);
    if (rc) { delete h; return rc; }
    std::vector<int32_t> slot_pix;
    make_pixel_order(p, slot_pix);
    // fixed-point travel times in device pixel order [D][Pd]
    std::vector<int32_t> tq((size_t)D * p->Pd);
    const double scale = std::ldexp(1.0, p->geom.F);

    for (int d = 0; d < D; ++d)
        for (int s = 0; s < p->Pd; ++s) {
            const int px = slot_pix[s];
            int32_t q = INT32_MIN;   // sentinel: decodes far below the bin range
            if (px >= 0) {
                const double v = tt[(size_t)d * p->P + px];
                if (std::isfinite(v)) q = (int32_t)std::llrint(v * scale);
            }
            tq[(size_t)d * p->Pd + s] = q;
        }
    if ((rc = dalloc(&p->d_pix, p->Pd, &p->table_bytes)) ||
        (rc = dalloc(&p->d_tq, (size_t)D * p->Pd, &p->table_bytes)) ||
        (rc = upload(p->d_pix, slot_pix.data(), slot_pix.size())) ||
        (rc = upload(p->d_tq, tq.data(), tq.size()))) {
        free_plan(p);
        delete h;
        return rc;
    }
    *out = h;
    return MBREC_OK;
}

int mbrec_plan_set_sir(mbrec_plan h, const int32_t *sirIdx, const double *sir, int Ls, int Nsir,
                       const int32_t *se)
{
    if (!h || !sirIdx || !sir || !se || Ls <= 0 || Nsir <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_set_sir: invalid arguments");
    Plan *p = &h->p;
    const size_t tot = (size_t)Ls * Nsir;
    std::vector<int32_t> win(2 * (size_t)Nsir);
    int maxlen = 0;
    for (int k = 0; k < Nsir; ++k) {
        const long long a = (long long)se[2 * (size_t)k] - 1, b = (long long)se[2 * (size_t)k + 1] - 1;
        const long long len = b >= a ? b - a + 1 : 0;
        if (len > 0 && ((long long)k * Ls + a < 0 || (long long)k * Ls + b >= (long long)tot))
            return set_error(MBREC_ERR_ARG, "mbrec:sirWindow",
                             "SIR window %d [%lld, %lld] reads outside the SIR table", k + 1, a + 1, b + 1);
        win[2 * (size_t)k] = (int32_t)a;
        win[2 * (size_t)k + 1] = (int32_t)len;
        maxlen = std::max(maxlen, (int)len);
    }
    std::vector<int32_t> sidx((size_t)p->D * p->Pd, 0);
    std::vector<int32_t> slot_pix;
    make_pixel_order(p, slot_pix);
    for (int d = 0; d < p->D; ++d)
        for (int s = 0; s < p->Pd; ++s) {
            const int px = slot_pix[s];
            if (px < 0) continue;
            const int32_t k = sirIdx[(size_t)d * p->P + px];
            if (k < 1 || k > Nsir)
                return set_error(MBREC_ERR_ARG, "mbrec:sirIndex",
                                 "SIR index %d at pixel %d detector %d outside [1, %d]", k, px, d, Nsir);
            sidx[(size_t)d * p->Pd + s] = k - 1;
        }
    std::vector<float> sirf(tot);
    for (size_t i = 0; i < tot; ++i) sirf[i] = (float)sir[i];
    if (p->d_sidx) { (void)hipFree(p->d_sidx); (void)hipFree(p->d_sir); (void)hipFree(p->d_swin); }
    int rc;
    if ((rc = dalloc(&p->d_sidx, sidx.size(), &p->table_bytes)) ||
        (rc = dalloc(&p->d_sir, sirf.size(), &p->table_bytes)) ||
        (rc = dalloc(&p->d_swin, win.size(), &p->table_bytes)) ||
        (rc = upload(p->d_sidx, sidx.data(), sidx.size())) ||
        (rc = upload(p->d_sir, sirf.data(), sirf.size())) ||
        (rc = upload(p->d_swin, win.data(), win.size())))
        return rc;
    p->use_sir = 1;
    p->Ls = Ls;
    p->Nsir = Nsir;
    p->sir_maxlen = maxlen;
    return MBREC_OK;
}

int mbrec_plan_set_ir(mbrec_plan h, const double *ir, int Lir, int per_det)
{
    if (!h || !ir || Lir <= 0 || (Lir % 2) == 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_set_ir: IR length must be odd and positive");
    Plan *p = &h->p;
    const size_t n = (size_t)Lir * (per_det ? p->D : 1);
    std::vector<float> f(n);
    // host layout [Lir x D] column-major = per detector contiguous
    for (size_t i = 0; i < n; ++i) f[i] = (float)ir[i];
    if (p->d_ir) (void)hipFree(p->d_ir);
    int rc;
    if ((rc = dalloc(&p->d_ir, n, &p->table_bytes)) || (rc = upload(p->d_ir, f.data(), n))) return rc;
    p->Lir = Lir;
    p->ir_per_det = per_det ? 1 : 0;
    return MBREC_OK;
}

int mbrec_plan_set_prefilter(mbrec_plan h, const double *kz, int Kz, const double *kx, int Kx)
{
    if (!h || !kz || !kx || Kz < 0 || Kx < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_set_prefilter: invalid arguments");
    Plan *p = &h->p;
    if (2 * Kz + 1 > p->Nz || 2 * Kx + 1 > p->Nx)
        return set_error(MBREC_ERR_ARG, "mbrec:prefilter",
                         "prefilter taps (%d, %d) wrap around the %dx%d grid", 2 * Kz + 1, 2 * Kx + 1,
                         p->Nz, p->Nx);
    std::vector<float> fz(2 * Kz + 1), fx(2 * Kx + 1);
    for (int i = 0; i < 2 * Kz + 1; ++i) fz[i] = (float)kz[i];
    for (int i = 0; i < 2 * Kx + 1; ++i) fx[i] = (float)kx[i];
    if (p->d_kz) { (void)hipFree(p->d_kz); (void)hipFree(p->d_kx); }
    int rc;
    if ((rc = dalloc(&p->d_kz, fz.size(), &p->table_bytes)) ||
        (rc = dalloc(&p->d_kx, fx.size(), &p->table_bytes)) ||
        (rc = upload(p->d_kz, fz.data(), fz.size())) || (rc = upload(p->d_kx, fx.data(), fx.size())))
        return rc;
    p->Kz = Kz;
    p->Kx = Kx;
    return MBREC_OK;
}

int mbrec_plan_set_sigma(mbrec_plan h, double sigma)
{
    if (!h || !(sigma > 0.0) || !std::isfinite(sigma))
        return set_error(MBREC_ERR_ARG, "mbrec:sigma", "normalisation factor must be positive");
    h->p.sigma = sigma;
    return MBREC_OK;
}

int mbrec_plan_destroy(mbrec_plan h)
{
    if (!h) return MBREC_OK;
    (void)hipDeviceSynchronize();
    free_plan(&h->p);
    delete h;
    return MBREC_OK;
}

size_t mbrec_plan_device_bytes(mbrec_plan h)
{
    if (!h) return 0;
    const Plan *p = &h->p;
    return p->table_bytes + (size_t)p->S * p->bins_bcap * p->D * p->geom.NB * sizeof(float) +
           (size_t)2 * p->tmp_bcap * p->P * sizeof(float);
}

int mbrec_plan_tau_frac_bits(mbrec_plan h) { return h ? h->p.geom.F : -1; }

int mbrec_mex_forward_dev(mbrec_plan h, const float *img, float *sino, int B, mbrec_stream s)
{
    if (!h || !img || !sino || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "invalid arguments");
    return op_forward(&h->p, img, sino, B, false, (hipStream_t)s);
}

int mbrec_mex_transpose_dev(mbrec_plan h, const float *sino, float *img, int B, mbrec_stream s)
{
    if (!h || !img || !sino || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "invalid arguments");
    return op_transpose(&h->p, sino, img, B, false, (hipStream_t)s);
}

int mbrec_model_forward_dev(mbrec_plan h, const float *img, float *sino, int B, mbrec_stream s)
{
    if (!h || !img || !sino || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "invalid arguments");
    return op_forward(&h->p, img, sino, B, true, (hipStream_t)s);
}

int mbrec_model_transpose_dev(mbrec_plan h, const float *sino, float *img, int B, mbrec_stream s)
{
    if (!h || !img || !sino || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "invalid arguments");
    return op_transpose(&h->p, sino, img, B, true, (hipStream_t)s);
}

int mbrec_plan_time_stage(mbrec_plan h, int stage, int B, int reps, mbrec_stream s, float *ms)
{
    if (!h || !ms || B <= 0 || reps <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_time_stage: invalid arguments");
    return time_stage(&h->p, stage, B, reps, (hipStream_t)s, ms);
}

}  // extern "C"
