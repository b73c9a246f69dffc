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

static void free_work(Plan *p)
{
    float *bufs[] = {p->d_bins, p->d_hpart, p->d_imgI, p->d_img_tmp, p->d_sino_tmp, p->d_gpart, p->d_sband, p->d_img_slot};
    for (float *q : bufs)
        if (q) (void)hipFree(q);
    p->d_bins = p->d_hpart = p->d_imgI = p->d_img_tmp = p->d_sino_tmp = p->d_gpart = p->d_sband = nullptr;
    p->d_img_slot = nullptr;
    p->sband_nwave = p->sband_nfb = 0;
    p->bcap = 0;
    p->work_bytes = 0;
}


int ensure_work(Plan *p, int B)
{
    const size_t need_h = std::max<size_t>((size_t)p->D * p->nbk * p->WL, 1);
    if (B <= p->bcap && p->d_hpart && p->work_bytes && need_h * batch_stride(p->bcap) * sizeof(float) <=
                                                             p->work_bytes)
        return MBREC_OK;
    if (p->pins)
        return set_error(MBREC_ERR_ARG, "mbrec:planPinned",
                         "plan workspace is held by a batch graph (batch size %d > %d); destroy the batch first", B,
                         p->bcap);
    if (p->bcap) {
        MBREC_HIP(hipDeviceSynchronize());
        free_work(p);
    }
    const int Bp = batch_stride(B);
    size_t acc = 0;
    int rc;
    if ((rc = dalloc(&p->d_bins, ((size_t)p->D * p->geom.NB + 2) * Bp, &acc)) ||
        (rc = dalloc(&p->d_hpart, need_h * Bp, nullptr)) ||
        (rc = dalloc(&p->d_imgI, (size_t)p->P * Bp, &acc)) ||
        (rc = dalloc(&p->d_img_tmp, (size_t)2 * B * p->P, &acc)) ||
        (rc = dalloc(&p->d_sino_tmp, (size_t)B * p->D * p->nT, &acc)) ||
        (rc = dalloc(&p->d_gpart, (size_t)std::max(8 * B, 16) * p->P, &acc)) ||
        (p->d_rec1 && (rc = dalloc(&p->d_img_slot, (size_t)p->Pd, &acc)))) {
        free_work(p);
        return rc;
    }
    // band-form SIR forward rows (batches): waves of nfb = TS + 1 - maxlen first samples each
    if (B > 1 && p->use_sir && p->d_srec && p->sir_maxlen <= kSirBandTS) {
        const int nfb = kSirBandTS + 1 - std::max(p->sir_maxlen, 1);
        const int nwave = (p->sp_n + nfb - 1) / nfb;
        if ((rc = dalloc(&p->d_sband, (size_t)p->D * nwave * (3 * kSirBandTS + 1) * Bp, &acc))) {
            free_work(p);
            return rc;
        }
        p->sband_nwave = nwave;
        p->sband_nfb = nfb;
    }
    // the two zero rows after the last detector's bins (pairs outside the bin range read them)
    MBREC_HIP(hipMemset(p->d_bins + (size_t)p->D * p->geom.NB * Bp, 0, sizeof(float) * 2 * Bp));
    p->bcap = B;
    p->work_bytes = need_h * Bp * sizeof(float);   // hpart size (reallocated when the window grows)
    return MBREC_OK;
}

void host_decode(int32_t q, const BinGeom &g, int &nn, int &m)
{
    // identical arithmetic to decode_tau on the device (exact float fraction, float compares)
    const int base = q >> g.F;
    const float fr = (float)(q & ((1 << g.F) - 1)) * std::ldexp(1.0f, -g.F);
    m = (fr >= g.phi1) + (fr >= g.phi2);
    nn = base + g.K;
}

// Pull records of the forward operator without SIR: per detector, the pairs
// whose sample nn lies in [0, nT + 2K) sorted by bin j = 3 nn + m (device tile
// order inside a bin), each stored as (pixel, interpolation weight u), with CSR
// offsets [D][NB + 1] into the record array.  Bin j of detector d then receives
// (1 - u) c from its own pairs and u c from the pairs of bin j - 1.
static int build_tile_forward(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq);

static int build_pull_records(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq)
{
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const int NB = g.NB;
    std::vector<int32_t> off((size_t)p->D * (NB + 1), 0);
    std::vector<int32_t> jbuf((size_t)p->D * p->Pd, -1);
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            if (slot_pix[sl] < 0) continue;
            const size_t at = (size_t)d * p->Pd + sl;
            int nn, m;
            host_decode(tq[at], g, nn, m);
            if ((unsigned)nn >= (unsigned)span) continue;
            const int j = 3 * nn + m;
            jbuf[at] = j;
            off[(size_t)d * (NB + 1) + j + 1]++;
        }
    size_t total = 0;
    for (int d = 0; d < p->D; ++d) {
        int32_t *o = &off[(size_t)d * (NB + 1)];
        if (total + (size_t)p->Pd > (size_t)INT32_MAX)
            return set_error(MBREC_ERR_ARG, "mbrec:size", "too many (pixel, detector) pairs for int32 offsets");
        o[0] = (int32_t)total;
        for (int j = 0; j < NB; ++j) o[j + 1] += o[j];
        total = o[NB];
    }
    std::vector<int2> rec(std::max<size_t>(total, 1)), rec1(rec.size());
    std::vector<int32_t> fill(off);
    const float scale = std::ldexp(1.0f, -g.F);
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            const size_t at = (size_t)d * p->Pd + sl;
            const int j = jbuf[at];
            if (j < 0) continue;
            // u exactly as decode_tau computes it on the device
            const int32_t q = tq[at];
            const float fr = (float)(q & ((1 << g.F) - 1)) * scale;
            const int m = j - 3 * (j / 3);
            const float ph = m == 0 ? 0.0f : (m == 1 ? g.phi1 : g.phi2);
            const float iw = m == 0 ? g.iw0 : (m == 1 ? g.iw1 : g.iw2);
            const float u = (fr - ph) * iw;
            int32_t ub;
            std::memcpy(&ub, &u, sizeof(ub));
            const size_t r = fill[(size_t)d * (NB + 1) + j]++;
            rec[r] = make_int2(slot_pix[sl], ub);
            rec1[r] = make_int2(sl, ub);
        }
    void *old[] = {p->d_rec, p->d_rec_off, p->d_rec_k, p->d_rec1};
    for (void *q : old)
        if (q) (void)hipFree(q);
    p->table_bytes -= p->rec_bytes;
    p->rec_bytes = 0;
    p->d_rec = p->d_rec1 = nullptr;
    p->d_rec_off = p->d_rec_k = nullptr;
    p->pull = 1;
    p->nbk = 0;
    p->nrec = total;
    p->lo_off = 0;
    p->hi_off = 1;
    p->WL = p->Wb + 1;
    rec.resize(total + kRecPad, make_int2(0, 0));   // zero pad: runs of records may read past the last bin
    rec1.resize(total + kRecPad, make_int2(0, 0));
    int rc;
    if ((rc = dalloc(&p->d_rec, rec.size(), &p->rec_bytes)) || (rc = dalloc(&p->d_rec_off, off.size(), &p->rec_bytes)) ||
        (rc = upload(p->d_rec, rec.data(), rec.size())) || (rc = upload(p->d_rec_off, off.data(), off.size())) ||
        (rc = dalloc(&p->d_rec1, rec1.size(), &p->rec_bytes)) || (rc = upload(p->d_rec1, rec1.data(), rec1.size())))
        return rc;
    p->table_bytes += p->rec_bytes;
    return build_tile_forward(p, slot_pix, tq);
}

// Tile forward tables for one sinogram (see the Plan fields).  Pairs as the pull records
// take them (nn in [0, nT + 2K), j = 3 nn + m, u exactly as decode_tau); blocks of 4 x 4
// device tiles in tile order; plans without a 2-D grid (Nx == 1) get none.
static int build_tile_forward(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq)
{
    void *old[] = {p->d_tfrec, p->d_tfroff, p->d_tfdesc, p->d_tfpart, p->d_tfsrec, p->d_tfsro};
    for (void *q : old)
        if (q) (void)hipFree(q);
    p->d_tfrec = nullptr;
    p->d_tfroff = nullptr;
    p->d_tfdesc = nullptr;
    p->d_tfpart = nullptr;
    p->d_tfsrec = nullptr;
    p->d_tfsro = nullptr;
    p->table_bytes -= p->tf_bytes;
    p->tf_bytes = 0;
    p->tf_nbz = p->tf_nbx = 0;
    if (p->Nx == 1 || p->TZ != 16) return MBREC_OK;
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const int ntz = (p->Nz + 15) / 16, ntx = (p->Nx + 15) / 16;
    const int nbz = (ntz + 3) / 4, nbx = (ntx + 3) / 4, nblk = nbz * nbx;
    const float scale = std::ldexp(1.0f, -g.F);
    std::vector<uint32_t> rec, srec, sro;
    std::vector<uint16_t> roff;
    std::vector<int4> desc((size_t)nblk * p->D + 1);
    std::vector<int> order;
    std::vector<int> js(4096);
    std::vector<uint32_t> rs(4096), srt(4096);
    std::vector<int> cnt;
    for (int bx = 0; bx < nbx; ++bx)
        for (int bz = 0; bz < nbz; ++bz) {
            const int T = bx * nbz + bz;
            for (int d = 0; d < p->D; ++d) {
                int n = 0, jmin = INT32_MAX, jmax = INT32_MIN;
                for (int lt = 0; lt < 16; ++lt) {
                    const int tx = 4 * bx + lt / 4, tz = 4 * bz + lt % 4;
                    if (tx >= ntx || tz >= ntz) continue;
                    const int tile = tx * ntz + tz;
                    for (int k = 0; k < 256; ++k) {
                        const int sl = tile * 256 + k;
                        if (slot_pix[sl] < 0) continue;
                        const int32_t q = tq[(size_t)d * p->Pd + sl];
                        int nn, m;
                        host_decode(q, g, nn, m);
                        if ((unsigned)nn >= (unsigned)span) continue;
                        const int j = 3 * nn + m;
                        const float fr = (float)(q & ((1 << g.F) - 1)) * scale;
                        const float ph = m == 0 ? 0.0f : (m == 1 ? g.phi1 : g.phi2);
                        const float iw = m == 0 ? g.iw0 : (m == 1 ? g.iw1 : g.iw2);
                        const float u = (fr - ph) * iw;
                        long long u20 = std::llrint((double)u * (double)(1 << 20));
                        u20 = std::min<long long>(std::max<long long>(u20, 0), (1 << 20) - 1);
                        js[n] = j;
                        rs[n] = (uint32_t)(lt * 256 + k) | ((uint32_t)u20 << 12);
                        jmin = std::min(jmin, j);
                        jmax = std::max(jmax, j);
                        ++n;
                    }
                }
                const size_t rstart = rec.size(), ostart = roff.size();
                if (rstart > (size_t)INT32_MAX || ostart > (size_t)INT32_MAX - 70000) return MBREC_OK;   // too large: no tile form
                const int nrow = n ? jmax - jmin + 1 : 0;
                if (nrow + 1 > kTfMaxRows) return MBREC_OK;   // windows too long for the kernel's LDS: no tile form
                desc[(size_t)T * p->D + d] = make_int4((int)rstart, (int)ostart, n ? jmin : 0, nrow);
                // counting sort by row; offsets relative to rstart (n <= 4096 fits uint16)
                cnt.assign(nrow + 1, 0);
                for (int i = 0; i < n; ++i) cnt[js[i] - jmin + 1]++;
                for (int r = 0; r < nrow; ++r) cnt[r + 1] += cnt[r];
                for (int r = 0; r <= nrow; ++r) roff.push_back((uint16_t)cnt[r]);
                for (int i = 0; i < n; ++i) srt[cnt[js[i] - jmin]++] = rs[i];
                rec.insert(rec.end(), srt.begin(), srt.begin() + n);
                // rows by decreasing record count (ties in row order); row r's records are
                // srt[roff[ostart + r] .. roff[ostart + r + 1])
                order.resize(nrow);
                for (int r = 0; r < nrow; ++r) order[r] = r;
                const uint16_t *ro = roff.data() + ostart;
                std::stable_sort(order.begin(), order.end(),
                                 [ro](int a, int b) { return ro[a + 1] - ro[a] > ro[b + 1] - ro[b]; });
                uint32_t at = 0;
                for (int k = 0; k < nrow; ++k) {
                    const int r = order[k];
                    sro.push_back(at | ((uint32_t)r << 16));
                    for (int i = ro[r]; i < ro[r + 1]; ++i) srec.push_back(srt[i]);
                    at += ro[r + 1] - ro[r];
                }
                sro.push_back((uint32_t)n);
            }
        }
    desc[(size_t)nblk * p->D] = make_int4((int)rec.size(), (int)roff.size(), 0, 0);
    rec.push_back(0u);
    roff.push_back(0);
    srec.push_back(0u);
    sro.push_back(0u);
    int rc;
    if ((rc = dalloc(&p->d_tfrec, rec.size(), &p->tf_bytes)) || (rc = dalloc(&p->d_tfroff, roff.size(), &p->tf_bytes)) ||
        (rc = dalloc(&p->d_tfdesc, desc.size(), &p->tf_bytes)) || (rc = dalloc(&p->d_tfpart, roff.size(), &p->tf_bytes)) ||
        (rc = dalloc(&p->d_tfsrec, srec.size(), &p->tf_bytes)) || (rc = dalloc(&p->d_tfsro, sro.size(), &p->tf_bytes)) ||
        (rc = upload(p->d_tfrec, rec.data(), rec.size())) || (rc = upload(p->d_tfroff, roff.data(), roff.size())) ||
        (rc = upload(p->d_tfdesc, desc.data(), desc.size())) || (rc = upload(p->d_tfsrec, srec.data(), srec.size())) ||
        (rc = upload(p->d_tfsro, sro.data(), sro.size())))
        return rc;
    p->table_bytes += p->tf_bytes;
    p->tf_nbz = nbz;
    p->tf_nbx = nbx;
    return MBREC_OK;
}

// Bucket the (pixel, travel time) pairs of every detector by bin j = 3 nn + m
// (bucket floor(j / Wb), clamped to the first / last bucket), keeping the device
// tile order inside a bucket.  Without SIR a pair contributes iff its sample
// nn lies in [0, nT + 2K); with SIR (sidx, win given) iff one of its taps
// nn + s - Ls/2, s in its window, does, and its SIR index is stored with it.
// The window margins lo_off / hi_off must already cover the tap displacements.
// SIR forward records in pull form (see the Plan fields): a pair's taps reach samples
// [nf, nf + len) with nf = nn + s0 - Ls/2 (bin-space sample index nn = n + K); pairs
// none of whose taps land in [0, nT + 2K) are left out.  u exactly as decode_tau.
// Records are grouped by (detector, phase m) and sorted by nf inside a group; the
// weights of window k sit between zero margins in their own table (d_sirw), and a
// record's base points at its window's first weight there.
static int build_sir_pull(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq,
                          const int32_t *sidx, const int32_t *win, int Ls, const float *sirf)
{
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const int half = Ls / 2;
    int maxlen = 0;
    for (int k = 0; k < p->Nsir_build; ++k) maxlen = std::max(maxlen, (int)win[2 * (size_t)k + 1]);
    // nf ranges over [-(maxlen - 1), span - 1] for pairs with a tap in range
    const int lo = -std::max(maxlen - 1, 0), n_nf = span - lo + 1;
    // weights table: window k at k * stride + kSirWinPad, kSirWinPad zeros before it and at least as many after
    const int stride = maxlen + 2 * kSirWinPad;
    // drop the previous tables first, so an early return below never leaves records of an
    // earlier SIR table in use
    if (p->d_srec) { (void)hipFree(p->d_srec); (void)hipFree(p->d_srec_off); (void)hipFree(p->d_sirw); }
    p->table_bytes -= p->srec_bytes;
    p->srec_bytes = 0;
    p->d_srec = nullptr;
    p->d_srec_off = nullptr;
    p->d_sirw = nullptr;
    p->sp_lo = 0;
    p->sp_n = 0;
    if ((size_t)p->Nsir_build * stride >= ((size_t)1 << 31) || maxlen >= (1 << 15) || p->P >= (1 << 28) ||
        span >= (1 << 15))
        return MBREC_OK;   // the packed record does not fit: the SIR forward keeps the bucketed form only
    std::vector<float> sw((size_t)p->Nsir_build * stride, 0.0f);
    for (int k = 0; k < p->Nsir_build; ++k)
        for (int i = 0; i < win[2 * (size_t)k + 1]; ++i)
            sw[(size_t)k * stride + kSirWinPad + i] = sirf[(size_t)k * Ls + win[2 * (size_t)k] + i];
    // CSR over (detector, phase m, nf): list (d, m) starts at off[(3 d + m) (n_nf + 1)]
    std::vector<int32_t> off((size_t)p->D * 3 * (n_nf + 1), 0);
    std::vector<int32_t> nfb((size_t)p->D * p->Pd, INT32_MIN);
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            if (slot_pix[sl] < 0) continue;
            const size_t at = (size_t)d * p->Pd + sl;
            const int k = sidx[at];
            const int s0 = win[2 * (size_t)k], len = win[2 * (size_t)k + 1];
            int nn, m;
            host_decode(tq[at], g, nn, m);
            if (tq[at] == INT32_MIN || len <= 0) continue;
            const long long nf = (long long)nn + s0 - half;
            if (nf + len - 1 < 0 || nf >= span) continue;   // no tap in range
            nfb[at] = (int32_t)nf;
            off[((size_t)d * 3 + m) * (n_nf + 1) + (nf - lo) + 1]++;
        }
    size_t total = 0;
    for (int dm = 0; dm < 3 * p->D; ++dm) {
        int32_t *o = &off[(size_t)dm * (n_nf + 1)];
        if (total + (size_t)p->Pd > (size_t)INT32_MAX) return MBREC_OK;
        o[0] = (int32_t)total;
        for (int i = 0; i < n_nf; ++i) o[i + 1] += o[i];
        total = o[n_nf];
    }
    std::vector<int4> rec(total + kRecPad, make_int4(0, 0, 0, 0));
    std::vector<int32_t> fill(off);
    const float scale = std::ldexp(1.0f, -g.F);
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            const size_t at = (size_t)d * p->Pd + sl;
            if (nfb[at] == INT32_MIN) continue;
            const int k = sidx[at];
            const int len = win[2 * (size_t)k + 1];   // nfb[at] already holds the window's first sample
            int nn, m;
            host_decode(tq[at], g, nn, m);
            const float fr = (float)(tq[at] & ((1 << g.F) - 1)) * scale;
            const float u = m == 0 ? fr * g.iw0 : (m == 1 ? (fr - g.phi1) * g.iw1 : (fr - g.phi2) * g.iw2);
            int32_t ub;
            std::memcpy(&ub, &u, sizeof(ub));
            const int base = k * stride + kSirWinPad;
            rec[fill[((size_t)d * 3 + m) * (n_nf + 1) + (nfb[at] - lo)]++] =
                make_int4(slot_pix[sl] | (m << 28), (nfb[at] & 0xffff) | (len << 16), ub, base);
        }
    int rc;
    if ((rc = dalloc(&p->d_srec, rec.size(), &p->srec_bytes)) ||
        (rc = dalloc(&p->d_srec_off, off.size(), &p->srec_bytes)) ||
        (rc = dalloc(&p->d_sirw, sw.size(), &p->srec_bytes)) || (rc = upload(p->d_srec, rec.data(), rec.size())) ||
        (rc = upload(p->d_srec_off, off.data(), off.size())) || (rc = upload(p->d_sirw, sw.data(), sw.size())))
        return rc;
    p->table_bytes += p->srec_bytes;
    p->sp_lo = lo;
    p->sp_n = n_nf;
    return MBREC_OK;
}

// SIR transpose tables for the LDS tile form (d_stwin, d_stab): for pair (slot, d) with
// decoded (nn, m, u) and window [s0, s0 + len) of its SIR index, the taps s whose sample
// nn + s - Ls/2 lies in [0, nT + 2K) form [s_lo, s_hi); they read bin rows
// 3 (nn + s - Ls/2) + m and the next.  The tile's window starts at the lowest such row
// of its 256 pairs.  Weights come from d_sirw (layout of build_sir_pull).
static int build_sir_tiles(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq,
                           const int32_t *sidx, const int32_t *win, int Ls)
{
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const int half = Ls / 2;
    const int ntile = p->Pd / 256;
    int maxlen = 0;
    for (int k = 0; k < p->Nsir_build; ++k) maxlen = std::max(maxlen, (int)win[2 * (size_t)k + 1]);
    const int stride = maxlen + 2 * kSirWinPad;
    if (p->d_stab) { (void)hipFree(p->d_stab); (void)hipFree(p->d_stwin); }
    p->table_bytes -= p->stab_bytes;
    p->stab_bytes = 0;
    p->d_stab = nullptr;
    p->d_stwin = nullptr;
    p->stile_zr = 0;
    if (!p->d_sirw || maxlen > 32) return MBREC_OK;   // the kernel's weight registers hold 32 taps per pair
    std::vector<int32_t> w((size_t)ntile * p->D, 0);
    std::vector<int4> tab((size_t)ntile * p->D * 256, make_int4(0, 0, 0, 0));
    const float scale = std::ldexp(1.0f, -g.F);
    int wmax = 0;
    for (int t = 0; t < ntile; ++t)
        for (int d = 0; d < p->D; ++d) {
            int lo = INT32_MAX, hi = INT32_MIN;
            int4 *e = tab.data() + ((size_t)t * p->D + d) * 256;
            for (int k = 0; k < 256; ++k) {
                const size_t at = (size_t)d * p->Pd + (size_t)t * 256 + k;
                if (slot_pix[(size_t)t * 256 + k] < 0) continue;
                const int sk = sidx[at];
                const int s0 = win[2 * (size_t)sk], len = win[2 * (size_t)sk + 1];
                if (tq[at] == INT32_MIN || len <= 0) continue;
                int nn, m;
                host_decode(tq[at], g, nn, m);
                const long long sl = std::max<long long>(s0, (long long)half - nn);
                const long long sh = std::min<long long>((long long)s0 + len, (long long)span + half - nn);
                if (sl >= sh) continue;
                const int jl = (int)(3 * (nn + sl - half) + m), jh = (int)(3 * (nn + sh - 1 - half) + m + 1);
                lo = std::min(lo, jl);
                hi = std::max(hi, jh);
                const float fr = (float)(tq[at] & ((1 << g.F) - 1)) * scale;
                const float u = m == 0 ? fr * g.iw0 : (m == 1 ? (fr - g.phi1) * g.iw1 : (fr - g.phi2) * g.iw2);
                int32_t ub;
                std::memcpy(&ub, &u, sizeof(ub));
                e[k] = make_int4(jl, ub, (int)((size_t)sk * stride + kSirWinPad + (sl - s0)), (int)(sh - sl));
            }
            if (lo > hi) continue;
            w[(size_t)t * p->D + d] = lo;
            wmax = std::max(wmax, hi - lo + 1);
            for (int k = 0; k < 256; ++k)
                if (e[k].w > 0) e[k].x -= lo;
        }
    const int zr = 64 * ((wmax + 63) / 64);
    p->stile_wmax = wmax;
    if (zr > 256 || (size_t)p->Nsir_build * stride >= ((size_t)1 << 31)) return MBREC_OK;
    int rc;
    if ((rc = dalloc(&p->d_stwin, w.size(), &p->stab_bytes)) || (rc = dalloc(&p->d_stab, tab.size(), &p->stab_bytes)) ||
        (rc = upload(p->d_stwin, w.data(), w.size())) || (rc = upload(p->d_stab, tab.data(), tab.size())))
        return rc;
    p->table_bytes += p->stab_bytes;
    p->stile_zr = zr;
    return MBREC_OK;
}

int build_records(Plan *p, const std::vector<int32_t> &slot_pix, const std::vector<int32_t> &tq,
                  const int32_t *sidx, const int32_t *win, int Ls)
{
    if (!sidx) return build_pull_records(p, slot_pix, tq);
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const int half = Ls / 2;
    const int nbk = (g.NB + p->Wb - 1) / p->Wb;
    std::vector<int32_t> off((size_t)p->D * (nbk + 1), 0);
    std::vector<int32_t> kbuf((size_t)p->D * p->Pd, -1);
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            if (slot_pix[sl] < 0) continue;
            const size_t at = (size_t)d * p->Pd + sl;
            int nn, m;
            host_decode(tq[at], g, nn, m);
            if (sidx) {
                const int k = sidx[at];
                const long long s0 = win[2 * k], len = win[2 * k + 1];
                if (len <= 0 || (long long)nn + s0 - half >= span || (long long)nn + s0 + len - 1 - half < 0)
                    continue;
            } else if ((unsigned)nn >= (unsigned)span) {
                continue;
            }
            const long long j = 3LL * nn + m;
            const long long kb = j < 0 ? 0 : std::min<long long>(nbk - 1, j / p->Wb);
            kbuf[at] = (int32_t)kb;
            off[(size_t)d * (nbk + 1) + kb + 1]++;
        }
    size_t total = 0;
    for (int d = 0; d < p->D; ++d) {
        int32_t *o = &off[(size_t)d * (nbk + 1)];
        if (total + (size_t)p->Pd > (size_t)INT32_MAX)
            return set_error(MBREC_ERR_ARG, "mbrec:size", "too many (pixel, detector) pairs for int32 offsets");
        o[0] = (int32_t)total;
        for (int k = 0; k < nbk; ++k) o[k + 1] += o[k];
        total = o[nbk];
    }
    std::vector<int2> rec(std::max<size_t>(total, 1));
    std::vector<int32_t> rk(sidx ? std::max<size_t>(total, 1) : 0);
    std::vector<int32_t> fill(off);
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            const size_t at = (size_t)d * p->Pd + sl;
            if (kbuf[at] < 0) continue;
            const int32_t r = fill[(size_t)d * (nbk + 1) + kbuf[at]]++;
            rec[r] = make_int2(slot_pix[sl], tq[at]);
            if (sidx) rk[r] = sidx[at];
        }
    // replace the previous tables (the tile forward serves plans without SIR only)
    void *oldtf[] = {p->d_tfrec, p->d_tfroff, p->d_tfdesc, p->d_tfpart, p->d_tfsrec, p->d_tfsro};
    for (void *q : oldtf)
        if (q) (void)hipFree(q);
    p->d_tfrec = nullptr;
    p->d_tfroff = nullptr;
    p->d_tfdesc = nullptr;
    p->d_tfpart = nullptr;
    p->d_tfsrec = nullptr;
    p->d_tfsro = nullptr;
    p->table_bytes -= p->tf_bytes;
    p->tf_bytes = 0;
    p->tf_nbz = p->tf_nbx = 0;
    void *old[] = {p->d_rec, p->d_rec_off, p->d_rec_k, p->d_rec1};
    for (void *q : old)
        if (q) (void)hipFree(q);
    p->table_bytes -= p->rec_bytes;
    p->rec_bytes = 0;
    p->d_rec = p->d_rec1 = nullptr;
    p->d_rec_off = p->d_rec_k = nullptr;
    p->pull = 0;
    p->nbk = nbk;
    p->nrec = total;
    p->WL = p->Wb + p->hi_off - p->lo_off;
    int rc;
    if ((rc = dalloc(&p->d_rec, rec.size(), &p->rec_bytes)) || (rc = dalloc(&p->d_rec_off, off.size(), &p->rec_bytes)) ||
        (rc = upload(p->d_rec, rec.data(), rec.size())) || (rc = upload(p->d_rec_off, off.data(), off.size())))
        return rc;
    if (sidx && ((rc = dalloc(&p->d_rec_k, rk.size(), &p->rec_bytes)) || (rc = upload(p->d_rec_k, rk.data(), rk.size()))))
        return rc;
    p->table_bytes += p->rec_bytes;
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
// Device pixel order: 16 x 16 tiles of 256 slots (tile-major, z fastest), each
// made of four 8 x 8 sub-tiles of 64 slots, so every aligned run of 64 slots is
// an 8 x 8 square and every run of 256 a 16 x 16 square.  Single-column grids
// (Nx = 1) use runs of 256 consecutive pixels.  Slots beyond the grid map to -1.
static void make_pixel_order(Plan *p, std::vector<int32_t> &slot_pix)
{
    if (p->Nx == 1) {
        p->TZ = 256;
        p->TX = 1;
        p->Pd = (p->Nz + 255) / 256 * 256;
        slot_pix.assign(p->Pd, -1);
        for (int iz = 0; iz < p->Nz; ++iz) slot_pix[iz] = iz;
        return;
    }
    p->TZ = 16;
    p->TX = 16;
    const int ntz = (p->Nz + 15) / 16, ntx = (p->Nx + 15) / 16;
    p->Pd = ntz * ntx * 256;
    slot_pix.assign(p->Pd, -1);
    int slot = 0;
    for (int tx = 0; tx < ntx; ++tx)
        for (int tz = 0; tz < ntz; ++tz)
            for (int sx = 0; sx < 2; ++sx)
                for (int sz = 0; sz < 2; ++sz)
                    for (int lx = 0; lx < 8; ++lx)
                        for (int lz = 0; lz < 8; ++lz, ++slot) {
                            const int iz = tz * 16 + sz * 8 + lz, ix = tx * 16 + sx * 8 + lx;
                            if (iz < p->Nz && ix < p->Nx) slot_pix[slot] = iz + p->Nz * ix;
                        }
}

// Decoded pairs, slot-major: entry (slot, d) = (bin row d * NB + j, u) of the
// pair's first bin j, u exactly as decode_tau computes it; pairs outside the
// bin range point at the zero row D * NB (the bins buffer carries two zero rows
// after the last detector), so a gather reads them without a test.
static int build_pair_table(Plan *p, const std::vector<int32_t> &tq)
{
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const float scale = std::ldexp(1.0f, -g.F);
    std::vector<int2> t((size_t)p->Pd * p->D + kRecPad, make_int2(0, 0));   // zero pad for run prefetches
    for (int d = 0; d < p->D; ++d)
        for (int sl = 0; sl < p->Pd; ++sl) {
            const int32_t q = tq[(size_t)d * p->Pd + sl];
            int nn, m;
            host_decode(q, g, nn, m);
            int2 e = make_int2(p->D * g.NB, 0);
            if ((unsigned)nn < (unsigned)span) {
                const float fr = (float)(q & ((1 << g.F) - 1)) * scale;
                const float u = m == 0 ? fr * g.iw0 : (m == 1 ? (fr - g.phi1) * g.iw1 : (fr - g.phi2) * g.iw2);
                int32_t ub;
                std::memcpy(&ub, &u, sizeof(ub));
                e = make_int2(d * g.NB + 3 * nn + m, ub);
            }
            t[(size_t)sl * p->D + d] = e;
        }
    int rc;
    if ((rc = dalloc(&p->d_ptab, t.size(), &p->table_bytes)) || (rc = upload(p->d_ptab, t.data(), t.size())))
        return rc;
    return MBREC_OK;
}

// Bin window of every (256-slot tile, detector): first bin and length covering
// bins j, j + 1 of all its pairs (length 0: no pair in range); kTileWinMax bounds
// what the tile gather stages in LDS.
static int build_tile_windows(Plan *p, const std::vector<int32_t> &tq)
{
    const BinGeom &g = p->geom;
    const int span = g.nT + 2 * g.K;
    const int ntile = p->Pd / 256;
    std::vector<int2> w((size_t)ntile * p->D);
    int wmax = 0;
    for (int t = 0; t < ntile; ++t)
        for (int d = 0; d < p->D; ++d) {
            int lo = INT32_MAX, hi = INT32_MIN;
            const int32_t *q = tq.data() + (size_t)d * p->Pd + (size_t)t * 256;
            for (int k = 0; k < 256; ++k) {
                int nn, m;
                host_decode(q[k], g, nn, m);
                if ((unsigned)nn >= (unsigned)span) continue;
                const int j = 3 * nn + m;
                lo = std::min(lo, j);
                hi = std::max(hi, j + 1);
            }
            w[(size_t)t * p->D + d] = lo <= hi ? make_int2(lo, hi - lo + 1) : make_int2(0, 0);
            if (lo <= hi) wmax = std::max(wmax, hi - lo + 1);
        }
    p->tile_wmax = wmax;
    p->tile_zr = 64 * ((wmax + 63) / 64);   // staged windows are whole 64-row rounds
    // decoded pairs per (tile, detector): row inside the tile's window, or the
    // zero row tile_zr for pairs outside the bin range; u exactly as decode_tau
    const float scale = std::ldexp(1.0f, -g.F);
    std::vector<int2> tt((size_t)ntile * p->D * 256);
    for (int t = 0; t < ntile; ++t)
        for (int d = 0; d < p->D; ++d) {
            const int lo = w[(size_t)t * p->D + d].x;
            const int32_t *q = tq.data() + (size_t)d * p->Pd + (size_t)t * 256;
            int2 *e = tt.data() + ((size_t)t * p->D + d) * 256;
            for (int k = 0; k < 256; ++k) {
                int nn, m;
                host_decode(q[k], g, nn, m);
                e[k] = make_int2(p->tile_zr, 0);
                if ((unsigned)nn >= (unsigned)span) continue;
                const float fr = (float)(q[k] & ((1 << g.F) - 1)) * scale;
                const float u = m == 0 ? fr * g.iw0 : (m == 1 ? (fr - g.phi1) * g.iw1 : (fr - g.phi2) * g.iw2);
                int32_t ub;
                std::memcpy(&ub, &u, sizeof(ub));
                e[k] = make_int2(3 * nn + m - lo, ub);
            }
        }
    int rc;
    if ((rc = dalloc(&p->d_twin, w.size(), &p->table_bytes)) || (rc = upload(p->d_twin, w.data(), w.size())) ||
        (rc = dalloc(&p->d_ttab, tt.size(), &p->table_bytes)) || (rc = upload(p->d_ttab, tt.data(), tt.size())))
        return rc;
    // launch order of the batch tile gather: the tiles listed in serpentine strips of four
    // tile rows (z), the list cut into 8 contiguous runs, one per XCD (workgroup L runs on
    // XCD L % 8 as its (L / 8)-th); the workgroups an XCD holds at once then cover a compact
    // patch of neighbouring tiles, whose bin windows for one detector overlap in its L2
    if (p->Nx > 1 && p->TZ == 16) {
        const int ntz = (p->Nz + 15) / 16, ntx = (p->Nx + 15) / 16;
        std::vector<int32_t> seq;
        seq.reserve(ntile);
        for (int z0 = 0, strip = 0; z0 < ntz; z0 += 4, ++strip)
            for (int i = 0; i < ntx; ++i) {
                const int tx = (strip & 1) ? ntx - 1 - i : i;
                for (int tz = z0; tz < std::min(ntz, z0 + 4); ++tz) seq.push_back(tx * ntz + tz);
            }
        const int ntp = 8 * ((ntile + 7) / 8), m = ntp / 8;
        std::vector<int32_t> ord(ntp);
        for (int L = 0; L < ntp; ++L) {
            const int idx = (L % 8) * m + L / 8;
            ord[L] = idx < ntile ? seq[idx] : -1;
        }
        if ((rc = dalloc(&p->d_tord, ord.size(), &p->table_bytes)) ||
            (rc = upload(p->d_tord, ord.data(), ord.size())))
            return rc;
        p->tile_launch = ntp;
    }
    // the same pairs packed into 32 bits for the one-sinogram gather: row (< 512) | u23 << 9,
    // u rounded to the nearest 2^-23 (within half an FP32 ulp at u ~ 1)
    if (p->tile_zr <= 256) {
        std::vector<uint32_t> t1(tt.size());
        for (size_t i = 0; i < tt.size(); ++i) {
            float u;
            std::memcpy(&u, &tt[i].y, sizeof(u));
            long long u23 = std::llrint((double)u * (double)(1 << 23));
            u23 = std::min<long long>(std::max<long long>(u23, 0), (1 << 23) - 1);
            t1[i] = (uint32_t)tt[i].x | ((uint32_t)u23 << 9);
        }
        if ((rc = dalloc(&p->d_tq1, t1.size(), &p->table_bytes)) || (rc = upload(p->d_tq1, t1.data(), t1.size())))
            return rc;
    }
    return MBREC_OK;
}

static void free_plan(Plan *p)
{
    void *ptrs[] = {p->d_pix, p->d_tq, p->d_sidx, p->d_sir, p->d_swin, p->d_ir, p->d_irf, p->d_ir64, p->d_irf64,
                    p->d_kz,  p->d_kx, p->d_rec, p->d_rec_off, p->d_rec_k, p->d_rec1, p->d_twin, p->d_ptab, p->d_ttab, p->d_srec, p->d_srec_off, p->d_sirw, p->d_stwin, p->d_stab,
                    p->d_tfrec, p->d_tfroff, p->d_tfdesc, p->d_tfpart, p->d_tfsrec, p->d_tfsro, p->d_tq1, p->d_tord};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    free_work(p);
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
        (rc = upload(p->d_tq, tq.data(), tq.size())) || (rc = build_tile_windows(p, tq)) ||
        (rc = build_pair_table(p, tq)) ||
        (rc = build_records(p, slot_pix, tq, nullptr, nullptr, 0))) {
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
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
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
    // a second call replaces the tables: the workspace (band-form rows sized for the old
    // window length and sample range) and the launch choices timed on the old tables go
    if (p->bcap) {
        MBREC_HIP(hipDeviceSynchronize());
        free_work(p);
    }
    p->tune = TuneChoice();
    p->use_sir = 0;
    if (p->d_sidx) { (void)hipFree(p->d_sidx); (void)hipFree(p->d_sir); (void)hipFree(p->d_swin); }
    int rc;
    if ((rc = dalloc(&p->d_sidx, sidx.size(), &p->table_bytes)) ||
        (rc = dalloc(&p->d_sir, sirf.size(), &p->table_bytes)) ||
        (rc = dalloc(&p->d_swin, win.size(), &p->table_bytes)) ||
        (rc = upload(p->d_sidx, sidx.data(), sidx.size())) ||
        (rc = upload(p->d_sir, sirf.data(), sirf.size())) ||
        (rc = upload(p->d_swin, win.data(), win.size())))
        return rc;
    // deposit margins: a record at bin j reaches j + 3 (s - Ls/2) + {0, 1} for s in its window
    int smin = INT32_MAX, smax = INT32_MIN;
    for (int k = 0; k < Nsir; ++k)
        if (win[2 * (size_t)k + 1] > 0) {
            smin = std::min(smin, (int)win[2 * (size_t)k]);
            smax = std::max(smax, (int)(win[2 * (size_t)k] + win[2 * (size_t)k + 1] - 1));
        }
    if (smin > smax) smin = smax = Ls / 2;
    p->lo_off = std::min(0, 3 * (smin - Ls / 2));
    p->hi_off = std::max(1, 3 * (smax - Ls / 2) + 1);
    p->WL = p->Wb + p->hi_off - p->lo_off;
    {
        std::vector<int32_t> tq((size_t)p->D * p->Pd);
        MBREC_HIP(hipMemcpy(tq.data(), p->d_tq, sizeof(int32_t) * tq.size(), hipMemcpyDeviceToHost));
        if ((rc = build_records(p, slot_pix, tq, sidx.data(), win.data(), Ls))) return rc;
        p->Nsir_build = Nsir;
        if ((rc = build_sir_pull(p, slot_pix, tq, sidx.data(), win.data(), Ls, sirf.data()))) return rc;
        if ((rc = build_sir_tiles(p, slot_pix, tq, sidx.data(), win.data(), Ls))) return rc;
    }
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
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    Plan *p = &h->p;
    const size_t n = (size_t)Lir * (per_det ? p->D : 1);
    std::vector<float> f(n);
    // host layout [Lir x D] column-major = per detector contiguous
    for (size_t i = 0; i < n; ++i) f[i] = (float)ir[i];
    std::vector<float> ff(n);
    const int rows = per_det ? p->D : 1;
    for (int r = 0; r < rows; ++r)
        for (int i = 0; i < Lir; ++i) ff[(size_t)r * Lir + i] = f[(size_t)r * Lir + (Lir - 1 - i)];
    // FP64 copies zero-padded to whole groups of 16 taps (exact: the FP32 taps widened)
    const int Lp = (Lir + 15) & ~15;
    std::vector<double> d64((size_t)rows * Lp, 0.0), df64((size_t)rows * Lp, 0.0);
    for (int r = 0; r < rows; ++r)
        for (int i = 0; i < Lir; ++i) {
            d64[(size_t)r * Lp + i] = (double)f[(size_t)r * Lir + i];
            df64[(size_t)r * Lp + i] = (double)ff[(size_t)r * Lir + i];
        }
    if (p->d_ir) { (void)hipFree(p->d_ir); (void)hipFree(p->d_irf); (void)hipFree(p->d_ir64); (void)hipFree(p->d_irf64); }
    int rc;
    if ((rc = dalloc(&p->d_ir, n, &p->table_bytes)) || (rc = upload(p->d_ir, f.data(), n)) ||
        (rc = dalloc(&p->d_irf, n, &p->table_bytes)) || (rc = upload(p->d_irf, ff.data(), n)) ||
        (rc = dalloc(&p->d_ir64, d64.size(), &p->table_bytes)) || (rc = upload(p->d_ir64, d64.data(), d64.size())) ||
        (rc = dalloc(&p->d_irf64, df64.size(), &p->table_bytes)) ||
        (rc = upload(p->d_irf64, df64.data(), df64.size())))
        return rc;
    p->Lir = Lir;
    p->ir_per_det = per_det ? 1 : 0;
    return MBREC_OK;
}

int mbrec_plan_set_prefilter(mbrec_plan h, const double *kz, int Kz, const double *kx, int Kx)
{
    if (!h || !kz || !kx || Kz < 0 || Kx < 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_set_prefilter: invalid arguments");
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    Plan *p = &h->p;
    // K <= N/2: every offset wraps at most once (at K = N/2, N even, the offsets -K and +K
    // reach the same sample; the caller zeroes one of them)
    if (2 * Kz > p->Nz || 2 * Kx > p->Nx)
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
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    h->p.sigma = sigma;
    return MBREC_OK;
}

int mbrec_plan_destroy(mbrec_plan h)
{
    if (!h) return MBREC_OK;
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    (void)hipDeviceSynchronize();
    free_plan(&h->p);
    delete h;
    return MBREC_OK;
}

size_t mbrec_plan_device_bytes(mbrec_plan h)
{
    if (!h) return 0;
    const Plan *p = &h->p;
    if (!p->bcap) return p->table_bytes;
    const size_t Bp = batch_stride(p->bcap);
    return p->table_bytes + p->work_bytes +
           sizeof(float) * ((size_t)p->D * p->geom.NB * Bp + (size_t)p->P * Bp + (size_t)2 * p->bcap * p->P +
                            (size_t)p->bcap * p->D * p->nT);
}

int mbrec_plan_autotune(mbrec_plan h, int B, int reps, mbrec_stream s, int *choice)
{
    if (!h || B <= 0 || reps <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_autotune: invalid arguments");
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    int g = 0, bc = 0;
    int rc = autotune(&h->p, B, reps, (hipStream_t)s, &g, &bc);
    if (rc) return rc;
    if (choice) {
        choice[0] = g;
        choice[1] = bc;
    }
    return MBREC_OK;
}

int mbrec_plan_set_launch(mbrec_plan h, int B, int gather_variant, int scatter_choice)
{
    if (!h || B <= 0) return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_set_launch: invalid arguments");
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    return set_launch(&h->p, B, gather_variant, scatter_choice);
}

int mbrec_plan_set_fir_kernel(mbrec_plan h, int mfma)
{
    if (!h || (mfma != 0 && mfma != 1))
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_set_fir_kernel: invalid arguments");
    if (h->p.pins) return set_error(MBREC_ERR_ARG, "mbrec:planPinned", "plan is held by a batch graph");
    h->p.fir_mfma = mfma;
    return MBREC_OK;
}

int mbrec_num_gather_variants(void) { return num_gather_variants(); }

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

// model operators on sinograms cropped to per-sinogram windows of the plan's sample axis
static int window_op(mbrec_plan h, const float *in, float *out, int B, const int32_t *windows, hipStream_t s,
                     bool fwd)
{
    if (!h || !in || !out || !windows || B <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "invalid arguments");
    Plan *p = &h->p;
    std::vector<int2> w(B);
    for (int b = 0; b < B; ++b) {
        const long long f = windows[2 * b], c = windows[2 * b + 1];
        if (f < 0 || c <= 0 || f + c > p->nT)
            return set_error(MBREC_ERR_ARG, "mbrec:window", "window %d [%lld, %lld) outside [0, %d)", b, f, f + c,
                             p->nT);
        w[b] = make_int2((int)f, (int)(f + c));
    }
    int2 *dw = nullptr;
    MBREC_HIP(hipMalloc((void **)&dw, sizeof(int2) * B));
    hipError_t e = hipMemcpyAsync(dw, w.data(), sizeof(int2) * B, hipMemcpyHostToDevice, s);
    int rc = e == hipSuccess ? (fwd ? op_forward(p, in, out, B, true, s, dw) : op_transpose(p, in, out, B, true, s, dw))
                             : hip_check(e, "hipMemcpyAsync");
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(dw);
    if (rc) return rc;
    MBREC_HIP(e2);
    return MBREC_OK;
}

int mbrec_model_forward_window_dev(mbrec_plan h, const float *img, float *sino, int B, const int32_t *windows,
                                   mbrec_stream s)
{
    return window_op(h, img, sino, B, windows, (hipStream_t)s, true);
}

int mbrec_model_transpose_window_dev(mbrec_plan h, const float *sino, float *img, int B, const int32_t *windows,
                                     mbrec_stream s)
{
    return window_op(h, sino, img, B, windows, (hipStream_t)s, false);
}

int mbrec_plan_time_stage(mbrec_plan h, int stage, int B, int reps, mbrec_stream s, float *ms)
{
    if (!h || !ms || B <= 0 || reps <= 0)
        return set_error(MBREC_ERR_ARG, "mbrec:args", "plan_time_stage: invalid arguments");
    return time_stage(&h->p, stage, B, reps, (hipStream_t)s, ms);
}

}  // extern "C"
