// ps_device.hip -- device half of the C ABI: uploads a problem to HBM, owns the per-slot scratch
// and launches the gfx950 fit/eval kernels of ps_kernels.h.
//
// Data layout in HBM (FP64, structure of arrays):
//   points  t, ex,ey,ez (Earth), sx,sy,sz (Sun), alpha, y, sig, wq = w/sig^2   [npts = ndata + 3 conv rows]
//   facets  nx,ny,nz, darea [nfac]; dsph [nfac][dpitch] (zero-padded); pleg [nfac][npairs]; fc, fs
//   per slot (one per resident workgroup): J [npts][64] raw model rows, ytemp [npts],
//            dave [nlc][ma+1] relative-lightcurve column sums, alpha [2][mfit^2], beta [2][mfit]
// Reference citations are relative to the reference repository root.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/periodsearch_amd.h"
#include "ps_internal.h"
#include "ps_kernels.h"

using psk::DevProblem;
using psk::Slots;

namespace {
constexpr int kWG = psk::kWG;
constexpr int kMaxFit = psk::kMaxFit;
constexpr int kMaxMa = psk::kMaxMa;
constexpr int kMaxFac = psk::kMaxFac;
constexpr double kPi = M_PI;

template <class T>
T *dev_upload(const T *h, size_t n, std::vector<void *> &owned) {
    T *d = nullptr;
    if (hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    owned.push_back(d);
    if (n && hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

// Coefficient layout of the shape Jacobian: NT full 16-column MFMA tiles + R VALU columns.
struct Variant {
    int nt, r;
    bool lds, lds_mix;  // Dsph fits in LDS as FP64 / as FP32
};
Variant pick_variant(int nc, int nf) {
    Variant v;
    switch (nc) {
        case 16: v = {1, 0, true}; break;
        case 25: v = {1, 9, true}; break;
        case 36: v = {2, 4, true}; break;
        case 49: v = {3, 1, true}; break;
        default: v = {4, 0, true}; break;  // any nc <= 64, zero-padded to four tiles
    }
    v.lds = nf * (16 * v.nt + v.r) <= psk::kDsphLds;
    v.lds_mix = nf * (16 * v.nt + v.r) <= 2 * psk::kDsphLds;
    return v;
}
}  // namespace

struct ps_handle {
    int device = 0;
    int precision = PS_FP64;
    bool ls = false;  // the fit needs the Lambert column (cl fitted)
    int ndata = 0, npts = 0, nlc = 0, nfac = 0, ncoef = 0, ma = 0, mfit = 0, nmax = 0, slots = 0;
    double freq_start = 0, freq_end = 0, freq_step = 0, escl = 0, conw_r0 = 0;
    std::vector<double> wq_host;  // (1/sig^2)*w per point (convexity rows, w = 1, set per call)
    double conw_loaded = -1;
    DevProblem P{};
    const DevProblem *dP = nullptr;  // device copy of P: the kernels read its fields on demand
    Slots S{};
    Variant var{};
    double *d_wq = nullptr;
    std::vector<void *> owned;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double kernel_ms = 0;
    int launches = 0;
    // per-call staging
    double *d_freq = nullptr;
    int *d_pole = nullptr;
    ps_fit_result *d_out = nullptr;
    int cap_fits = 0;
};

namespace {
#define HIPCHECK(x)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return ps_fail(PS_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_));          \
    } while (0)

template <int NT, int R, bool LDSB, bool MIX, bool LS>
int launch_fit(ps_handle *h, int grid, int nfits, double conw_r) {
    hipLaunchKernelGGL((psk::fit_kernel<NT, R, LDSB, MIX, LS>), dim3(grid), dim3(kWG), 0, h->stream, h->dP, h->S, h->d_freq,
                       h->d_pole, nfits, conw_r, h->d_out);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
}
template <int NT, int R, bool LDSB, bool MIX>
int launch_eval(ps_handle *h, int grid, const double *a, int nvec, double *al, double *be, double *ch, double *ry,
                double *rd) {
    hipLaunchKernelGGL((psk::eval_kernel<NT, R, LDSB, MIX>), dim3(grid), dim3(kWG), 0, h->stream, h->dP, h->S, a, nvec,
                       al, be, ch, ry, rd);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
}

// Calls f.template operator()<NT, R, LDSB, MIX>() for the handle's kernel variant and precision.
template <class F>
int dispatch(const Variant &v, bool mix, F &&f) {
#define PS_VARIANT(NT_, R_)                                                                              \
    if (v.nt == NT_ && v.r == R_) {                                                                      \
        if (mix)                                                                                         \
            return v.lds_mix ? f.template operator()<NT_, R_, true, true>()                              \
                             : f.template operator()<NT_, R_, false, true>();                            \
        return v.lds ? f.template operator()<NT_, R_, true, false>() : f.template operator()<NT_, R_, false, false>(); \
    }
    PS_VARIANT(1, 0)
    PS_VARIANT(1, 9)
    PS_VARIANT(2, 4)
    PS_VARIANT(3, 1)
    PS_VARIANT(4, 0)
#undef PS_VARIANT
    return PS_ERR_UNSUPPORTED;
}

int set_conw(ps_handle *h, double conw_r) {
    if (h->conw_loaded == conw_r) return PS_OK;
    const double sig = 1 / conw_r;  // period_search_BOINC.cpp:721-726
    for (int j = 0; j < 3; ++j) h->wq_host[h->ndata + j] = 1 / (sig * sig);
    HIPCHECK(hipMemcpyAsync(h->d_wq + h->ndata, h->wq_host.data() + h->ndata, 3 * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
    h->conw_loaded = conw_r;
    return PS_OK;
}
}  // namespace

extern "C" int ps_prepare(const ps_input *in, const ps_geometry *geo, int device, ps_handle **out) {
    if (!in || !geo || !out) return ps_fail(PS_ERR_ARG, "ps_prepare: null argument");
    *out = nullptr;
    const int ncoef = geo->ncoef, ma = ncoef + 8;
    if (ncoef < 1) return ps_fail(PS_ERR_ARG, "ps_prepare: geometry has %d shape coefficients", ncoef);
    // a caller may fill ps_input itself: the lightcurve table must describe exactly the ndata points
    // the arrays hold (the kernels index points through it)
    if (in->ndata < 1 || in->lcurves < 1 || !in->lpoints || !in->inrel || !in->tim || !in->brightness || !in->sig ||
        !in->weight || !in->ee || !in->ee0)
        return ps_fail(PS_ERR_ARG, "ps_prepare: input has no points or a null array");
    {
        long long total = 0;
        for (int L = 0; L < in->lcurves; ++L) {
            if (in->lpoints[L] < 1) return ps_fail(PS_ERR_ARG, "ps_prepare: lightcurve %d has %d points", L, in->lpoints[L]);
            total += in->lpoints[L];
        }
        if (total != in->ndata)
            return ps_fail(PS_ERR_ARG, "ps_prepare: lightcurves hold %lld points, ndata is %d", total, in->ndata);
    }
    std::vector<int> ia(ma);
    for (int i = 0; i < ncoef; ++i) ia[i] = 1;
    if (in->onlyrel == 1) ia[0] = 0;  // period_search_BOINC.cpp:788-802
    ia[ncoef] = 1;
    ia[ncoef + 1] = 1;
    ia[ncoef + 2] = in->ia_prd;
    for (int i = 0; i < 3; ++i) ia[ncoef + 3 + i] = in->ia_par[i];
    ia[ncoef + 6] = in->ia_cl;
    ia[ncoef + 7] = 0;
    std::vector<int> fit_idx;
    for (int l = 0; l < ma; ++l)
        if (ia[l]) fit_idx.push_back(l);
    const int mfit = int(fit_idx.size());
    // J rows hold ma raw columns plus the residual in column kRes, one lane per column in phase C
    if (mfit > kMaxFit || ma > psk::kRes)
        return ps_fail(PS_ERR_UNSUPPORTED,
                       "GPU kernels support up to %d fitted of %d parameters (got %d of %d; Lmax <= 6)", kMaxFit,
                       psk::kRes, mfit, ma);
    if (geo->nfac > kMaxFac) return ps_fail(PS_ERR_UNSUPPORTED, "GPU kernels support nrows <= 8 (nfac <= %d)", kMaxFac);
    // the per-lightcurve normalisation table (nlc x (ma + 1) sums + two factors and a flag per curve)
    // lives in LDS
    if ((in->lcurves + 1) * (ma + 4) > psk::kDsphLds)
        return ps_fail(PS_ERR_UNSUPPORTED, "GPU kernels support up to %d lightcurves at ma = %d", psk::kDsphLds / (ma + 4) - 1,
                       ma);
    // the kernels take every column of a layout with leftover VALU columns (R > 0) as a coefficient
    {
        const Variant v = pick_variant(ncoef, geo->nfac);
        if ((v.r > 0 && 16 * v.nt + v.r != ncoef) || 16 * v.nt + v.r < ncoef)
            return ps_fail(PS_ERR_UNSUPPORTED, "no kernel layout for %d shape coefficients", ncoef);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ps_fail(PS_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return ps_fail(PS_ERR_ARG, "device %d out of range", device);
    HIPCHECK(hipSetDevice(device));

    auto *h = new ps_handle();
    h->device = device;
    h->ndata = in->ndata;
    h->npts = in->ndata + 3;
    h->nlc = in->lcurves + 1;
    h->nfac = geo->nfac;
    h->ncoef = ncoef;
    h->ma = ma;
    h->mfit = mfit;
    h->freq_start = 1 / in->per_start;  // :784-786
    h->freq_end = 1 / in->per_end;
    h->freq_step = 0.5 / (in->jd_max - in->jd_min) / 24 * in->per_step_coef;
    h->nmax = int((h->freq_start - h->freq_end) / h->freq_step) + 1;  // :807
    h->escl = in->escl;
    h->conw_r0 = in->conw / in->escl / in->escl;  // :485
    auto fail = [&](int rc) {
        ps_release(h);
        return rc;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
        return fail(ps_fail(PS_ERR_DEVICE, "stream/event creation failed"));

    const int npts = h->npts, nobs = in->ndata, nf = geo->nfac;
    // points (SoA); the convexity rows have y = 0, weight 1, t/e/e0 unused (:455-456, :721-726)
    std::vector<double> t(npts, 0), ex(npts, 0), ey(npts, 0), ez(npts, 0), sx(npts, 0), sy(npts, 0), sz(npts, 0),
        alpha(npts, 0), y(npts, 0), w(npts, 1), sig(npts, 1);
    h->wq_host.assign(npts, 0);
    for (int p = 0; p < nobs; ++p) {
        t[p] = in->tim[p];
        ex[p] = in->ee[3 * p];
        ey[p] = in->ee[3 * p + 1];
        ez[p] = in->ee[3 * p + 2];
        sx[p] = in->ee0[3 * p];
        sy[p] = in->ee0[3 * p + 1];
        sz[p] = in->ee0[3 * p + 2];
        // phase angle: a constant of the point, from the ecliptic-frame directions (bright.c:30-31,
        // dot_product.c:14), with the host libm's acos
        alpha[p] = std::acos(ex[p] * sx[p] + ey[p] * sy[p] + ez[p] * sz[p]);
        y[p] = in->brightness[p];
        w[p] = in->weight[p];
        sig[p] = in->sig[p];
        // the weight of the point in alpha, beta and chi2: (1/sig^2)*w, the product mrqcof forms per
        // evaluation (mrqcof.c:112-113, 121, 134), formed once
        h->wq_host[p] = (1 / (in->sig[p] * in->sig[p])) * w[p];
    }
    std::vector<int> lc_start(h->nlc + 1), lc_inrel(h->nlc), lc_of(npts);
    lc_start[0] = 0;
    for (int L = 0; L < in->lcurves; ++L) {
        lc_start[L + 1] = lc_start[L] + in->lpoints[L];
        lc_inrel[L] = in->inrel[L];
    }
    lc_start[h->nlc] = npts;
    lc_inrel[h->nlc - 1] = 0;
    for (int L = 0; L < h->nlc; ++L)
        for (int p = lc_start[L]; p < lc_start[L + 1]; ++p) lc_of[p] = L;
    // runs of one lightcurve within the 16-point tiles of the model pass (segments); two guard entries
    // at the end of seg_first cover the tile index one past the last
    const int ntiles = (nobs + 15) / 16;
    std::vector<int> seg_first(ntiles + 2), seg_lc, seg_range, lc_seg(h->nlc + 1);
    for (int tl = 0; tl < ntiles; ++tl) {
        seg_first[tl] = int(seg_lc.size());
        const int p0 = 16 * tl, p1 = std::min(nobs, p0 + 16);
        for (int p = p0; p < p1; ++p)
            if (p == p0 || lc_of[p] != lc_of[p - 1]) {
                if (p != p0) seg_range.back() |= (p - p0) << 8;
                seg_lc.push_back(lc_of[p]);
                seg_range.push_back(p - p0);
            }
        seg_range.back() |= (p1 - p0) << 8;
    }
    const int nseg = int(seg_lc.size());
    seg_first[ntiles] = seg_first[ntiles + 1] = nseg;
    for (int L = 0, s = 0; L <= h->nlc; ++L) {
        while (s < nseg && seg_lc[s] < L) ++s;
        lc_seg[L] = s;
    }
    std::vector<double> nx(nf), ny(nf), nz(nf);
    for (int i = 0; i < nf; ++i) {
        nx[i] = geo->nor[3 * i];
        ny[i] = geo->nor[3 * i + 1];
        nz[i] = geo->nor[3 * i + 2];
    }
    // initial poles (:197-206): beta measured from the pole, radians (:828-831)
    static const double bp[PS_N_POLES] = {0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
    static const double lp[PS_N_POLES] = {0, 90, 180, 270, 60, 180, 300, 60, 180, 300};
    double pb[PS_N_POLES], pl[PS_N_POLES];
    for (int m = 0; m < PS_N_POLES; ++m) {
        double b = 90 - bp[m];
        pb[m] = (kPi / 180) * b;
        pl[m] = (kPi / 180) * lp[m];
    }

    DevProblem &P = h->P;
    P.npts = npts;
    P.nobs = nobs;
    P.nlc = h->nlc;
    P.nf = nf;
    P.nc = ncoef;
    P.ma = ma;
    P.mfit = mfit;
    P.npairs = geo->npairs;
    P.mmax1 = geo->mmax + 1;
    P.ia_cl = ia[ncoef + 6];
    h->ls = P.ia_cl != 0;
    P.ia_ls = ia[ncoef + 7];
    // LM stop rule (:847-860)
    P.n_iter_max = 0;
    P.iter_diff_max = -1;
    P.n_iter_min = in->n_iter_min;
    if (in->stop_condition > 1) {
        P.n_iter_max = int(in->stop_condition);
        P.iter_diff_max = 0;
        P.n_iter_min = 0;
    }
    if (in->stop_condition < 1) {
        P.n_iter_max = 100;  // MAX_N_ITER, constants.h:6
        P.iter_diff_max = in->stop_condition;
    }
    P.phi0 = in->phi0;
    P.alamda_start = in->alamda_start;
    P.alamda_incr = in->alamda_incr;
    P.par1 = in->par[0];
    P.par2 = in->par[1];
    P.par3 = in->par[2];
    P.log_cl = std::log(in->cl);
    auto &o = h->owned;
    P.t = dev_upload(t.data(), npts, o);
    P.ex = dev_upload(ex.data(), npts, o);
    P.ey = dev_upload(ey.data(), npts, o);
    P.ez = dev_upload(ez.data(), npts, o);
    P.sx = dev_upload(sx.data(), npts, o);
    P.sy = dev_upload(sy.data(), npts, o);
    P.sz = dev_upload(sz.data(), npts, o);
    P.alpha = dev_upload(alpha.data(), npts, o);
    P.y = dev_upload(y.data(), npts, o);
    P.sig = dev_upload(sig.data(), npts, o);
    h->d_wq = dev_upload(h->wq_host.data(), npts, o);
    P.wq = h->d_wq;
    P.lc_start = dev_upload(lc_start.data(), lc_start.size(), o);
    P.lc_inrel = dev_upload(lc_inrel.data(), lc_inrel.size(), o);
    P.lc_of = dev_upload(lc_of.data(), lc_of.size(), o);
    P.nseg = nseg;
    P.seg_first = dev_upload(seg_first.data(), seg_first.size(), o);
    P.seg_lc = dev_upload(seg_lc.data(), seg_lc.size(), o);
    P.seg_range = dev_upload(seg_range.data(), seg_range.size(), o);
    P.lc_seg = dev_upload(lc_seg.data(), lc_seg.size(), o);
    P.nx = dev_upload(nx.data(), nf, o);
    P.ny = dev_upload(ny.data(), nf, o);
    P.nz = dev_upload(nz.data(), nf, o);
    P.darea = dev_upload(geo->darea, nf, o);
    h->var = pick_variant(ncoef, nf);
    P.dpitch = 16 * h->var.nt + h->var.r;
    std::vector<double> dsph_pad(size_t(nf) * P.dpitch, 0.0);
    for (int i = 0; i < nf; ++i)
        for (int k = 0; k < ncoef; ++k) dsph_pad[size_t(i) * P.dpitch + k] = geo->dsph[size_t(i) * ncoef + k];
    P.dsph = dev_upload(dsph_pad.data(), dsph_pad.size(), o);
    std::vector<float> dsph_f(dsph_pad.begin(), dsph_pad.end());
    P.dsphf = dev_upload(dsph_f.data(), dsph_f.size(), o);
    P.pleg = dev_upload(geo->pleg, size_t(nf) * geo->npairs, o);
    P.fc = dev_upload(geo->fc, size_t(nf) * (geo->mmax + 1), o);
    P.fs = dev_upload(geo->fs, size_t(nf) * (geo->mmax + 1), o);
    P.pair_m = dev_upload(geo->pair_m, geo->npairs, o);
    P.fit_idx = dev_upload(fit_idx.data(), fit_idx.size(), o);
    P.cg_first = dev_upload(geo->cg_first, ncoef, o);
    P.pole_beta = dev_upload(pb, PS_N_POLES, o);
    P.pole_lambda = dev_upload(pl, PS_N_POLES, o);
    for (void *ptr : o)
        if (!ptr) return fail(ps_fail(PS_ERR_DEVICE, "device upload failed"));
    h->dP = dev_upload(&P, 1, o);
    if (!h->dP) return fail(ps_fail(PS_ERR_DEVICE, "device upload failed"));

    // resident workgroups: one scratch slot each
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(ps_fail(PS_ERR_DEVICE, "device query failed"));
    int per_cu = 1;
    dispatch(h->var, false, [&]<int NT, int R, bool LDSB, bool MIX>() {
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, psk::fit_kernel<NT, R, LDSB, MIX, false>, kWG, 0);
        return 0;
    });
    h->slots = std::max(1, per_cu) * prop.multiProcessorCount;
    Slots &S = h->S;
    S.j_stride = size_t(npts) * psk::kMap;
    S.y_stride = size_t(npts);
    S.dave_stride = size_t(nseg) * psk::kMap;  // per-segment partial column sums
    S.a_stride = size_t(mfit) * mfit;
    S.b_stride = size_t(mfit);
    const size_t ns = size_t(h->slots);
    if (hipMalloc(&S.J, sizeof(double) * S.j_stride * ns) != hipSuccess ||
        hipMalloc(&S.ytemp, sizeof(double) * S.y_stride * ns) != hipSuccess ||
        hipMalloc(&S.dave, sizeof(double) * S.dave_stride * ns) != hipSuccess ||
        hipMalloc(&S.alpha, sizeof(double) * S.a_stride * 2 * ns) != hipSuccess ||
        hipMalloc(&S.beta, sizeof(double) * S.b_stride * 2 * ns) != hipSuccess)
        return fail(ps_fail(PS_ERR_NOMEM, "scratch allocation failed (%d slots)", h->slots));
    o.push_back(S.J);
    o.push_back(S.ytemp);
    o.push_back(S.dave);
    o.push_back(S.alpha);
    o.push_back(S.beta);
    *out = h;
    return PS_OK;
}

extern "C" int ps_set_precision(ps_handle *h, int precision) {
    if (!h) return ps_fail(PS_ERR_ARG, "ps_set_precision: null handle");
    if (precision != PS_FP64 && precision != PS_MIXED)
        return ps_fail(PS_ERR_ARG, "precision %d: use PS_FP64 or PS_MIXED", precision);
    h->precision = precision;
    return PS_OK;
}

extern "C" int ps_get_info(const ps_handle *h, ps_info *info) {
    if (!h || !info) return ps_fail(PS_ERR_ARG, "ps_get_info: null argument");
    info->ndata = h->ndata;
    info->npts = h->npts;
    info->nlc = h->nlc;
    info->nfac = h->nfac;
    info->ncoef = h->ncoef;
    info->ma = h->ma;
    info->mfit = h->mfit;
    info->nmax = h->nmax;
    info->freq_start = h->freq_start;
    info->freq_end = h->freq_end;
    info->freq_step = h->freq_step;
    info->escl = h->escl;
    info->conw_r0 = h->conw_r0;
    info->slots = h->slots;
    info->precision = h->precision;
    return PS_OK;
}

namespace {
// Uploads one batch of fits of a problem and launches the fit kernel on the problem's stream, without
// waiting: the results land in h->d_out.  Validation is the caller's.
int enqueue_fits(ps_handle *h, int nfits, const double *freq, const int *pole, double conw_r) {
    if (nfits > h->cap_fits) {
        (void)hipFree(h->d_freq);
        (void)hipFree(h->d_pole);
        (void)hipFree(h->d_out);
        h->d_freq = nullptr;
        h->d_pole = nullptr;
        h->d_out = nullptr;
        h->cap_fits = 0;
        HIPCHECK(hipMalloc(&h->d_freq, sizeof(double) * nfits));
        HIPCHECK(hipMalloc(&h->d_pole, sizeof(int) * nfits));
        HIPCHECK(hipMalloc(&h->d_out, sizeof(ps_fit_result) * nfits));
        h->cap_fits = nfits;
    }
    int rc = set_conw(h, conw_r);
    if (rc != PS_OK) return rc;
    HIPCHECK(hipMemcpyAsync(h->d_freq, freq, sizeof(double) * nfits, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(hipMemcpyAsync(h->d_pole, pole, sizeof(int) * nfits, hipMemcpyHostToDevice, h->stream));
    const int grid = std::min(nfits, h->slots);
    HIPCHECK(hipEventRecord(h->ev0, h->stream));
    rc = dispatch(h->var, h->precision == PS_MIXED, [&]<int NT, int R, bool LDSB, bool MIX>() {
        return h->ls ? launch_fit<NT, R, LDSB, MIX, true>(h, grid, nfits, conw_r)
                     : launch_fit<NT, R, LDSB, MIX, false>(h, grid, nfits, conw_r);
    });
    if (rc != PS_OK) return ps_fail(rc, "fit kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    HIPCHECK(hipEventRecord(h->ev1, h->stream));
    return PS_OK;
}

// Waits for the batch enqueued last on h and copies its nfits results to out.
int collect_fits(ps_handle *h, int nfits, ps_fit_result *out) {
    HIPCHECK(hipMemcpyAsync(out, h->d_out, sizeof(ps_fit_result) * nfits, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->kernel_ms += ms;
    h->launches += 1;
    return PS_OK;
}

int check_fits(const char *fn, int nfits, const double *freq, const int *pole) {
    for (int i = 0; i < nfits; ++i)
        if (pole[i] < 1 || pole[i] > PS_N_POLES || !(freq[i] > 0))
            return ps_fail(PS_ERR_ARG, "%s: fit %d has pole %d / freq %g", fn, i, pole[i], freq[i]);
    return PS_OK;
}
}  // namespace

extern "C" int ps_fit_batch(ps_handle *h, int nfits, const double *freq, const int *pole, double conw_r,
                            ps_fit_result *out) {
    if (!h || nfits < 0 || (nfits && (!freq || !pole || !out))) return ps_fail(PS_ERR_ARG, "ps_fit_batch: bad argument");
    if (nfits == 0) return PS_OK;
    int rc = check_fits("ps_fit_batch", nfits, freq, pole);
    if (rc != PS_OK) return rc;
    if (!(conw_r > 0)) return ps_fail(PS_ERR_ARG, "ps_fit_batch: conw_r must be positive");
    HIPCHECK(hipSetDevice(h->device));
    rc = enqueue_fits(h, nfits, freq, pole, conw_r);
    return rc != PS_OK ? rc : collect_fits(h, nfits, out);
}

extern "C" int ps_fit_multi(ps_handle *const *hs, int nh, const double *conw_r, int njobs, const int *prob,
                            const double *freq, const int *pole, ps_fit_result *out) {
    if (!hs || nh < 1 || !conw_r || njobs < 0 || (njobs && (!prob || !freq || !pole || !out)))
        return ps_fail(PS_ERR_ARG, "ps_fit_multi: bad argument");
    for (int i = 0; i < nh; ++i) {
        if (!hs[i]) return ps_fail(PS_ERR_ARG, "ps_fit_multi: problem %d is null", i);
        if (hs[i]->device != hs[0]->device) return ps_fail(PS_ERR_ARG, "ps_fit_multi: problems on different devices");
        if (!(conw_r[i] > 0)) return ps_fail(PS_ERR_ARG, "ps_fit_multi: conw_r[%d] must be positive", i);
        for (int k = 0; k < i; ++k)
            if (hs[k] == hs[i]) return ps_fail(PS_ERR_ARG, "ps_fit_multi: problem %d repeats problem %d", i, k);
    }
    for (int j = 0; j < njobs; ++j)
        if (prob[j] < 0 || prob[j] >= nh) return ps_fail(PS_ERR_ARG, "ps_fit_multi: job %d has problem %d", j, prob[j]);
    int rc = check_fits("ps_fit_multi", njobs, freq, pole);
    if (rc != PS_OK) return rc;
    if (njobs == 0) return PS_OK;
    HIPCHECK(hipSetDevice(hs[0]->device));
    // the jobs of each problem, and the problems in order of decreasing work (longest first)
    std::vector<std::vector<int>> jobs(nh);
    for (int j = 0; j < njobs; ++j) jobs[prob[j]].push_back(j);
    std::vector<int> order(nh);
    for (int i = 0; i < nh; ++i) order[i] = i;
    auto work = [&](int i) { return double(jobs[i].size()) * hs[i]->npts * hs[i]->nfac * hs[i]->ncoef; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    // every problem's batch goes onto its own stream before any result is waited for, so the launches
    // of different problems run concurrently on the device
    std::vector<std::vector<double>> f(nh);
    std::vector<std::vector<int>> pl(nh);
    for (int i : order) {
        if (jobs[i].empty()) continue;
        for (int j : jobs[i]) {
            f[i].push_back(freq[j]);
            pl[i].push_back(pole[j]);
        }
        rc = enqueue_fits(hs[i], int(jobs[i].size()), f[i].data(), pl[i].data(), conw_r[i]);
        if (rc != PS_OK) return rc;
    }
    std::vector<ps_fit_result> res;
    for (int i : order) {
        if (jobs[i].empty()) continue;
        res.resize(jobs[i].size());
        rc = collect_fits(hs[i], int(jobs[i].size()), res.data());
        if (rc != PS_OK) return rc;
        for (size_t k = 0; k < jobs[i].size(); ++k) out[jobs[i][k]] = res[k];
    }
    return PS_OK;
}

extern "C" int ps_eval(ps_handle *h, int nvec, const double *a, double conw_r, double *alpha, double *beta,
                       double *chisq, double *ymod, double *dyda) {
    if (!h || nvec < 0 || (nvec && !a)) return ps_fail(PS_ERR_ARG, "ps_eval: bad argument");
    if (nvec == 0) return PS_OK;
    if ((ymod == nullptr) != (dyda == nullptr)) return ps_fail(PS_ERR_ARG, "ps_eval: ymod and dyda go together");
    if (!(conw_r > 0)) return ps_fail(PS_ERR_ARG, "ps_eval: conw_r must be positive");
    HIPCHECK(hipSetDevice(h->device));
    int rc = set_conw(h, conw_r);
    if (rc != PS_OK) return rc;
    const size_t mf2 = size_t(h->mfit) * h->mfit;
    double *d_a = nullptr, *d_al = nullptr, *d_be = nullptr, *d_ch = nullptr, *d_y = nullptr, *d_dy = nullptr;
    auto cleanup = [&]() {
        hipFree(d_a);
        hipFree(d_al);
        hipFree(d_be);
        hipFree(d_ch);
        hipFree(d_y);
        hipFree(d_dy);
    };
    bool ok = hipMalloc(&d_a, sizeof(double) * nvec * h->ma) == hipSuccess &&
              (!alpha || hipMalloc(&d_al, sizeof(double) * nvec * mf2) == hipSuccess) &&
              (!beta || hipMalloc(&d_be, sizeof(double) * nvec * h->mfit) == hipSuccess) &&
              (!chisq || hipMalloc(&d_ch, sizeof(double) * nvec) == hipSuccess) &&
              (!ymod || hipMalloc(&d_y, sizeof(double) * nvec * h->npts) == hipSuccess) &&
              (!dyda || hipMalloc(&d_dy, sizeof(double) * nvec * h->npts * h->ma) == hipSuccess);
    if (!ok) {
        cleanup();
        return ps_fail(PS_ERR_NOMEM, "ps_eval: device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(d_a, a, sizeof(double) * nvec * h->ma, hipMemcpyHostToDevice, h->stream);
    const int grid = std::min(nvec, h->slots);
    if (e == hipSuccess)
        rc = dispatch(h->var, h->precision == PS_MIXED, [&]<int NT, int R, bool LDSB, bool MIX>() {
            return launch_eval<NT, R, LDSB, MIX>(h, grid, d_a, nvec, d_al, d_be, d_ch, d_y, d_dy);
        });
    if (e == hipSuccess && rc == PS_OK && alpha)
        e = hipMemcpyAsync(alpha, d_al, sizeof(double) * nvec * mf2, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && rc == PS_OK && beta)
        e = hipMemcpyAsync(beta, d_be, sizeof(double) * nvec * h->mfit, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && rc == PS_OK && chisq)
        e = hipMemcpyAsync(chisq, d_ch, sizeof(double) * nvec, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && rc == PS_OK && ymod)
        e = hipMemcpyAsync(ymod, d_y, sizeof(double) * nvec * h->npts, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && rc == PS_OK && dyda)
        e = hipMemcpyAsync(dyda, d_dy, sizeof(double) * nvec * h->npts * h->ma, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    cleanup();
    if (e != hipSuccess) return ps_fail(PS_ERR_DEVICE, "ps_eval: %s", hipGetErrorString(e));
    return rc;
}

extern "C" int ps_kernel_time(ps_handle *h, double *ms, int *launches, int reset) {
    if (!h) return ps_fail(PS_ERR_ARG, "ps_kernel_time: null handle");
    if (ms) *ms = h->kernel_ms;
    if (launches) *launches = h->launches;
    if (reset) {
        h->kernel_ms = 0;
        h->launches = 0;
    }
    return PS_OK;
}

extern "C" void ps_release(ps_handle *h) {
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void *p : h->owned) hipFree(p);
    hipFree(h->d_freq);
    hipFree(h->d_pole);
    hipFree(h->d_out);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}
