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
#include <array>
#include <map>
#include <tuple>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/periodsearch_amd.h"
#include "ps_internal.h"
#include "ps_kernels.h"
#include "ps_fit2.h"

using psk::DevProblem;
using psk::Slots;
using psk::Slots2;

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

// Pass schedule of the J-free kernels: the observed points and the three convexity rows (index
// nobs .. nobs+2) placed into passes of at most kPassRows positions (kPassTiles tiles of 16), every
// relative lightcurve whole inside one pass.  Relative lightcurves are packed first, in order; the
// calibrated points then fill each pass up to its last started tile, then whole tiles of the passes
// with room, then new passes; the convexity rows go to the end of the last pass.  perm[q] = the
// point at position q (-1: padding).  Returns false when a relative lightcurve is longer than a pass.
struct Schedule {
    std::vector<int> perm, pass_tile0, pass_unit0, unit_rows, unit_lc;
    int conv_q = 0;
};
bool build_schedule(const std::vector<int> &lc_start, const std::vector<int> &lc_inrel, int nlc_obs, Schedule &s) {
    constexpr int cap = psk::kPassRows;
    std::vector<std::vector<int>> pass;   // positions of each pass (point indices)
    std::vector<std::vector<int>> units;  // per pass: relative lightcurves
    for (int L = 0; L < nlc_obs; ++L) {
        if (!lc_inrel[L]) continue;
        const int n = lc_start[L + 1] - lc_start[L];
        if (n > cap) return false;
        if (pass.empty() || int(pass.back().size()) + n > cap) {
            pass.emplace_back();
            units.emplace_back();
        }
        units.back().push_back(L);
        for (int p = lc_start[L]; p < lc_start[L + 1]; ++p) pass.back().push_back(p);
    }
    std::vector<int> cal;
    for (int L = 0; L < nlc_obs; ++L)
        if (!lc_inrel[L])
            for (int p = lc_start[L]; p < lc_start[L + 1]; ++p) cal.push_back(p);
    size_t next = 0;
    auto room = [&](size_t k, int upto) { return std::max(0, upto - int(pass[k].size())); };
    for (size_t k = 0; k < pass.size() && next < cal.size(); ++k) {  // to the end of the started tile
        int n = room(k, (int(pass[k].size()) + 15) / 16 * 16);
        while (n-- > 0 && next < cal.size()) pass[k].push_back(cal[next++]);
    }
    for (size_t k = 0; k < pass.size() && next < cal.size(); ++k) {  // whole tiles
        int n = room(k, cap);
        while (n-- > 0 && next < cal.size()) pass[k].push_back(cal[next++]);
    }
    while (next < cal.size()) {
        pass.emplace_back();
        units.emplace_back();
        while (int(pass.back().size()) < cap && next < cal.size()) pass.back().push_back(cal[next++]);
    }
    const int nobs = lc_start[nlc_obs];
    if (pass.empty() || int(pass.back().size()) + 3 > cap) {
        pass.emplace_back();
        units.emplace_back();
    }
    for (int c = 0; c < 3; ++c) pass.back().push_back(nobs + c);
    s.perm.clear();
    s.pass_tile0.assign(1, 0);
    s.pass_unit0.assign(1, 0);
    s.unit_rows.clear();
    s.unit_lc.clear();
    for (size_t k = 0; k < pass.size(); ++k) {
        const int base = int(s.perm.size());
        // the relative lightcurves were placed first, in order, so each is a run from its first point
        for (int L : units[k]) {
            int r0 = -1;
            for (int i = 0; i < int(pass[k].size()); ++i)
                if (pass[k][i] == lc_start[L]) r0 = i;
            s.unit_rows.push_back(r0 | ((r0 + lc_start[L + 1] - lc_start[L]) << 8));
            s.unit_lc.push_back(L);
        }
        for (int p : pass[k]) {
            if (p == nobs) s.conv_q = int(s.perm.size());
            s.perm.push_back(p);
        }
        const int ntile = (int(pass[k].size()) + 15) / 16;
        s.perm.resize(size_t(base) + 16 * ntile, -1);
        s.pass_tile0.push_back(s.pass_tile0.back() + ntile);
        s.pass_unit0.push_back(int(s.unit_lc.size()));
    }
    return true;
}
}  // namespace

// The per-workgroup buffers of ps_fit_batch_noalloc: a file-scope device array, one copy per GPU (held
// until the process ends), 40 MiB = 512 workgroup slots of the largest supported shape (mfit 56, 512
// facets: 9968 doubles per slot).  Host code never uses the array's name as an address: each device's
// copy is found once, at ps_prepare, with hipGetSymbolAddress.
constexpr size_t kNoAllocDoubles = size_t(5) << 20;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
namespace {
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices];  // each device's copy of scratch (nullptr: not looked up yet)
std::mutex g_scratch_mu;

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return ps_fail(PS_ERR_DEVICE, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return PS_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return ps_fail(PS_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(p);
    return PS_OK;
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
    // the J-free kernels (ps_fit2.h): FP64 fits and evaluations when every relative lightcurve fits
    // one pass; P2 holds the points in pass-schedule order
    bool v2 = false;
    DevProblem P2{};
    const DevProblem *dP2 = nullptr;
    Slots2 S2{};
    int slots2 = 0;
    int nseg = 0;
    int kernel = PS_KERNEL_AUTO;  // kernel form of single-problem FP64 fits (ps_set_kernel / ps_autotune)
    struct EvalBuffers {          // ps_eval's device staging, kept between calls
        double *a = nullptr, *al = nullptr, *be = nullptr, *ch = nullptr, *y = nullptr, *dy = nullptr;
        size_t ca = 0, cal = 0, cbe = 0, cch = 0, cy = 0, cdy = 0;
    } ev;
    std::vector<double> wq2_host;
    double *d_wq2 = nullptr;
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

// Calls f.template operator()<NT, R>() for the coefficient layout of the J-free kernels.
template <class F>
int dispatch2(const Variant &v, F &&f) {
#define PS_VARIANT2(NT_, R_) \
    if (v.nt == NT_ && v.r == R_) return f.template operator()<NT_, R_>();
    PS_VARIANT2(1, 0)
    PS_VARIANT2(1, 9)
    PS_VARIANT2(2, 4)
    PS_VARIANT2(3, 1)
    PS_VARIANT2(4, 0)
#undef PS_VARIANT2
    return PS_ERR_UNSUPPORTED;
}

// The queue kernel's index of a coefficient layout (the order of its switch).
int layout_id(const Variant &v) {
    if (v.nt == 1 && v.r == 0) return 0;
    if (v.nt == 1 && v.r == 9) return 1;
    if (v.nt == 2 && v.r == 4) return 2;
    if (v.nt == 3 && v.r == 1) return 3;
    return 4;
}

template <int NT, int R, bool LS>
int launch_fit2(ps_handle *h, int grid, int nfits, double conw_r) {
    hipLaunchKernelGGL((psk::fit_kernel2<NT, R, LS, kMaxFac>), dim3(grid), dim3(psk::kWG2), 0, h->stream, h->dP2, h->S2,
                       h->d_freq, h->d_pole, nfits, conw_r, h->d_out);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
}

// The J-free kernels can run the FP64 fits and evaluations of a problem whose schedule they accept
// (batches and ps_fit_batch_noalloc use nothing else); single-problem fits use them unless the kernel
// form was set (or tuned) to the materialised-Jacobian kernels.
bool v2_ok(const ps_handle *h) { return h->v2 && h->precision == PS_FP64; }
bool use_v2(const ps_handle *h) { return v2_ok(h) && h->kernel != PS_KERNEL_LEGACY; }

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
    {
        const int rc = noalloc_setup(device);  // this device's copy of the no-allocation scratch
        if (rc != PS_OK) return rc;
    }

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
    // four zero facets past the end: the J-free facet loop loads the next chunk of Dsph unconditionally
    std::vector<double> dsph_pad(size_t(nf + 4) * P.dpitch, 0.0);
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

    // J-free kernels: the points in pass-schedule order (padding positions all zero, the arrays one tile
    // longer than the schedule for the second tile a wave reads past the end of an odd pass)
    // PSAMD_JFREE=0 in the environment keeps the materialised-Jacobian kernels (A/B comparisons)
    const char *jf = std::getenv("PSAMD_JFREE");
    const bool allow_v2 = !(jf && std::strcmp(jf, "0") == 0);
    Schedule sc;
    if (allow_v2 && mfit + 1 <= psk::kJP && nf <= kMaxFac && build_schedule(lc_start, lc_inrel, in->lcurves, sc)) {
        // the arrays run a whole pass past the schedule: a wave reads its second tile, and the per-row
        // loads of the last pass read 128 rows, whatever the pass holds
        const int nsched = int(sc.perm.size()), nq = nsched + psk::kPassRows;
        std::vector<double> qt(nq, 0), qex(nq, 0), qey(nq, 0), qez(nq, 0), qsx(nq, 0), qsy(nq, 0), qsz(nq, 0),
            qal(nq, 0), qy(nq, 0), qsig(nq, 1), qsl(nq, 0);
        std::vector<int> qrel(nq, 0), qperm(nq, -1);
        h->wq2_host.assign(nq, 0.0);
        for (int q = 0; q < nsched; ++q) {
            const int p = sc.perm[q];
            if (p < 0 || p >= nobs) continue;  // padding; the convexity rows keep zeros (their weight: set_conw)
            qt[q] = t[p];
            qex[q] = ex[p];
            qey[q] = ey[p];
            qez[q] = ez[p];
            qsx[q] = sx[p];
            qsy[q] = sy[p];
            qsz[q] = sz[p];
            qal[q] = alpha[p];
            qy[q] = y[p];
            qsig[q] = sig[p];
            h->wq2_host[q] = h->wq_host[p];
            qrel[q] = lc_inrel[lc_of[p]];
            qperm[q] = p;
            if (qrel[q]) qsl[q] = sig[p] * (lc_start[lc_of[p] + 1] - lc_start[lc_of[p]]);  // mrqcof.c:94
        }
        std::vector<int> fpos(ma, -1);
        for (int j = 0; j < mfit; ++j) fpos[fit_idx[j]] = j;
        DevProblem &Q = h->P2;
        Q = P;
        Q.t = dev_upload(qt.data(), nq, o);
        Q.ex = dev_upload(qex.data(), nq, o);
        Q.ey = dev_upload(qey.data(), nq, o);
        Q.ez = dev_upload(qez.data(), nq, o);
        Q.sx = dev_upload(qsx.data(), nq, o);
        Q.sy = dev_upload(qsy.data(), nq, o);
        Q.sz = dev_upload(qsz.data(), nq, o);
        Q.alpha = dev_upload(qal.data(), nq, o);
        Q.y = dev_upload(qy.data(), nq, o);
        Q.sig = dev_upload(qsig.data(), nq, o);
        h->d_wq2 = dev_upload(h->wq2_host.data(), nq, o);
        Q.wq = h->d_wq2;
        Q.npass = int(sc.pass_tile0.size()) - 1;
        Q.nsched = nsched;
        Q.conv_q = sc.conv_q;
        Q.c00col = fpos[0];
        Q.pass_tile0 = dev_upload(sc.pass_tile0.data(), sc.pass_tile0.size(), o);
        Q.pass_unit0 = dev_upload(sc.pass_unit0.data(), sc.pass_unit0.size(), o);
        Q.unit_rows = dev_upload(sc.unit_rows.data(), sc.unit_rows.size(), o);
        Q.unit_lc = dev_upload(sc.unit_lc.data(), sc.unit_lc.size(), o);
        Q.qrel = dev_upload(qrel.data(), nq, o);
        Q.qperm = dev_upload(qperm.data(), nq, o);
        Q.fpos = dev_upload(fpos.data(), fpos.size(), o);
        Q.qsl = dev_upload(qsl.data(), nq, o);
        for (void *ptr : o)
            if (!ptr) return fail(ps_fail(PS_ERR_DEVICE, "device upload failed"));
        h->dP2 = dev_upload(&Q, 1, o);
        if (!h->dP2) return fail(ps_fail(PS_ERR_DEVICE, "device upload failed"));
        h->v2 = true;
    }

    // resident workgroups: one scratch slot each; the slots are allocated on first use (a handle of a
    // multi-problem batch may never run a launch of its own)
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(ps_fail(PS_ERR_DEVICE, "device query failed"));
    int per_cu = 1;
    dispatch(h->var, false, [&]<int NT, int R, bool LDSB, bool MIX>() {
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, psk::fit_kernel<NT, R, LDSB, MIX, false>, kWG, 0);
        return 0;
    });
    h->slots = std::max(1, per_cu) * prop.multiProcessorCount;
    h->nseg = nseg;
    if (h->v2) {
        int per_cu2 = 1;
        dispatch2(h->var, [&]<int NT, int R>() {
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, psk::fit_kernel2<NT, R, false, kMaxFac>, psk::kWG2, 0);
            return 0;
        });
        h->slots2 = std::max(1, per_cu2) * prop.multiProcessorCount;
    }
    *out = h;
    return PS_OK;
}

namespace {
// Allocates the per-slot buffers of the kernels the handle is about to launch (once).
int ensure_scratch(ps_handle *h, bool v2) {
    auto alloc = [&](double **p, size_t n) {
        if (hipMalloc(p, sizeof(double) * std::max<size_t>(n, 1)) != hipSuccess) return false;
        h->owned.push_back(*p);
        return true;
    };
    if (v2 && !h->S2.alpha) {
        Slots2 T{};
        T.a_stride = size_t(h->mfit) * h->mfit;
        T.b_stride = size_t(h->mfit);
        T.f_stride = size_t(h->nfac);
        const size_t n = size_t(h->slots2);
        if (!alloc(&T.alpha, T.a_stride * 2 * n) || !alloc(&T.beta, T.b_stride * 2 * n) ||
            !alloc(&T.stash, psk::kStashPerSlot * n) || !alloc(&T.area, T.f_stride * n))
            return ps_fail(PS_ERR_NOMEM, "scratch allocation failed (%d slots)", h->slots2);
        h->S2 = T;
    }
    if (!v2 && !h->S.J) {
        Slots S{};
        S.j_stride = size_t(h->npts) * psk::kMap;
        S.y_stride = size_t(h->npts);
        S.dave_stride = size_t(h->nseg) * psk::kMap;  // per-segment partial column sums
        S.a_stride = size_t(h->mfit) * h->mfit;
        S.b_stride = size_t(h->mfit);
        const size_t n = size_t(h->slots);
        if (!alloc(&S.ytemp, S.y_stride * n) || !alloc(&S.dave, S.dave_stride * n) ||
            !alloc(&S.alpha, S.a_stride * 2 * n) || !alloc(&S.beta, S.b_stride * 2 * n) ||
            !alloc(&S.J, S.j_stride * n))
            return ps_fail(PS_ERR_NOMEM, "scratch allocation failed (%d slots)", h->slots);
        h->S = S;
    }
    return PS_OK;
}
}  // namespace

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
    info->slots = use_v2(h) ? h->slots2 : h->slots;
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
    if (rc == PS_OK) rc = ensure_scratch(h, use_v2(h));
    if (rc != PS_OK) return rc;
    HIPCHECK(hipMemcpyAsync(h->d_freq, freq, sizeof(double) * nfits, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(hipMemcpyAsync(h->d_pole, pole, sizeof(int) * nfits, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(hipEventRecord(h->ev0, h->stream));
    if (use_v2(h)) {
        const int grid = std::min(nfits, h->slots2);
        rc = dispatch2(h->var, [&]<int NT, int R>() {
            return h->ls ? launch_fit2<NT, R, true>(h, grid, nfits, conw_r) : launch_fit2<NT, R, false>(h, grid, nfits, conw_r);
        });
    } else {
        const int grid = std::min(nfits, h->slots);
        rc = dispatch(h->var, h->precision == PS_MIXED, [&]<int NT, int R, bool LDSB, bool MIX>() {
            return h->ls ? launch_fit<NT, R, LDSB, MIX, true>(h, grid, nfits, conw_r)
                         : launch_fit<NT, R, LDSB, MIX, false>(h, grid, nfits, conw_r);
        });
    }
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
    // FP64 problems of the J-free kernels: one persistent launch over a device work queue
    bool all_v2 = true;
    for (int i = 0; i < nh; ++i) all_v2 = all_v2 && v2_ok(hs[i]);
    if (all_v2) {
        ps_batch *b = nullptr;
        rc = ps_batch_create(hs, nh, &b);
        if (rc == PS_OK) rc = ps_batch_run(b, conw_r, njobs, prob, freq, pole, out);
        ps_batch_release(b);
        return rc;
    }
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

extern "C" int ps_fit_batch_noalloc(ps_handle *h, int nfits, const double *d_freq, const int *d_pole, double conw_r,
                                    ps_fit_result *d_out) {
    if (!h || nfits < 0 || (nfits && (!d_freq || !d_pole || !d_out)))
        return ps_fail(PS_ERR_ARG, "ps_fit_batch_noalloc: bad argument");
    if (nfits == 0) return PS_OK;
    if (!(conw_r > 0)) return ps_fail(PS_ERR_ARG, "ps_fit_batch_noalloc: conw_r must be positive");
    if (!v2_ok(h))
        return ps_fail(PS_ERR_UNSUPPORTED, "ps_fit_batch_noalloc: needs the J-free kernels (precision PS_FP64, every "
                                           "relative lightcurve <= %d points)", psk::kPassRows);
    double *base = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch[h->device] : nullptr;
    if (!base) return ps_fail(PS_ERR_DEVICE, "ps_fit_batch_noalloc: no scratch address for device %d", h->device);
    // the slots carved from the array, each region rounded to 256 bytes
    auto r32 = [](size_t n) { return (n + 31) & ~size_t(31); };
    Slots2 T{};
    T.a_stride = size_t(h->mfit) * h->mfit;
    T.b_stride = size_t(h->mfit);
    T.f_stride = size_t(h->nfac);
    const size_t per_slot = 2 * T.a_stride + 2 * T.b_stride + psk::kStashPerSlot + T.f_stride;
    size_t slots = std::min<size_t>(size_t(h->slots2), kNoAllocDoubles / (per_slot + 128));
    if (slots < 1)
        return ps_fail(PS_ERR_LIMIT, "ps_fit_batch_noalloc: one workgroup needs %zu bytes, the scratch holds %zu",
                       per_slot * sizeof(double), kNoAllocDoubles * sizeof(double));
    T.alpha = base;
    T.beta = T.alpha + r32(T.a_stride * 2 * slots);
    T.stash = T.beta + r32(T.b_stride * 2 * slots);
    T.area = T.stash + r32(psk::kStashPerSlot * slots);
    if (size_t(T.area + T.f_stride * slots - base) > kNoAllocDoubles)
        return ps_fail(PS_ERR_LIMIT, "ps_fit_batch_noalloc: %zu slots do not fit the scratch", slots);
    HIPCHECK(hipSetDevice(h->device));
    const int grid = int(std::min<size_t>(size_t(nfits), slots));
    const int rc = dispatch2(h->var, [&]<int NT, int R>() {
        if (h->ls)
            hipLaunchKernelGGL((psk::fit_kernel2<NT, R, true, kMaxFac>), dim3(grid), dim3(psk::kWG2), 0, h->stream,
                               h->dP2, T, d_freq, d_pole, nfits, conw_r, d_out);
        else
            hipLaunchKernelGGL((psk::fit_kernel2<NT, R, false, kMaxFac>), dim3(grid), dim3(psk::kWG2), 0, h->stream,
                               h->dP2, T, d_freq, d_pole, nfits, conw_r, d_out);
        return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
    });
    if (rc != PS_OK) return ps_fail(rc, "ps_fit_batch_noalloc: launch failed: %s", hipGetErrorString(hipGetLastError()));
    HIPCHECK(hipStreamSynchronize(h->stream));
    return PS_OK;
}

extern "C" int ps_set_kernel(ps_handle *h, int kernel) {
    if (!h) return ps_fail(PS_ERR_ARG, "ps_set_kernel: null handle");
    if (kernel != PS_KERNEL_AUTO && kernel != PS_KERNEL_LEGACY && kernel != PS_KERNEL_JFREE)
        return ps_fail(PS_ERR_ARG, "ps_set_kernel: unknown kernel form %d", kernel);
    if (kernel == PS_KERNEL_JFREE && !h->v2)
        return ps_fail(PS_ERR_UNSUPPORTED, "ps_set_kernel: the J-free kernels do not accept this problem (a relative "
                                           "lightcurve longer than %d points)", psk::kPassRows);
    h->kernel = kernel;
    return PS_OK;
}

namespace {
// Kernel forms chosen by ps_autotune, per problem shape, for the life of the process.
struct TuneKey {
    int ncoef, nfac, npts, mfit, ls, iters;
    bool operator<(const TuneKey &o) const {
        return std::tie(ncoef, nfac, npts, mfit, ls, iters) < std::tie(o.ncoef, o.nfac, o.npts, o.mfit, o.ls, o.iters);
    }
};
std::map<TuneKey, std::array<double, 2>> g_tuned;  // ms per fit of the two forms
std::mutex g_tuned_mu;
}  // namespace

extern "C" int ps_autotune(ps_handle *h, double conw_r, int *kernel, double *ms_per_fit) {
    if (!h || !(conw_r > 0)) return ps_fail(PS_ERR_ARG, "ps_autotune: bad argument");
    std::array<double, 2> ms{0, 0};
    int choice = PS_KERNEL_LEGACY;
    if (h->precision == PS_FP64 && h->v2) {
        const TuneKey key{h->ncoef, h->nfac, h->npts, h->mfit, int(h->ls), h->P.n_iter_max};
        bool known = false;
        {
            std::lock_guard<std::mutex> lk(g_tuned_mu);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) {
                ms = it->second;
                known = true;
            }
        }
        if (!known) {
            // a sample of the problem's own fits: the first trial periods x 10 poles, two launches of each
            // form at full occupancy of the J-free form (the second one timed)
            const int nfits = 2 * h->slots2;
            std::vector<double> freq(nfits);
            std::vector<int> pole(nfits);
            for (int i = 0; i < nfits; ++i) {
                freq[i] = h->freq_start - (i / PS_N_POLES) * h->freq_step;
                pole[i] = 1 + i % PS_N_POLES;
            }
            std::vector<ps_fit_result> out(nfits);
            const int saved = h->kernel;
            const double km = h->kernel_ms;
            const int kl = h->launches;
            for (int f = 0; f < 2; ++f) {
                h->kernel = f == 0 ? PS_KERNEL_LEGACY : PS_KERNEL_JFREE;
                for (int rep = 0; rep < 2; ++rep) {
                    const double before = h->kernel_ms;
                    const int rc = ps_fit_batch(h, nfits, freq.data(), pole.data(), conw_r, out.data());
                    if (rc != PS_OK) {
                        h->kernel = saved;
                        return rc;
                    }
                    ms[f] = (h->kernel_ms - before) / nfits;
                }
            }
            h->kernel = saved;
            h->kernel_ms = km;
            h->launches = kl;
            std::lock_guard<std::mutex> lk(g_tuned_mu);
            g_tuned[key] = ms;
        }
        choice = ms[1] <= ms[0] ? PS_KERNEL_JFREE : PS_KERNEL_LEGACY;
    }
    h->kernel = choice;
    if (kernel) *kernel = choice;
    if (ms_per_fit) {
        ms_per_fit[0] = ms[0];
        ms_per_fit[1] = ms[1];
    }
    return PS_OK;
}

extern "C" int ps_get_kernel(const ps_handle *h) {
    if (!h) return ps_fail(PS_ERR_ARG, "ps_get_kernel: null handle");
    return use_v2(h) ? PS_KERNEL_JFREE : PS_KERNEL_LEGACY;
}

// ---------------------------------------------------------------------------------------------
// Batches of many problems: one persistent launch draining a device work queue, captured in a graph.
// ---------------------------------------------------------------------------------------------
struct ps_batch {
    int device = 0;
    std::vector<ps_handle *> hs;
    std::vector<double> cost;  // estimated work of one fit of each problem
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void *> owned;
    const DevProblem **d_probs = nullptr;
    int *d_layout = nullptr;
    double *d_conw = nullptr;
    int *d_counter = nullptr;
    Slots2 pool{};
    bool mixed = false;  // mixed precision: queue_kernel_mixed over the legacy problem descriptions
    Slots lpool{};       // its per-slot buffers
    int slots = 0;
    psk::QueueJob *d_jobs = nullptr;
    ps_fit_result *d_out = nullptr;
    int cap_jobs = 0;
    hipGraphExec_t exec = nullptr;
    int exec_n0 = -1, exec_n1 = -1;
    double kernel_ms = 0;
    int launches = 0;
};

extern "C" void ps_batch_release(ps_batch *b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    for (void *p : b->owned) hipFree(p);
    hipFree(b->d_jobs);
    hipFree(b->d_out);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int ps_batch_create(ps_handle *const *hs, int nh, ps_batch **out) {
    if (!hs || nh < 1 || !out) return ps_fail(PS_ERR_ARG, "ps_batch_create: bad argument");
    *out = nullptr;
    const bool mixed = hs[0] && hs[0]->precision == PS_MIXED;
    for (int i = 0; i < nh; ++i) {
        if (!hs[i]) return ps_fail(PS_ERR_ARG, "ps_batch_create: problem %d is null", i);
        if (hs[i]->device != hs[0]->device) return ps_fail(PS_ERR_ARG, "ps_batch_create: problems on different devices");
        if ((hs[i]->precision == PS_MIXED) != mixed)
            return ps_fail(PS_ERR_ARG, "ps_batch_create: problems %d and 0 have different precisions", i);
        if (mixed) {
            if (layout_id(hs[i]->var) > 3 || !hs[i]->var.lds_mix)
                return ps_fail(PS_ERR_UNSUPPORTED, "ps_batch_create: problem %d: a mixed batch needs Lmax <= 6", i);
        } else if (!v2_ok(hs[i])) {
            return ps_fail(PS_ERR_UNSUPPORTED, "ps_batch_create: problem %d needs the FP64 J-free kernels (precision FP64, "
                                               "every relative lightcurve <= %d points)", i, psk::kPassRows);
        }
    }
    HIPCHECK(hipSetDevice(hs[0]->device));
    auto *b = new ps_batch();
    b->device = hs[0]->device;
    b->mixed = mixed;
    b->hs.assign(hs, hs + nh);
    auto fail = [&](int rc) {
        ps_batch_release(b);
        return rc;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return fail(ps_fail(PS_ERR_DEVICE, "stream/event creation failed"));
    std::vector<const DevProblem *> probs(nh);
    std::vector<int> layout(nh);
    size_t amax = 1, bmax = 1, fmax = 1, nptmax = 1, segmax = 1;
    b->cost.resize(nh);
    for (int i = 0; i < nh; ++i) {
        const ps_handle *h = hs[i];
        probs[i] = mixed ? h->dP : h->dP2;
        layout[i] = layout_id(h->var);
        amax = std::max(amax, size_t(h->mfit) * h->mfit);
        bmax = std::max(bmax, size_t(h->mfit));
        fmax = std::max(fmax, size_t(h->nfac));
        nptmax = std::max(nptmax, size_t(h->npts));
        segmax = std::max(segmax, size_t(h->nseg));
        const double iters = h->P.n_iter_max > 0 ? h->P.n_iter_max : h->P.n_iter_min;
        const double rows = mixed ? double(h->npts) : 16.0 * h->P2.pass_tile0[h->P2.npass];
        b->cost[i] = (iters + 1) * rows * (2.0 * h->nfac * h->ncoef + double(h->mfit) * (h->mfit + 1));
    }
    auto &o = b->owned;
    b->d_probs = dev_upload(probs.data(), nh, o);
    b->d_layout = dev_upload(layout.data(), nh, o);
    std::vector<double> zero(nh, 0.0);
    b->d_conw = dev_upload(zero.data(), nh, o);
    int z2[2] = {0, 0};
    b->d_counter = dev_upload(z2, 2, o);
    for (void *p : o)
        if (!p) return fail(ps_fail(PS_ERR_DEVICE, "device upload failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) return fail(ps_fail(PS_ERR_DEVICE, "device query failed"));
    int per_cu = 1;
    if (mixed)
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, psk::queue_kernel_mixed<false>, kWG, 0);
    else
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, psk::queue_kernel2<false>, psk::kWG2, 0);
    b->slots = std::max(1, per_cu) * prop.multiProcessorCount;
    if (mixed) {
        Slots &L = b->lpool;
        L.j_stride = nptmax * psk::kMap;
        L.y_stride = nptmax;
        L.dave_stride = segmax * psk::kMap;
        L.a_stride = amax;
        L.b_stride = bmax;
        const size_t n = size_t(b->slots);
        double *pj = nullptr, *py = nullptr, *pd = nullptr, *pa = nullptr, *pb = nullptr;
        if (hipMalloc(&pj, sizeof(double) * L.j_stride * n) != hipSuccess || (o.push_back(pj), false) ||
            hipMalloc(&py, sizeof(double) * L.y_stride * n) != hipSuccess || (o.push_back(py), false) ||
            hipMalloc(&pd, sizeof(double) * L.dave_stride * n) != hipSuccess || (o.push_back(pd), false) ||
            hipMalloc(&pa, sizeof(double) * amax * 2 * n) != hipSuccess || (o.push_back(pa), false) ||
            hipMalloc(&pb, sizeof(double) * bmax * 2 * n) != hipSuccess || (o.push_back(pb), false))
            return fail(ps_fail(PS_ERR_NOMEM, "batch scratch allocation failed (%d slots)", b->slots));
        L.J = pj;
        L.ytemp = py;
        L.dave = pd;
        L.alpha = pa;
        L.beta = pb;
        *out = b;
        return PS_OK;
    }
    Slots2 &T = b->pool;
    T.a_stride = amax;
    T.b_stride = bmax;
    T.f_stride = fmax;
    const size_t n = size_t(b->slots);
    double *pa = nullptr, *pb = nullptr, *ps = nullptr, *pf = nullptr;
    if (hipMalloc(&pa, sizeof(double) * amax * 2 * n) != hipSuccess || (o.push_back(pa), false) ||
        hipMalloc(&pb, sizeof(double) * bmax * 2 * n) != hipSuccess || (o.push_back(pb), false) ||
        hipMalloc(&ps, sizeof(double) * psk::kStashPerSlot * n) != hipSuccess || (o.push_back(ps), false) ||
        hipMalloc(&pf, sizeof(double) * fmax * n) != hipSuccess || (o.push_back(pf), false))
        return fail(ps_fail(PS_ERR_NOMEM, "batch scratch allocation failed (%d slots)", b->slots));
    T.alpha = pa;
    T.beta = pb;
    T.stash = ps;
    T.area = pf;
    *out = b;
    return PS_OK;
}

extern "C" int ps_batch_run(ps_batch *b, const double *conw_r, int njobs, const int *prob, const double *freq,
                            const int *pole, ps_fit_result *out) {
    if (!b || !conw_r || njobs < 0 || (njobs && (!prob || !freq || !pole || !out)))
        return ps_fail(PS_ERR_ARG, "ps_batch_run: bad argument");
    const int nh = int(b->hs.size());
    for (int i = 0; i < nh; ++i)
        if (!(conw_r[i] > 0)) return ps_fail(PS_ERR_ARG, "ps_batch_run: conw_r[%d] must be positive", i);
    for (int j = 0; j < njobs; ++j)
        if (prob[j] < 0 || prob[j] >= nh) return ps_fail(PS_ERR_ARG, "ps_batch_run: job %d has problem %d", j, prob[j]);
    int rc = check_fits("ps_batch_run", njobs, freq, pole);
    if (rc != PS_OK) return rc;
    if (njobs == 0) return PS_OK;
    HIPCHECK(hipSetDevice(b->device));
    if (b->mixed) {
        // the materialised-Jacobian kernels weight the convexity rows from each problem's weight array,
        // which holds the rows' weight for one conw_r at a time (the J-free kernels form it themselves)
        bool changed = false;
        for (int i = 0; i < nh; ++i) {
            changed |= b->hs[i]->conw_loaded != conw_r[i];
            rc = set_conw(b->hs[i], conw_r[i]);
            if (rc != PS_OK) return rc;
        }
        if (changed)
            for (int i = 0; i < nh; ++i) HIPCHECK(hipStreamSynchronize(b->hs[i]->stream));
    }
    // the job queue: problems without the Lambert column first, then those with it; each group longest first
    std::vector<int> order(njobs);
    for (int j = 0; j < njobs; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const bool lx = b->hs[prob[x]]->ls, ly = b->hs[prob[y]]->ls;
        if (lx != ly) return ly;
        return b->cost[prob[x]] > b->cost[prob[y]];
    });
    std::vector<psk::QueueJob> jobs(njobs);
    int n0 = 0;
    for (int k = 0; k < njobs; ++k) {
        const int j = order[k];
        jobs[k] = {freq[j], prob[j], pole[j], j, 0};
        n0 += !b->hs[prob[j]]->ls;
    }
    const int n1 = njobs - n0;
    if (njobs > b->cap_jobs) {
        if (b->exec) hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        (void)hipFree(b->d_jobs);
        (void)hipFree(b->d_out);
        b->d_jobs = nullptr;
        b->d_out = nullptr;
        b->cap_jobs = 0;
        HIPCHECK(hipMalloc(&b->d_jobs, sizeof(psk::QueueJob) * njobs));
        HIPCHECK(hipMalloc(&b->d_out, sizeof(ps_fit_result) * njobs));
        b->cap_jobs = njobs;
    }
    HIPCHECK(hipMemcpyAsync(b->d_conw, conw_r, sizeof(double) * nh, hipMemcpyHostToDevice, b->stream));
    HIPCHECK(hipMemcpyAsync(b->d_jobs, jobs.data(), sizeof(psk::QueueJob) * njobs, hipMemcpyHostToDevice, b->stream));
    if (!b->exec || b->exec_n0 != n0 || b->exec_n1 != n1) {
        if (b->exec) hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(psk::queue_reset, dim3(1), dim3(64), 0, b->stream, b->d_counter, 2);
        if (b->mixed) {
            if (n0)
                hipLaunchKernelGGL(psk::queue_kernel_mixed<false>, dim3(std::min(n0, b->slots)), dim3(kWG), 0, b->stream,
                                   b->d_probs, b->d_layout, b->d_conw, b->lpool, b->d_jobs, n0, b->d_counter, b->d_out);
            if (n1)
                hipLaunchKernelGGL(psk::queue_kernel_mixed<true>, dim3(std::min(n1, b->slots)), dim3(kWG), 0, b->stream,
                                   b->d_probs, b->d_layout, b->d_conw, b->lpool, b->d_jobs + n0, n1, b->d_counter + 1,
                                   b->d_out);
        } else {
            if (n0)
                hipLaunchKernelGGL(psk::queue_kernel2<false>, dim3(std::min(n0, b->slots)), dim3(psk::kWG2), 0, b->stream,
                                   b->d_probs, b->d_layout, b->d_conw, b->pool, b->d_jobs, n0, b->d_counter, b->d_out);
            if (n1)
                hipLaunchKernelGGL(psk::queue_kernel2<true>, dim3(std::min(n1, b->slots)), dim3(psk::kWG2), 0, b->stream,
                                   b->d_probs, b->d_layout, b->d_conw, b->pool, b->d_jobs + n0, n1, b->d_counter + 1,
                                   b->d_out);
        }
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        if (e != hipSuccess) return ps_fail(PS_ERR_DEVICE, "ps_batch_run: graph capture failed: %s", hipGetErrorString(e));
        b->exec_n0 = n0;
        b->exec_n1 = n1;
    }
    HIPCHECK(hipEventRecord(b->ev0, b->stream));
    HIPCHECK(hipGraphLaunch(b->exec, b->stream));
    HIPCHECK(hipEventRecord(b->ev1, b->stream));
    HIPCHECK(hipMemcpyAsync(out, b->d_out, sizeof(ps_fit_result) * njobs, hipMemcpyDeviceToHost, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->kernel_ms += ms;
    b->launches += 1;
    return PS_OK;
}

extern "C" int ps_batch_kernel_time(ps_batch *b, double *ms, int *launches, int reset) {
    if (!b) return ps_fail(PS_ERR_ARG, "ps_batch_kernel_time: null batch");
    if (ms) *ms = b->kernel_ms;
    if (launches) *launches = b->launches;
    if (reset) {
        b->kernel_ms = 0;
        b->launches = 0;
    }
    return PS_OK;
}

namespace {
// A device buffer of the handle that grows on demand (ps_eval's staging) and is freed by ps_release.
int grow(ps_handle *h, double *&p, size_t &cap, size_t n) {
    if (n <= cap) return PS_OK;
    if (p) {
        (void)hipFree(p);
        h->owned.erase(std::remove(h->owned.begin(), h->owned.end(), static_cast<void *>(p)), h->owned.end());
        p = nullptr;
        cap = 0;
    }
    if (hipMalloc(&p, sizeof(double) * n) != hipSuccess) {
        p = nullptr;
        return ps_fail(PS_ERR_NOMEM, "ps_eval: device allocation of %zu doubles failed", n);
    }
    h->owned.push_back(p);
    cap = n;
    return PS_OK;
}
}  // namespace

extern "C" int ps_eval(ps_handle *h, int nvec, const double *a, double conw_r, double *alpha, double *beta,
                       double *chisq, double *ymod, double *dyda) {
    if (!h || nvec < 0 || (nvec && !a)) return ps_fail(PS_ERR_ARG, "ps_eval: bad argument");
    if (nvec == 0) return PS_OK;
    if ((ymod == nullptr) != (dyda == nullptr)) return ps_fail(PS_ERR_ARG, "ps_eval: ymod and dyda go together");
    if (!(conw_r > 0)) return ps_fail(PS_ERR_ARG, "ps_eval: conw_r must be positive");
    HIPCHECK(hipSetDevice(h->device));
    int rc = set_conw(h, conw_r);
    if (rc == PS_OK) rc = ensure_scratch(h, use_v2(h));
    if (rc != PS_OK) return rc;
    const size_t mf2 = size_t(h->mfit) * h->mfit, nv = size_t(nvec);
    // staging buffers kept by the handle: allocated on first use, grown when a call needs more
    auto &E = h->ev;
    if ((rc = grow(h, E.a, E.ca, nv * h->ma)) != PS_OK || (alpha && (rc = grow(h, E.al, E.cal, nv * mf2)) != PS_OK) ||
        (beta && (rc = grow(h, E.be, E.cbe, nv * h->mfit)) != PS_OK) || (chisq && (rc = grow(h, E.ch, E.cch, nv)) != PS_OK) ||
        (ymod && (rc = grow(h, E.y, E.cy, nv * h->npts)) != PS_OK) ||
        (dyda && (rc = grow(h, E.dy, E.cdy, nv * h->npts * h->ma)) != PS_OK))
        return rc;
    double *d_al = alpha ? E.al : nullptr, *d_be = beta ? E.be : nullptr, *d_ch = chisq ? E.ch : nullptr;
    double *d_y = ymod ? E.y : nullptr, *d_dy = dyda ? E.dy : nullptr;
    HIPCHECK(hipMemcpyAsync(E.a, a, sizeof(double) * nv * h->ma, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(hipEventRecord(h->ev0, h->stream));
    if (use_v2(h)) {
        const int grid = std::min(nvec, h->slots2);
        rc = dispatch2(h->var, [&]<int NT, int R>() {
            hipLaunchKernelGGL((psk::eval_kernel2<NT, R, kMaxFac>), dim3(grid), dim3(psk::kWG2), 0, h->stream, h->dP2,
                               h->S2, E.a, nvec, conw_r, d_al, d_be, d_ch, d_y, d_dy);
            return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
        });
    } else {
        const int grid = std::min(nvec, h->slots);
        rc = dispatch(h->var, h->precision == PS_MIXED, [&]<int NT, int R, bool LDSB, bool MIX>() {
            return launch_eval<NT, R, LDSB, MIX>(h, grid, E.a, nvec, d_al, d_be, d_ch, d_y, d_dy);
        });
    }
    if (rc != PS_OK) return ps_fail(rc, "ps_eval: launch failed: %s", hipGetErrorString(hipGetLastError()));
    HIPCHECK(hipEventRecord(h->ev1, h->stream));
    if (alpha) HIPCHECK(hipMemcpyAsync(alpha, d_al, sizeof(double) * nv * mf2, hipMemcpyDeviceToHost, h->stream));
    if (beta) HIPCHECK(hipMemcpyAsync(beta, d_be, sizeof(double) * nv * h->mfit, hipMemcpyDeviceToHost, h->stream));
    if (chisq) HIPCHECK(hipMemcpyAsync(chisq, d_ch, sizeof(double) * nv, hipMemcpyDeviceToHost, h->stream));
    if (ymod) HIPCHECK(hipMemcpyAsync(ymod, d_y, sizeof(double) * nv * h->npts, hipMemcpyDeviceToHost, h->stream));
    if (dyda)
        HIPCHECK(hipMemcpyAsync(dyda, d_dy, sizeof(double) * nv * h->npts * h->ma, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->kernel_ms += ms;
    h->launches += 1;
    return PS_OK;
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
