// ps_device.hip -- gfx950 kernels of the period scan and the device half of the C ABI.
//
// One 256-lane workgroup owns one (period, pole) Levenberg-Marquardt fit and runs it to termination
// without returning to the host: model + Jacobian (bright/conv), relative-lightcurve normalisation and
// normal equations (mrqcof), damped full-pivot Gauss-Jordan solve (gauss_errc), accept/reject and the
// driver's convexity/rms bookkeeping (mrqmin + period_search_BOINC.cpp:868-911).  Fits are drained by a
// grid-stride loop over the batch, so a launch sized to the resident-workgroup count keeps every CU busy
// with one per-slot scratch area per workgroup.
//
// Data layout in HBM (all FP64, SoA):
//   points  t, ex,ey,ez (Earth), sx,sy,sz (Sun), y, s2i=1/sig^2, w   [npts = ndata + 3 convexity rows]
//   facets  nx,ny,nz, darea [nfac]; dsph [nfac][nc]; pleg [nfac][npairs]; fc, fs [nfac][mmax+1]
//   per slot: J [npts][map] (raw then normalised Jacobian rows), ytemp [npts], dave [nlc][ma+1],
//             alpha [2][mfit][mfit], beta [2][mfit]
// Reference citations are relative to the reference repository root.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/periodsearch_amd.h"
#include "ps_internal.h"

namespace {

constexpr int kWG = 256;       // lanes per workgroup (4 wave64)
constexpr int kMaxFit = 64;    // fitted parameters (Gauss-Jordan matrix kept in LDS)
constexpr int kMaxMa = 72;     // all parameters
constexpr int kMaxFac = 512;   // facets (nrows <= 8)
constexpr double kTiny = 1e-8; // period_search/constants.h:10
constexpr double kPi = M_PI;

struct DevProblem {
    int npts, nobs, nlc, nf, nc, ma, mfit, map, npairs, mmax1;
    int ia_cl, ia_ls;          // cl / Lommel-Seeliger weights fitted
    int n_iter_max, n_iter_min;
    double iter_diff_max;
    double phi0, alamda_start, alamda_incr;
    double par1, par2, par3, log_cl;
    const double *t, *ex, *ey, *ez, *sx, *sy, *sz, *y, *s2i, *w;
    const int *lc_start, *lc_inrel, *lc_of;
    const double *nx, *ny, *nz, *darea, *dsph, *pleg, *fc, *fs;
    const int *pair_m;
    const int *fit_idx;
    const unsigned short *tri_jk;  // lower-triangle pair list (jj << 8 | kk)
    const double *cg_first;
    const double *pole_beta, *pole_lambda;  // [10] radians, beta measured from the pole
    const double *sig;                      // [npts] (observed rows only are read)
};

struct Slots {
    double *J, *ytemp, *dave, *alpha, *beta;
    size_t j_stride, y_stride, dave_stride, a_stride, b_stride;
};

struct Shared {
    double a[kMaxMa];      // parameters under evaluation
    double cg[kMaxMa];     // accepted parameters
    double atry[kMaxMa];
    double area[kMaxFac];
    double g[kMaxFac];
    double cov[kMaxFit * kMaxFit];
    double da[kMaxFit];
    double col[kMaxFit];
    double red_v[kWG / 64];
    int red_q[kWG / 64];
    int ipiv[kMaxFit];
    double scal[8];        // chisq, chck[3], piv
    int iscal[4];          // irow, icol, err
};

// ---------------------------------------------------------------------------------------------
// curv (period_search/curv.c:10-39): g = exp(sum_{m,l} Pleg*(c*Fc + s*Fs)), Area = Darea*g.
// ---------------------------------------------------------------------------------------------
__device__ void curv_lds(const DevProblem &P, const double *a, Shared &sh) {
    for (int i = threadIdx.x; i < P.nf; i += kWG) {
        double gsum = 0;
        int n = 0;
        const double *pl = P.pleg + size_t(i) * P.npairs;
        const double *fc = P.fc + size_t(i) * P.mmax1, *fs = P.fs + size_t(i) * P.mmax1;
        for (int q = 0; q < P.npairs; ++q) {
            const int m = P.pair_m[q];
            double fsum = a[n] * fc[m];
            ++n;
            if (m != 0) {
                fsum = fsum + a[n] * fs[m];
                ++n;
            }
            gsum = gsum + pl[q] * fsum;
        }
        const double gv = exp(gsum);
        sh.g[i] = gv;
        sh.area[i] = P.darea[i] * gv;
    }
}

// Pole rotation and its derivatives (blmatrix.c:9-46); bl[3][3], dbl[2][3][3].
__device__ void blmatrix(double bet, double lam, double bl[3][3], double dbl[2][3][3]) {
    double sb, cb, sl, cl;
    sincos(bet, &sb, &cb);
    sincos(lam, &sl, &cl);
    bl[0][0] = cb * cl; bl[0][1] = cb * sl; bl[0][2] = -sb;
    bl[1][0] = -sl;     bl[1][1] = cl;      bl[1][2] = 0;
    bl[2][0] = sb * cl; bl[2][1] = sb * sl; bl[2][2] = cb;
    dbl[0][0][0] = -sb * cl; dbl[0][0][1] = -sb * sl; dbl[0][0][2] = -cb;
    dbl[0][1][0] = 0;        dbl[0][1][1] = 0;        dbl[0][1][2] = 0;
    dbl[0][2][0] = cb * cl;  dbl[0][2][1] = cb * sl;  dbl[0][2][2] = -sb;
    dbl[1][0][0] = -cb * sl; dbl[1][0][1] = cb * cl;  dbl[1][0][2] = 0;
    dbl[1][1][0] = -cl;      dbl[1][1][1] = -sl;      dbl[1][1][2] = 0;
    dbl[1][2][0] = -sb * sl; dbl[1][2][1] = sb * cl;  dbl[1][2][2] = 0;
}

// ---------------------------------------------------------------------------------------------
// bright (period_search/bright.c:14-126) for one observed point: brightness and dyda[0..ma-1].
// The three rotation derivatives use S = sum_vis Area*dsmu*n and S0 = sum_vis Area*dsmu0*n, then
// dyda_k = Scale*(S.de_k + S0.de0_k) -- algebraically the reference's per-facet sum (bright.c:88-104).
// The shape derivatives contract X_j = Area_j*s_j (visible facets) with Dsph (Dg = g*Dsph, curv.c:35).
// ---------------------------------------------------------------------------------------------
template <int NC, bool EXACT>
__device__ void bright_point(const DevProblem &P, int p, const double *a, const Shared &sh, const double bl[3][3],
                             const double dbl[2][3][3], bool all_derivs, double *jrow, double &ymod) {
    const int nc = EXACT ? NC : P.nc;
    const double ee[3] = {P.ex[p], P.ey[p], P.ez[p]};
    const double ee0[3] = {P.sx[p], P.sy[p], P.sz[p]};
    const double t = P.t[p];
    const double cl = exp(a[nc + 6]);
    const double cls = a[nc + 7];
    // phase function (phasec.c:11-24)
    const double cos_alpha = ee[0] * ee0[0] + ee[1] * ee0[1] + ee[2] * ee0[2];
    const double alpha = acos(cos_alpha);
    const double p1 = a[nc + 3], p2 = a[nc + 4], p3 = a[nc + 5];
    const double ephi = exp(-alpha / p2);
    const double scale = 1 + p1 * ephi + p3 * alpha;
    const double dphp[3] = {ephi, p1 * ephi * alpha / (p2 * p2), alpha};
    // rotation (matrix.c:11-58); the rotation phase stays an unfused FP64 multiply-add
    double f = __dadd_rn(__dmul_rn(a[nc + 2], t), P.phi0);
    f = fmod(f, 2 * kPi);
    double sf, cf;
    sincos(f, &sf, &cf);
    const double fm[3][3] = {{cf, sf, 0}, {-sf, cf, 0}, {0, 0, 1}};
    const double dfm[3][3] = {{-t * sf, t * cf, 0}, {-t * cf, -t * sf, 0}, {0, 0, 0}};
    double tm[3][3], dt[3][3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s0 = s0 + fm[i][k] * bl[k][j];
                s1 = s1 + fm[i][k] * dbl[0][k][j];
                s2 = s2 + fm[i][k] * dbl[1][k][j];
                s3 = s3 + dfm[i][k] * bl[k][j];
            }
            tm[i][j] = s0;
            dt[0][i][j] = s1;
            dt[1][i][j] = s2;
            dt[2][i][j] = s3;
        }
    double e[3], e0[3], de[3][3], de0[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double s = 0, s0 = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            s = s + tm[i][j] * ee[j];
            s0 = s0 + tm[i][j] * ee0[j];
            double d = 0, d0 = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                d = d + dt[j][i][k] * ee[k];
                d0 = d0 + dt[j][i][k] * ee0[k];
            }
            de[i][j] = d;
            de0[i][j] = d0;
        }
        e[i] = s;
        e0[i] = s0;
    }
    double acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = 0;
    double br = 0, S0 = 0, S1 = 0, S2 = 0, T0 = 0, T1 = 0, T2 = 0, d1 = 0, d2 = 0;
    const int nf = P.nf;
    for (int j = 0; j < nf; ++j) {
        const double n0 = P.nx[j], n1 = P.ny[j], n2 = P.nz[j];
        const double mu = e[0] * n0 + e[1] * n1 + e[2] * n2;
        const double mu0 = e0[0] * n0 + e0[1] * n1 + e0[2] * n2;
        const bool vis = (mu > kTiny) && (mu0 > kTiny);
        const double rd = 1.0 / (mu + mu0);
        const double A = sh.area[j];
        const double s = mu * mu0 * (cl + cls * rd);
        const double X = vis ? A * s : 0.0;
        br += X;
        const double q0 = mu0 * rd, q = mu * rd;
        const double ad = vis ? A * (cls * (q0 * q0) + cl * mu0) : 0.0;
        const double ad0 = vis ? A * (cls * (q * q) + cl * mu) : 0.0;
        S0 += ad * n0; S1 += ad * n1; S2 += ad * n2;
        T0 += ad0 * n0; T1 += ad0 * n1; T2 += ad0 * n2;
        if (P.ia_cl || all_derivs) d1 += vis ? mu * mu0 * A : 0.0;
        if (P.ia_ls || all_derivs) d2 += vis ? A * mu * mu0 * rd : 0.0;
        const double *ds = P.dsph + size_t(j) * nc;
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (EXACT || k < nc) acc[k] += X * ds[k];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k)
        if (EXACT || k < nc) jrow[k] = scale * acc[k];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        jrow[nc + k] = scale * (S0 * de[0][k] + S1 * de[1][k] + S2 * de[2][k] +
                                (T0 * de0[0][k] + T1 * de0[1][k] + T2 * de0[2][k]));
#pragma unroll
    for (int i = 0; i < 3; ++i) jrow[nc + 3 + i] = br * dphp[i];
    jrow[nc + 6] = scale * d1 * cl;
    jrow[nc + 7] = scale * d2;
    ymod = br * scale;
}

// Convexity pseudo-observation (conv.c:12-29): component c of sum Area*n, derivatives for the shape.
template <int NC, bool EXACT>
__device__ void conv_point(const DevProblem &P, int c, const Shared &sh, double *jrow, double &ymod) {
    const int nc = EXACT ? NC : P.nc;
    const double *nn = (c == 0) ? P.nx : (c == 1) ? P.ny : P.nz;
    double res = 0;
    double acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = 0;
    for (int i = 0; i < P.nf; ++i) {
        const double n = nn[i];
        res += sh.area[i] * n;
        const double dg = P.darea[i] * sh.g[i];
        const double *ds = P.dsph + size_t(i) * nc;
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (EXACT || k < nc) acc[k] += dg * ds[k] * n;
    }
#pragma unroll
    for (int k = 0; k < NC; ++k)
        if (EXACT || k < nc) jrow[k] = acc[k];
    for (int k = nc; k < P.ma; ++k) jrow[k] = 0;
    ymod = res;
}

// ---------------------------------------------------------------------------------------------
// mrqcof (period_search/mrqcof.c:16-148) for the parameters in sh.a.  Writes the full symmetric
// alpha[mfit][mfit] and beta[mfit] and returns chi2.  When raw_y / raw_dyda are given, the raw
// bright/conv outputs are copied there before normalisation.
// ---------------------------------------------------------------------------------------------
template <int NC, bool EXACT>
__device__ double mrqcof(const DevProblem &P, double *J, double *ytemp, double *dave, double *alpha, double *beta,
                         Shared &sh, double *raw_y, double *raw_dyda) {
    const int tid = threadIdx.x;
    const int nc = EXACT ? NC : P.nc;
    const int ma = P.ma, map = P.map, npts = P.npts;
    curv_lds(P, sh.a, sh);
    __syncthreads();
    double bl[3][3], dbl[2][3][3];
    blmatrix(sh.a[nc], sh.a[nc + 1], bl, dbl);
    // model and raw Jacobian rows
    for (int p = tid; p < npts; p += kWG) {
        double *jrow = J + size_t(p) * map;
        double ym;
        if (p < P.nobs)
            bright_point<NC, EXACT>(P, p, sh.a, sh, bl, dbl, raw_y != nullptr, jrow, ym);
        else
            conv_point<NC, EXACT>(P, p - P.nobs, sh, jrow, ym);
        ytemp[p] = ym;
        if (raw_y) {
            raw_y[p] = ym;
            for (int k = 0; k < ma; ++k) raw_dyda[size_t(p) * ma + k] = jrow[k];
        }
    }
    __syncthreads();
    // relative lightcurves: column sums in point order (mrqcof.c:53-85)
    const int nlc = P.nlc;
    for (int it = tid; it < nlc * (ma + 1); it += kWG) {
        const int L = it / (ma + 1), l = it - L * (ma + 1);
        if (!P.lc_inrel[L]) continue;
        double s = 0;
        for (int p = P.lc_start[L]; p < P.lc_start[L + 1]; ++p) s = s + (l < ma ? J[size_t(p) * map + l] : ytemp[p]);
        dave[it] = s;
    }
    __syncthreads();
    // normalisation (mrqcof.c:89-103): dy' = coef*(dy - y*dave/ave), coef = sig*Lpoints/ave
    for (int it = tid; it < P.nobs * ma; it += kWG) {
        const int p = it / ma, l = it - p * ma;
        const int L = P.lc_of[p];
        if (!P.lc_inrel[L]) continue;
        const double *dv = dave + size_t(L) * (ma + 1);
        const double ave = dv[ma];
        const double coef = P.sig[p] * (P.lc_start[L + 1] - P.lc_start[L]) / ave;
        double *jr = J + size_t(p) * map;
        jr[l] = coef * (jr[l] - ytemp[p] * dv[l] / ave);
    }
    __syncthreads();
    for (int p = tid; p < P.nobs; p += kWG) {
        const int L = P.lc_of[p];
        if (!P.lc_inrel[L]) continue;
        const double *dv = dave + size_t(L) * (ma + 1);
        const double ave = dv[ma];
        const double coef = P.sig[p] * (P.lc_start[L + 1] - P.lc_start[L]) / ave;
        ytemp[p] = coef * ytemp[p];
        J[size_t(p) * map] = 0;
    }
    __syncthreads();
    // normal equations in point order (mrqcof.c:105-145)
    const int mfit = P.mfit;
    const int ntri = mfit * (mfit + 1) / 2;
    for (int it = tid; it < ntri + mfit + 1; it += kWG) {
        if (it < ntri) {
            const unsigned jk = P.tri_jk[it];
            const int jj = jk >> 8, kk = jk & 255;
            const int l = P.fit_idx[jj], m = P.fit_idx[kk];
            double acc = 0;
            for (int p = 0; p < npts; ++p) {
                const double *jr = J + size_t(p) * map;
                const double wt = jr[l] * P.s2i[p] * P.w[p];
                acc += wt * jr[m];
            }
            alpha[jj * mfit + kk] = acc;
            alpha[kk * mfit + jj] = acc;
        } else if (it < ntri + mfit) {
            const int jj = it - ntri;
            const int l = P.fit_idx[jj];
            double acc = 0;
            for (int p = 0; p < npts; ++p) {
                const double dy = P.y[p] - ytemp[p];
                acc += dy * (J[size_t(p) * map + l] * P.s2i[p] * P.w[p]);
            }
            beta[jj] = acc;
        } else {
            double acc = 0;
            for (int p = 0; p < npts; ++p) {
                const double dy = P.y[p] - ytemp[p];
                acc += dy * dy * P.s2i[p] * P.w[p];
            }
            sh.scal[0] = acc;
        }
    }
    __syncthreads();
    return sh.scal[0];
}

// ---------------------------------------------------------------------------------------------
// Full-pivot Gauss-Jordan on sh.cov (n x n) with right-hand side sh.da (gauss_errc.c:9-73).
// Pivot: the largest |a| over unpivoted rows/columns, ties to the last one in row-major order
// (the reference's `>=` scan).  Returns 0, 1 (no valid pivot: all candidates NaN) or 2 (zero pivot).
// The inverse's column unscramble is skipped: only the solution is used (mrqmin.c:61-79).
// ---------------------------------------------------------------------------------------------
__device__ int gauss_jordan(Shared &sh, int n) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *a = sh.cov, *b = sh.da;
    for (int j = tid; j < n; j += kWG) sh.ipiv[j] = 0;
    __syncthreads();
    for (int i = 0; i < n; ++i) {
        double bv = -1.0;
        int bq = -1;
        for (int q = tid; q < n * n; q += kWG) {
            const int j = q / n, k = q - j * n;
            if (sh.ipiv[j] != 1 && sh.ipiv[k] == 0) {
                const double v = fabs(a[q]);
                if (v >= bv) {
                    bv = v;
                    bq = q;
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int oq = __shfl_xor(bq, off);
            if (ov > bv || (ov == bv && oq > bq)) {
                bv = ov;
                bq = oq;
            }
        }
        if (lane == 0) {
            sh.red_v[wave] = bv;
            sh.red_q[wave] = bq;
        }
        __syncthreads();
        if (tid == 0) {
            double v = sh.red_v[0];
            int q = sh.red_q[0];
            for (int w = 1; w < kWG / 64; ++w)
                if (sh.red_v[w] > v || (sh.red_v[w] == v && sh.red_q[w] > q)) {
                    v = sh.red_v[w];
                    q = sh.red_q[w];
                }
            sh.iscal[0] = q < 0 ? -1 : q / n;
            sh.iscal[1] = q < 0 ? -1 : q % n;
        }
        __syncthreads();
        const int irow = sh.iscal[0], icol = sh.iscal[1];
        if (irow < 0) return 1;
        if (irow != icol)
            for (int l = tid; l <= n; l += kWG) {
                double *x = (l < n) ? &a[irow * n + l] : &b[irow];
                double *y = (l < n) ? &a[icol * n + l] : &b[icol];
                const double tmp = *x;
                *x = *y;
                *y = tmp;
            }
        __syncthreads();
        const double piv = a[icol * n + icol];
        if (piv == 0.0) return 2;
        const double pivinv = 1.0 / piv;
        if (tid == 0) sh.ipiv[icol] += 1;
        for (int l = tid; l <= n; l += kWG) {
            if (l < n)
                a[icol * n + l] = (l == icol ? 1.0 : a[icol * n + l]) * pivinv;
            else
                b[icol] *= pivinv;
        }
        for (int ll = tid; ll < n; ll += kWG)
            if (ll != icol) sh.col[ll] = a[ll * n + icol];
        __syncthreads();
        for (int it = tid; it < n * (n + 1); it += kWG) {
            const int ll = it / (n + 1), l = it - ll * (n + 1);
            if (ll == icol) continue;
            const double dum = sh.col[ll];
            if (l < n)
                a[ll * n + l] = (l == icol ? 0.0 : a[ll * n + l]) - a[icol * n + l] * dum;
            else
                b[ll] -= b[icol] * dum;
        }
        __syncthreads();
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// One (period, pole) fit per workgroup: period_search_BOINC.cpp:819-911 with mrqmin.c:11-110.
// ---------------------------------------------------------------------------------------------
template <int NC, bool EXACT>
__global__ __launch_bounds__(kWG) void fit_kernel(DevProblem P, Slots S, const double *freq, const int *pole,
                                                  int nfits, double conw_r, ps_fit_result *out) {
    __shared__ Shared sh;
    const int tid = threadIdx.x;
    const int nc = EXACT ? NC : P.nc;
    const int ma = P.ma, mfit = P.mfit;
    double *J = S.J + S.j_stride * blockIdx.x;
    double *ytemp = S.ytemp + S.y_stride * blockIdx.x;
    double *dave = S.dave + S.dave_stride * blockIdx.x;
    double *alpha[2] = {S.alpha + S.a_stride * (2 * blockIdx.x), S.alpha + S.a_stride * (2 * blockIdx.x + 1)};
    double *beta[2] = {S.beta + S.b_stride * (2 * blockIdx.x), S.beta + S.b_stride * (2 * blockIdx.x + 1)};

    for (int fi = blockIdx.x; fi < nfits; fi += gridDim.x) {
        // initial parameters (:819-844)
        if (tid < ma) {
            double v;
            const int m = pole[fi] - 1;
            if (tid < nc)
                v = P.cg_first[tid];
            else if (tid == nc)
                v = P.pole_beta[m];
            else if (tid == nc + 1)
                v = P.pole_lambda[m];
            else if (tid == nc + 2)
                v = 24 * 2 * kPi / (1 / freq[fi]);
            else if (tid < nc + 6)
                v = (tid == nc + 3) ? P.par1 : (tid == nc + 4) ? P.par2 : P.par3;
            else if (tid == nc + 6)
                v = P.log_cl;
            else
                v = 1;
            sh.cg[tid] = v;
        }
        __syncthreads();
        int cur = 0;  // alpha[cur], beta[cur] hold the accepted normal equations
        double alamda = -1, ochisq = 0, chisq = 0, rchisq = -1;
        double iter_diff = 1e40, dev_old = 1e30, dev_new = 0;
        double chck0 = 0, chck1 = 0, chck2 = 0;
        int niter = 0, nfail = 0;
        while (((niter < P.n_iter_max) && (iter_diff > P.iter_diff_max)) || (niter < P.n_iter_min)) {
            // ---- mrqmin ----
            if (alamda < 0) {
                alamda = P.alamda_start;
                if (tid < ma) sh.a[tid] = sh.cg[tid];
                __syncthreads();
                ochisq = mrqcof<NC, EXACT>(P, J, ytemp, dave, alpha[cur], beta[cur], sh, nullptr, nullptr);
                if (tid < ma) sh.atry[tid] = sh.cg[tid];
            }
            for (int it = tid; it < mfit * mfit; it += kWG) {
                const int j = it / mfit, k = it - j * mfit;
                const double v = alpha[cur][it];
                sh.cov[it] = (j == k) ? v * (1 + alamda) : v;
            }
            for (int j = tid; j < mfit; j += kWG) sh.da[j] = beta[cur][j];
            __syncthreads();
            const int err = gauss_jordan(sh, mfit);
            __syncthreads();
            if (err == 0) {
                if (tid < mfit) {
                    const int l = P.fit_idx[tid];
                    sh.atry[l] = sh.cg[l] + sh.da[tid];
                }
                __syncthreads();
                if (tid < ma) sh.a[tid] = sh.atry[tid];
                __syncthreads();
                chisq = mrqcof<NC, EXACT>(P, J, ytemp, dave, alpha[cur ^ 1], beta[cur ^ 1], sh, nullptr, nullptr);
                if (chisq < ochisq) {
                    alamda = alamda / P.alamda_incr;
                    cur ^= 1;
                    if (tid < ma) sh.cg[tid] = sh.atry[tid];
                } else {
                    alamda = P.alamda_incr * alamda;
                    chisq = ochisq;
                }
                __syncthreads();
            } else {
                ++nfail;
            }
            ++niter;
            // ---- driver bookkeeping (:873-892) ----
            if (niter == 1 || chisq < ochisq) {
                ochisq = chisq;
                if (tid < ma) sh.a[tid] = sh.cg[tid];
                __syncthreads();
                curv_lds(P, sh.a, sh);
                __syncthreads();
                if (tid < 3) {
                    const double *nn = (tid == 0) ? P.nx : (tid == 1) ? P.ny : P.nz;
                    double s = 0;
                    for (int j = 0; j < P.nf; ++j) s = s + sh.area[j] * nn[j];
                    sh.scal[1 + tid] = s;
                }
                __syncthreads();
                chck0 = sh.scal[1];
                chck1 = sh.scal[2];
                chck2 = sh.scal[3];
                rchisq = chisq - (chck0 * chck0 + chck1 * chck1 + chck2 * chck2) * (conw_r * conw_r);
                __syncthreads();
            }
            dev_new = sqrt(rchisq / P.nobs);
            if (dev_old - dev_new > 1e-10) {
                iter_diff = dev_old - dev_new;
                dev_old = dev_new;
            }
        }
        // ---- fit result (:900-911); totarea from the last curv evaluation ----
        if (tid == 0) {
            double totarea = 0;
            for (int j = 0; j < P.nf; ++j) totarea = totarea + sh.area[j];
            ps_fit_result r;
            r.dev = dev_new;
            r.prd = 2 * kPi / sh.cg[nc + 2];
            r.dark = sqrt(chck0 * chck0 + chck1 * chck1 + chck2 * chck2);
            r.totarea = totarea;
            r.la = (180 / kPi) * sh.cg[nc + 1];
            r.be = 90 - (180 / kPi) * sh.cg[nc];
            r.chisq = chisq;
            r.niter = niter;
            r.nsolve_fail = nfail;
            out[fi] = r;
        }
        __syncthreads();
    }
}

// mrqcof for externally supplied parameter vectors (ps_eval).
template <int NC, bool EXACT>
__global__ __launch_bounds__(kWG) void eval_kernel(DevProblem P, Slots S, const double *avec, int nvec, double *alpha,
                                                   double *beta, double *chisq, double *raw_y, double *raw_dyda) {
    __shared__ Shared sh;
    const int tid = threadIdx.x;
    const int mfit = P.mfit;
    double *J = S.J + S.j_stride * blockIdx.x;
    double *ytemp = S.ytemp + S.y_stride * blockIdx.x;
    double *dave = S.dave + S.dave_stride * blockIdx.x;
    double *al = S.alpha + S.a_stride * (2 * blockIdx.x);
    double *be = S.beta + S.b_stride * (2 * blockIdx.x);
    for (int v = blockIdx.x; v < nvec; v += gridDim.x) {
        if (tid < P.ma) sh.a[tid] = avec[size_t(v) * P.ma + tid];
        __syncthreads();
        const double c = mrqcof<NC, EXACT>(P, J, ytemp, dave, al, be, sh, raw_y ? raw_y + size_t(v) * P.npts : nullptr,
                                           raw_dyda ? raw_dyda + size_t(v) * P.npts * P.ma : nullptr);
        for (int it = tid; it < mfit * mfit; it += kWG)
            if (alpha) alpha[size_t(v) * mfit * mfit + it] = al[it];
        for (int it = tid; it < mfit; it += kWG)
            if (beta) beta[size_t(v) * mfit + it] = be[it];
        if (tid == 0 && chisq) chisq[v] = c;
        __syncthreads();
    }
}

template <class T>
T *dev_upload(const T *h, size_t n, std::vector<void *> &owned) {
    T *d = nullptr;
    if (hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    owned.push_back(d);
    if (n && hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

}  // namespace

struct ps_handle {
    int device = 0;
    int precision = PS_FP64;
    int ndata = 0, npts = 0, nlc = 0, nfac = 0, ncoef = 0, ma = 0, mfit = 0, nmax = 0, slots = 0;
    double freq_start = 0, freq_end = 0, freq_step = 0, escl = 0, conw_r0 = 0;
    std::vector<double> s2i_host;  // 1/sig^2 per point (convexity rows set per call)
    double conw_loaded = -1;
    DevProblem P{};
    Slots S{};
    double *d_s2i = nullptr;
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

template <int NC, bool EXACT>
int launch_fit(ps_handle *h, int grid, int nfits, double conw_r) {
    hipLaunchKernelGGL((fit_kernel<NC, EXACT>), dim3(grid), dim3(kWG), 0, h->stream, h->P, h->S, h->d_freq, h->d_pole,
                       nfits, conw_r, h->d_out);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
}
template <int NC, bool EXACT>
int launch_eval(ps_handle *h, int grid, const double *a, int nvec, double *al, double *be, double *ch, double *ry,
                double *rd) {
    hipLaunchKernelGGL((eval_kernel<NC, EXACT>), dim3(grid), dim3(kWG), 0, h->stream, h->P, h->S, a, nvec, al, be, ch,
                       ry, rd);
    return hipGetLastError() == hipSuccess ? PS_OK : PS_ERR_DEVICE;
}

// Shape-specialised dispatch: exact coefficient counts of Lmax = Mmax = 3..6, else padded bounds.
template <class F>
int dispatch_nc(int nc, F &&f) {
    switch (nc) {
        case 16: return f.template operator()<16, true>();
        case 25: return f.template operator()<25, true>();
        case 36: return f.template operator()<36, true>();
        case 49: return f.template operator()<49, true>();
        default: break;
    }
    if (nc <= 16) return f.template operator()<16, false>();
    if (nc <= 32) return f.template operator()<32, false>();
    return f.template operator()<64, false>();
}

int set_conw(ps_handle *h, double conw_r) {
    if (h->conw_loaded == conw_r) return PS_OK;
    const double sig = 1 / conw_r;  // period_search_BOINC.cpp:721-726
    for (int j = 0; j < 3; ++j) h->s2i_host[h->ndata + j] = 1 / (sig * sig);
    HIPCHECK(hipMemcpyAsync(h->d_s2i + h->ndata, h->s2i_host.data() + h->ndata, 3 * sizeof(double),
                            hipMemcpyHostToDevice, h->stream));
    h->conw_loaded = conw_r;
    return PS_OK;
}
}  // namespace

extern "C" int ps_prepare(const ps_input *in, const ps_geometry *geo, int device, ps_handle **out) {
    if (!in || !geo || !out) return ps_fail(PS_ERR_ARG, "ps_prepare: null argument");
    *out = nullptr;
    const int ncoef = geo->ncoef, ma = ncoef + 8;
    int ia[kMaxMa + 8];
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
    if (mfit > kMaxFit || ma > kMaxMa)
        return ps_fail(PS_ERR_UNSUPPORTED, "GPU kernels support up to %d fitted parameters (got %d; Lmax <= 6)",
                       kMaxFit, mfit);
    if (geo->nfac > kMaxFac) return ps_fail(PS_ERR_UNSUPPORTED, "GPU kernels support nrows <= 8 (nfac <= %d)", kMaxFac);
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
        y(npts, 0), w(npts, 1), sig(npts, 1);
    h->s2i_host.assign(npts, 0);
    for (int p = 0; p < nobs; ++p) {
        t[p] = in->tim[p];
        ex[p] = in->ee[3 * p];
        ey[p] = in->ee[3 * p + 1];
        ez[p] = in->ee[3 * p + 2];
        sx[p] = in->ee0[3 * p];
        sy[p] = in->ee0[3 * p + 1];
        sz[p] = in->ee0[3 * p + 2];
        y[p] = in->brightness[p];
        w[p] = in->weight[p];
        sig[p] = in->sig[p];
        h->s2i_host[p] = 1 / (in->sig[p] * in->sig[p]);
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
    std::vector<double> nx(nf), ny(nf), nz(nf);
    for (int i = 0; i < nf; ++i) {
        nx[i] = geo->nor[3 * i];
        ny[i] = geo->nor[3 * i + 1];
        nz[i] = geo->nor[3 * i + 2];
    }
    std::vector<unsigned short> tri;
    for (int j = 0; j < mfit; ++j)
        for (int k = 0; k <= j; ++k) tri.push_back((unsigned short)((j << 8) | k));
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
    P.map = (ma + 1) & ~1;
    P.npairs = geo->npairs;
    P.mmax1 = geo->mmax + 1;
    P.ia_cl = ia[ncoef + 6];
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
    P.y = dev_upload(y.data(), npts, o);
    P.w = dev_upload(w.data(), npts, o);
    P.sig = dev_upload(sig.data(), npts, o);
    h->d_s2i = dev_upload(h->s2i_host.data(), npts, o);
    P.s2i = h->d_s2i;
    P.lc_start = dev_upload(lc_start.data(), lc_start.size(), o);
    P.lc_inrel = dev_upload(lc_inrel.data(), lc_inrel.size(), o);
    P.lc_of = dev_upload(lc_of.data(), lc_of.size(), o);
    P.nx = dev_upload(nx.data(), nf, o);
    P.ny = dev_upload(ny.data(), nf, o);
    P.nz = dev_upload(nz.data(), nf, o);
    P.darea = dev_upload(geo->darea, nf, o);
    P.dsph = dev_upload(geo->dsph, size_t(nf) * ncoef, o);
    P.pleg = dev_upload(geo->pleg, size_t(nf) * geo->npairs, o);
    P.fc = dev_upload(geo->fc, size_t(nf) * (geo->mmax + 1), o);
    P.fs = dev_upload(geo->fs, size_t(nf) * (geo->mmax + 1), o);
    P.pair_m = dev_upload(geo->pair_m, geo->npairs, o);
    P.fit_idx = dev_upload(fit_idx.data(), fit_idx.size(), o);
    P.tri_jk = dev_upload(tri.data(), tri.size(), o);
    P.cg_first = dev_upload(geo->cg_first, ncoef, o);
    P.pole_beta = dev_upload(pb, PS_N_POLES, o);
    P.pole_lambda = dev_upload(pl, PS_N_POLES, o);
    for (void *ptr : o)
        if (!ptr) return fail(ps_fail(PS_ERR_DEVICE, "device upload failed"));

    // resident workgroups: one scratch slot each
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(ps_fail(PS_ERR_DEVICE, "device query failed"));
    int per_cu = 1;
    dispatch_nc(ncoef, [&]<int NC, bool EXACT>() {
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fit_kernel<NC, EXACT>, kWG, 0);
        return 0;
    });
    h->slots = std::max(1, per_cu) * prop.multiProcessorCount;
    Slots &S = h->S;
    S.j_stride = size_t(npts) * P.map;
    S.y_stride = size_t(npts);
    S.dave_stride = size_t(h->nlc) * (ma + 1);
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
    if (precision != PS_FP64) return ps_fail(PS_ERR_UNSUPPORTED, "precision %d not available", precision);
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

extern "C" int ps_fit_batch(ps_handle *h, int nfits, const double *freq, const int *pole, double conw_r,
                            ps_fit_result *out) {
    if (!h || nfits < 0 || (nfits && (!freq || !pole || !out))) return ps_fail(PS_ERR_ARG, "ps_fit_batch: bad argument");
    if (nfits == 0) return PS_OK;
    for (int i = 0; i < nfits; ++i)
        if (pole[i] < 1 || pole[i] > PS_N_POLES || !(freq[i] > 0))
            return ps_fail(PS_ERR_ARG, "ps_fit_batch: fit %d has pole %d / freq %g", i, pole[i], freq[i]);
    if (!(conw_r > 0)) return ps_fail(PS_ERR_ARG, "ps_fit_batch: conw_r must be positive");
    HIPCHECK(hipSetDevice(h->device));
    if (nfits > h->cap_fits) {
        hipFree(h->d_freq);
        hipFree(h->d_pole);
        hipFree(h->d_out);
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
    rc = dispatch_nc(h->ncoef, [&]<int NC, bool EXACT>() { return launch_fit<NC, EXACT>(h, grid, nfits, conw_r); });
    if (rc != PS_OK) return ps_fail(rc, "fit kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
    HIPCHECK(hipEventRecord(h->ev1, h->stream));
    HIPCHECK(hipMemcpyAsync(out, h->d_out, sizeof(ps_fit_result) * nfits, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->kernel_ms += ms;
    h->launches += 1;
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
        rc = dispatch_nc(h->ncoef, [&]<int NC, bool EXACT>() {
            return launch_eval<NC, EXACT>(h, grid, d_a, nvec, d_al, d_be, d_ch, d_y, d_dy);
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
