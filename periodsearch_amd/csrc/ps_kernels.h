// ps_kernels.h -- gfx950 device code of the period scan (included by ps_device.hip only).
//
// One workgroup of 512 lanes (8 wave64, two per SIMD) owns one (period, pole) Levenberg-Marquardt fit
// and runs it to termination on the device: model + Jacobian (bright/conv), relative-lightcurve
// normalisation and normal equations (mrqcof), damped full-pivot Gauss-Jordan (gauss_errc),
// accept/reject and the driver's convexity/rms bookkeeping (mrqmin + period_search_BOINC.cpp:868-911).
//
// One evaluation (mrqcof) is three phases:
//  B  model and raw Jacobian.  Each wave takes 16-point tiles; lane l works on point (l & 15) and
//     facet (4c + (l >> 4)) of facet chunk c.  It computes X = Area*s (the visible, illuminated
//     Lommel-Seeliger + Lambert term) elementwise.  The shape Jacobian X*Dsph (16 points x 16
//     coefficients per tile) runs on v_mfma_f64_16x16x4f64 with Dsph staged in LDS; leftover
//     coefficients, the brightness and the two 3-vector sums that give the rotation derivatives stay
//     on the VALU and are reduced across the four lane groups with shuffles.  Raw rows go to the
//     slot's J scratch (row pitch kMap).
//  C  per relative lightcurve: column sums of the raw rows (period_search/mrqcof.c:53-85).
//  D  normal equations.  Each wave takes 4-point K-steps.  Lanes build the normalised rows of the
//     fitted parameters plus the residual as an extra column, and v_mfma_f64_16x16x4f64 accumulates
//     sum_p w_p v_p v_p^T over the lower 16x16 blocks.  The residual row of that product is beta and
//     its corner is chi2.  Per-wave partial sums are added in wave order (deterministic).
//
// Operand layouts (measured): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k; FP64 result
// D[4r + (l >> 4)][l & 15] in register r of lane l.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace psk {

constexpr int kWG = 512;
constexpr int kWaves = kWG / 64;
constexpr int kMaxFit = 63;     // fitted parameters; +1 residual column = 64
constexpr int kMaxMa = 72;
constexpr int kMap = 64;        // J row pitch (doubles)
constexpr int kMaxFac = 512;    // nrows <= 8
constexpr int kDsphLds = 320 * 49;  // FP64 Dsph held in LDS (facets x pitch), else read from HBM/L2
constexpr double kTiny = 1e-8;  // period_search/constants.h:10
constexpr double kPi = M_PI;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct DevProblem {
    int npts, nobs, nlc, nf, nc, ma, mfit, npairs, mmax1, dpitch;
    int ia_cl, ia_ls;
    int n_iter_max, n_iter_min;
    double iter_diff_max;
    double phi0, alamda_start, alamda_incr;
    double par1, par2, par3, log_cl;
    const double *t, *ex, *ey, *ez, *sx, *sy, *sz, *y, *s2i, *w, *sig;
    const int *lc_start, *lc_inrel, *lc_of;
    const double *nx, *ny, *nz, *darea;
    const double *dsph;  // [nf][dpitch], zero-padded beyond nc
    const float *dsphf;  // the same in FP32 (mixed precision)
    const double *pleg, *fc, *fs;
    const int *pair_m;
    const int *fit_idx;
    const double *cg_first;
    const double *pole_beta, *pole_lambda;
};

struct Slots {
    double *J, *ytemp, *dave, *alpha, *beta;
    size_t j_stride, y_stride, dave_stride, a_stride, b_stride;
};

struct Shared {
    double a[kMaxMa];     // parameters under evaluation
    double cg[kMaxMa];    // accepted parameters
    double atry[kMaxMa];
    double nx[kMaxFac], ny[kMaxFac], nz[kMaxFac], area[kMaxFac], g[kMaxFac];
    float nxf[kMaxFac], nyf[kMaxFac], nzf[kMaxFac], areaf[kMaxFac];  // FP32 copies (mixed precision)
    double da[64], col[64], b0[64];
    int indxr[64], indxc[64];
    double chi_part[kWaves];
    double red_v[kWaves];
    int red_q[kWaves];
    int ipiv[64];
    double scal[8];
    int iscal[4];
    double bl[9], dbl[18];  // pole rotation and its beta/lambda derivatives (blmatrix)
    union {
        double dsph[kDsphLds];       // phase B, FP64
        float dsphf[2 * kDsphLds];   // phase B, FP32
        struct {
            double cov[64 * 64];     // Gauss-Jordan: the damped matrix, then its inverse
            double cov0[64 * 64];    // the damped matrix kept for iterative refinement
        } g;
        struct {
            double stage[kWaves][4 * 64];
            double red[64 * 64];
        } d;                         // phase D
    } u;
};

// Precision traits of the two MFMA contractions.  FP64: v_mfma_f64_16x16x4f64, result row 4r + (l>>4)
// in register r; FP32: v_mfma_f32_16x16x4f32, result row 4(l>>4) + r (lane layouts measured on gfx950).
template <bool MIX>
struct Prec;
template <>
struct Prec<false> {
    using T = double;
    using V4 = d4;
    __device__ static V4 mma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int r, int kg) { return 4 * r + kg; }
    __device__ static V4 zero() { return V4{0, 0, 0, 0}; }
    __device__ static double rcp(double x) {  // v_rcp_f64 and two Newton steps: within an ulp of 1/x
        double y = __builtin_amdgcn_rcp(x);
        double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
        e = fma(-x, y, 1.0);
        return fma(y, e, y);
    }
};
template <>
struct Prec<true> {
    using T = float;
    using V4 = f4;
    __device__ static V4 mma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int r, int kg) { return 4 * kg + r; }
    __device__ static V4 zero() { return V4{0, 0, 0, 0}; }
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

// ---------------------------------------------------------------------------------------------
// curv (period_search/curv.c:10-39): g = exp(sum_{m,l} Pleg*(c*Fc + s*Fs)), Area = Darea*g.
// ---------------------------------------------------------------------------------------------
__device__ inline void curv_lds(const DevProblem &P, const double *a, Shared &sh) {
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

// Pole rotation and its derivatives (blmatrix.c:9-46).
__device__ inline void blmatrix(double bet, double lam, double bl[3][3], double dbl[2][3][3]) {
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

// Pole rotation into LDS once per evaluation (blmatrix.c:9-46).
__device__ inline void blmatrix_lds(double bet, double lam, Shared &sh) {
    if (threadIdx.x == 0) {
        double bl[3][3], dbl[2][3][3];
        blmatrix(bet, lam, bl, dbl);
        for (int i = 0; i < 9; ++i) sh.bl[i] = bl[i / 3][i % 3];
        for (int i = 0; i < 18; ++i) sh.dbl[i] = dbl[i / 9][(i / 3) % 3][i % 3];
    }
}

// Rotation phase of epoch t: an unfused FP64 multiply-add reduced modulo 2 pi exactly (matrix.c:18-19).
__device__ inline void rotation_phase(double omg, double t, double phi0, double &sf, double &cf) {
    double f = __dadd_rn(__dmul_rn(omg, t), phi0);
    f = fmod(f, 2 * kPi);
    sincos(f, &sf, &cf);
}

// Earth and Sun directions of point p in the body frame (bright.c:41-57 with matrix.c).
__device__ inline void point_dirs(const DevProblem &P, const double *a, int p, const Shared &sh, double e[3],
                                  double e0[3]) {
    const double ee[3] = {P.ex[p], P.ey[p], P.ez[p]};
    const double ee0[3] = {P.sx[p], P.sy[p], P.sz[p]};
    double sf, cf;
    rotation_phase(a[P.nc + 2], P.t[p], P.phi0, sf, cf);
    const double fm[3][3] = {{cf, sf, 0}, {-sf, cf, 0}, {0, 0, 1}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double s = 0, s0 = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double tm = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) tm = tm + fm[i][k] * sh.bl[3 * k + j];
            s = s + tm * ee[j];
            s0 = s0 + tm * ee0[j];
        }
        e[i] = s;
        e0[i] = s0;
    }
}

// Derivatives of the body-frame directions w.r.t. beta, lambda, omega (de[i][k]), the phase
// function and its derivatives (bright.c:30-57, matrix.c:11-58, phasec.c:11-24).
struct PointDer {
    double de[3][3], de0[3][3], scale, dphp[3];
};

__device__ inline void point_derivs(const DevProblem &P, const double *a, int p, const Shared &sh, PointDer &G) {
    const int nc = P.nc;
    const double ee[3] = {P.ex[p], P.ey[p], P.ez[p]};
    const double ee0[3] = {P.sx[p], P.sy[p], P.sz[p]};
    const double t = P.t[p];
    const double cos_alpha = ee[0] * ee0[0] + ee[1] * ee0[1] + ee[2] * ee0[2];
    const double alpha = acos(cos_alpha);
    const double p1 = a[nc + 3], p2 = a[nc + 4], p3 = a[nc + 5];
    const double ephi = exp(-alpha / p2);
    G.scale = 1 + p1 * ephi + p3 * alpha;
    G.dphp[0] = ephi;
    G.dphp[1] = p1 * ephi * alpha / (p2 * p2);
    G.dphp[2] = alpha;
    double sf, cf;
    rotation_phase(a[nc + 2], t, P.phi0, sf, cf);
    const double fm[3][3] = {{cf, sf, 0}, {-sf, cf, 0}, {0, 0, 1}};
    const double dfm[3][3] = {{-t * sf, t * cf, 0}, {-t * cf, -t * sf, 0}, {0, 0, 0}};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {  // de[i][j] = sum_k dtm[j][i][k] ee[k]
            double d = 0, d0 = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double dt = 0;
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    dt = dt + (j < 2 ? fm[i][q] * sh.dbl[9 * j + 3 * q + k] : dfm[i][q] * sh.bl[3 * q + k]);
                d = d + dt * ee[k];
                d0 = d0 + dt * ee0[k];
            }
            G.de[i][j] = d;
            G.de0[i][j] = d0;
        }
}

template <class T>
__device__ inline T sum_groups(T v) {  // sum over the four 16-lane groups of a wave
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// ---------------------------------------------------------------------------------------------
// Phase B: bright (period_search/bright.c:14-126) for every observed point, conv (conv.c:12-29) for
// the three convexity rows.  NT full 16-column MFMA tiles + R leftover coefficients (NT*16 + R >= nc).
// MIX: facet terms, rotation sums and the shape contraction in FP32; the body-frame directions and
// the brightness (the model value entering the residual) stay FP64.
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool LDSB, bool MIX>
__device__ void model_rows(const DevProblem &P, Shared &sh, double *J, double *ytemp, bool all_derivs) {
    using Pr = Prec<MIX>;
    using T = typename Pr::T;
    using V4 = typename Pr::V4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int pl = lane & 15, kg = lane >> 4;
    const int nc = P.nc, pitch = P.dpitch;
    const double *a = sh.a;
    const T *dsph, *snx, *sny, *snz, *sar;
    if constexpr (MIX) {
        dsph = LDSB ? sh.u.dsphf : P.dsphf;
        snx = sh.nxf; sny = sh.nyf; snz = sh.nzf; sar = sh.areaf;
    } else {
        dsph = LDSB ? sh.u.dsph : P.dsph;
        snx = sh.nx; sny = sh.ny; snz = sh.nz; sar = sh.area;
    }
    const double cl_d = exp(a[nc + 6]);
    const T cl = T(cl_d), cls = T(a[nc + 7]);
    const bool want_d1 = P.ia_cl || all_derivs, want_d2 = P.ia_ls || all_derivs;
    const int ntiles = (P.nobs + 15) >> 4;
    const int nchunk = P.nf >> 2;
    for (int tile = wave; tile < ntiles; tile += kWaves) {
        const int p = tile * 16 + pl;
        const bool valid = p < P.nobs;
        double ed[3], e0d[3];
        point_dirs(P, a, valid ? p : 0, sh, ed, e0d);
        const T e[3] = {T(ed[0]), T(ed[1]), T(ed[2])}, e0[3] = {T(e0d[0]), T(e0d[1]), T(e0d[2])};
        V4 acc[NT];
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) acc[tt] = Pr::zero();
        T racc[R > 0 ? R : 1];
#pragma unroll
        for (int r = 0; r < R; ++r) racc[r] = 0;
        double br = 0;
        T S0 = 0, S1 = 0, S2 = 0, T0 = 0, T1 = 0, T2 = 0, d1 = 0, d2 = 0;
        for (int c = 0; c < nchunk; ++c) {
            const int j = 4 * c + kg;
            const T n0 = snx[j], n1 = sny[j], n2 = snz[j], A = sar[j];
            const T mu = e[0] * n0 + e[1] * n1 + e[2] * n2;
            const T mu0 = e0[0] * n0 + e0[1] * n1 + e0[2] * n2;
            const bool vis = valid && (mu > T(kTiny)) && (mu0 > T(kTiny));
            const T rd = Pr::rcp(mu + mu0);
            const T s = mu * mu0 * (cl + cls * rd);
            const T X = vis ? A * s : T(0);
            br += double(X);
            const T q0 = mu0 * rd, q = mu * rd;
            const T ad = vis ? A * (cls * (q0 * q0) + cl * mu0) : T(0);
            const T ad0 = vis ? A * (cls * (q * q) + cl * mu) : T(0);
            S0 += ad * n0; S1 += ad * n1; S2 += ad * n2;
            T0 += ad0 * n0; T1 += ad0 * n1; T2 += ad0 * n2;
            if (want_d1) d1 += vis ? mu * mu0 * A : T(0);
            if (want_d2) d2 += vis ? A * mu * mu0 * rd : T(0);
            const T *brow = dsph + j * pitch;
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) acc[tt] = Pr::mma(X, brow[16 * tt + pl], acc[tt]);
#pragma unroll
            for (int r = 0; r < R; ++r) racc[r] += X * brow[16 * NT + r];
        }
        br = sum_groups(br);
        S0 = sum_groups(S0); S1 = sum_groups(S1); S2 = sum_groups(S2);
        T0 = sum_groups(T0); T1 = sum_groups(T1); T2 = sum_groups(T2);
        if (want_d1) d1 = sum_groups(d1);
        if (want_d2) d2 = sum_groups(d2);
#pragma unroll
        for (int r = 0; r < R; ++r) racc[r] = sum_groups(racc[r]);
        PointDer G;
        point_derivs(P, a, valid ? p : 0, sh, G);
        const double scale = G.scale;
        if (kg == 0 && valid) {
            double *jr = J + size_t(p) * kMap;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (16 * NT + r < nc) jr[16 * NT + r] = scale * double(racc[r]);
            const double s0 = S0, s1 = S1, s2 = S2, t0 = T0, t1 = T1, t2 = T2;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                jr[nc + k] = scale * (s0 * G.de[0][k] + s1 * G.de[1][k] + s2 * G.de[2][k] +
                                      (t0 * G.de0[0][k] + t1 * G.de0[1][k] + t2 * G.de0[2][k]));
#pragma unroll
            for (int i = 0; i < 3; ++i) jr[nc + 3 + i] = br * G.dphp[i];
            jr[nc + 6] = scale * double(d1) * cl_d;
            jr[nc + 7] = scale * double(d2);
            ytemp[p] = br * scale;
        }
        // shape columns: register r of lane l holds point Pr::row(r, kg), coefficient 16tt + pl
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = Pr::row(r, kg);
            const double sc = __shfl(scale, i);
            const int pi = tile * 16 + i;
            if (pi < P.nobs) {
                double *jr = J + size_t(pi) * kMap;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt)
                    if (16 * tt + pl < nc) jr[16 * tt + pl] = sc * double(acc[tt][r]);
            }
        }
    }
    // convexity rows: res_c = sum Area*n_c, dres_k = sum Area*Dsph_k*n_c (conv.c; Area = Darea*g), FP64
    for (int it = threadIdx.x; it < 3 * (nc + 1); it += kWG) {
        const int c = it / (nc + 1), k = it - c * (nc + 1);
        const double *nn = (c == 0) ? sh.nx : (c == 1) ? sh.ny : sh.nz;
        double s = 0;
        if (k == nc) {
            for (int i = 0; i < P.nf; ++i) s += sh.area[i] * nn[i];
            ytemp[P.nobs + c] = s;
        } else {
            for (int i = 0; i < P.nf; ++i) s += sh.area[i] * P.dsph[i * pitch + k] * nn[i];
            J[size_t(P.nobs + c) * kMap + k] = s;
        }
    }
    for (int it = threadIdx.x; it < 3 * (P.ma - nc); it += kWG) {
        const int c = it / (P.ma - nc), k = nc + it - c * (P.ma - nc);
        J[size_t(P.nobs + c) * kMap + k] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// mrqcof (period_search/mrqcof.c:16-148) for the parameters in sh.a: writes alpha[mfit][mfit]
// (full, symmetric) and beta[mfit], returns chi2.  raw_y / raw_dyda receive the raw model rows.
// MIX: the normal-equation contraction runs on FP32 MFMA with FP32 accumulation; chi2 (the LM
// acceptance test) is summed in FP64 from the FP64 residuals.
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool LDSB, bool MIX>
__device__ double mrqcof(const DevProblem &P, double *J, double *ytemp, double *dave, double *alpha, double *beta,
                         Shared &sh, double *raw_y, double *raw_dyda) {
    using Pr = Prec<MIX>;
    using T = typename Pr::T;
    using V4 = typename Pr::V4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int pl = lane & 15, kg = lane >> 4;
    const int ma = P.ma, npts = P.npts, mfit = P.mfit;
    curv_lds(P, sh.a, sh);
    blmatrix_lds(sh.a[P.nc], sh.a[P.nc + 1], sh);
    if (LDSB) {
        if constexpr (MIX)
            for (int i = tid; i < P.nf * P.dpitch; i += kWG) sh.u.dsphf[i] = P.dsphf[i];
        else
            for (int i = tid; i < P.nf * P.dpitch; i += kWG) sh.u.dsph[i] = P.dsph[i];
    }
    if constexpr (MIX)
        for (int i = tid; i < P.nf; i += kWG) sh.areaf[i] = float(sh.area[i]);
    __syncthreads();
    model_rows<NT, R, LDSB, MIX>(P, sh, J, ytemp, raw_y != nullptr);
    __syncthreads();
    // Phase C: column sums of the raw rows of each relative lightcurve, in point order
    for (int it = tid; it < P.nlc * (ma + 1); it += kWG) {
        const int L = it / (ma + 1), l = it - L * (ma + 1);
        if (!P.lc_inrel[L]) continue;
        double s = 0;
        for (int p = P.lc_start[L]; p < P.lc_start[L + 1]; ++p) s += (l < ma) ? J[size_t(p) * kMap + l] : ytemp[p];
        dave[it] = s;
    }
    __syncthreads();
    // Phase D: alpha/beta/chi2 from the normalised rows (mrqcof.c:89-145)
    const int nb = (mfit + 16) >> 4;  // 16-column blocks covering mfit + 1 columns
    V4 acc[10];
#pragma unroll
    for (int b = 0; b < 10; ++b) acc[b] = Pr::zero();
    T *st = reinterpret_cast<T *>(sh.u.d.stage[wave]);
    double chi = 0;  // MIX: this lane's share of sum w*r^2 (FP64)
    const int nsteps = (npts + 3) >> 2;
    for (int s = wave; s < nsteps; s += kWaves) {
        const int p = 4 * s + kg;
        double wq = 0, coef = 1, yt = 0, ave = 1;
        const double *dv = nullptr;
        bool rel = false;
        if (p < npts) {
            wq = P.s2i[p] * P.w[p];
            yt = ytemp[p];
            if (p < P.nobs) {
                const int L = P.lc_of[p];
                if (P.lc_inrel[L]) {
                    rel = true;
                    dv = dave + size_t(L) * (ma + 1);
                    ave = dv[ma];
                    coef = P.sig[p] * (P.lc_start[L + 1] - P.lc_start[L]) / ave;
                }
            }
        }
        const double *jr = J + size_t(p < npts ? p : 0) * kMap;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = pl + 16 * u;
            double v = 0;
            if (p < npts) {
                if (c < mfit) {
                    const int col = P.fit_idx[c];
                    v = jr[col];
                    if (rel) v = (col == 0) ? 0.0 : coef * (v - yt * dv[col] / ave);
                } else if (c == mfit) {
                    v = P.y[p] - (rel ? coef * yt : yt);
                    if (MIX) chi += wq * v * v;
                }
            }
            st[kg * 64 + c] = T(v);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        T vb[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) vb[b] = (b < nb) ? st[kg * 64 + 16 * b + pl] : T(0);
        const T wt = T(wq);
        int idx = 0;
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj <= bi; ++bj, ++idx)
                if (bi < nb) acc[idx] = Pr::mma(vb[bi], wt * vb[bj], acc[idx]);
        __builtin_amdgcn_wave_barrier();
    }
    if (MIX) {
        for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off);
        if (lane == 0) sh.chi_part[wave] = chi;
    }
    __syncthreads();
    // deterministic cross-wave sum into u.d.red (lower blocks; diagonal blocks in full)
    for (int w = 0; w < kWaves; ++w) {
        if (wave == w) {
            int idx = 0;
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj <= bi; ++bj, ++idx)
                    if (bi < nb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int o = (16 * bi + Pr::row(r, kg)) * 64 + 16 * bj + pl;
                            const double v = double(acc[idx][r]);
                            sh.u.d.red[o] = (w == 0) ? v : sh.u.d.red[o] + v;
                        }
        }
        __syncthreads();
    }
    for (int it = tid; it < mfit * mfit; it += kWG) {
        const int j = it / mfit, k = it - j * mfit;
        alpha[it] = (j >= k) ? sh.u.d.red[j * 64 + k] : sh.u.d.red[k * 64 + j];
    }
    for (int j = tid; j < mfit; j += kWG) beta[j] = sh.u.d.red[mfit * 64 + j];
    double chisq;
    if (MIX) {
        chisq = 0;
        for (int w = 0; w < kWaves; ++w) chisq += sh.chi_part[w];
    } else {
        chisq = sh.u.d.red[mfit * 64 + mfit];
    }
    if (raw_y) {
        for (int p = tid; p < npts; p += kWG) raw_y[p] = ytemp[p];
        for (int it = tid; it < npts * ma; it += kWG) {
            const int p = it / ma, k = it - p * ma;
            raw_dyda[it] = J[size_t(p) * kMap + k];
        }
    }
    __syncthreads();
    return chisq;
}

// ---------------------------------------------------------------------------------------------
// Full-pivot Gauss-Jordan on sh.u.cov (n x n) with right-hand side sh.da (gauss_errc.c:9-73).
// Pivot: the largest |a| over unpivoted rows/columns, ties to the last one in row-major order
// (the reference's `>=` scan).  Returns 0, 1 (no valid pivot: all candidates NaN) or 2 (zero pivot).
// The inverse's column unscramble is skipped: only the solution is used (mrqmin.c:61-79).
// ---------------------------------------------------------------------------------------------
__device__ inline int gauss_jordan(Shared &sh, int n) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *a = sh.u.g.cov, *b = sh.da;
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
            for (int w = 1; w < kWaves; ++w)
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
        if (tid == 0) {
            sh.indxr[i] = irow;
            sh.indxc[i] = icol;
        }
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

// One step of iterative refinement of the Gauss-Jordan solution x (in sh.da) of A x = b, with A and b
// kept in sh.u.g.cov0 / sh.b0: unscramble the inverse's columns (gauss_errc.c:63-67), r = b - A x in
// FP64, x += A^-1 r.
__device__ inline void refine_solution(Shared &sh, int n) {
    const int tid = threadIdx.x;
    double *inv = sh.u.g.cov, *A = sh.u.g.cov0;
    for (int l = n - 1; l >= 0; --l) {
        const int r0 = sh.indxr[l], c0 = sh.indxc[l];
        if (r0 != c0)
            for (int k = tid; k < n; k += kWG) {
                const double t = inv[k * n + r0];
                inv[k * n + r0] = inv[k * n + c0];
                inv[k * n + c0] = t;
            }
        __syncthreads();
    }
    for (int j = tid; j < n; j += kWG) {
        double s = sh.b0[j];
        for (int k = 0; k < n; ++k) s -= A[j * n + k] * sh.da[k];
        sh.col[j] = s;
    }
    __syncthreads();
    double dx = 0;
    if (tid < n)
        for (int k = 0; k < n; ++k) dx += inv[tid * n + k] * sh.col[k];
    __syncthreads();
    if (tid < n) sh.da[tid] += dx;
    __syncthreads();
}

__device__ inline void load_facets(const DevProblem &P, Shared &sh) {
    for (int i = threadIdx.x; i < P.nf; i += kWG) {
        sh.nx[i] = P.nx[i];
        sh.ny[i] = P.ny[i];
        sh.nz[i] = P.nz[i];
        sh.nxf[i] = float(P.nx[i]);
        sh.nyf[i] = float(P.ny[i]);
        sh.nzf[i] = float(P.nz[i]);
    }
}

// ---------------------------------------------------------------------------------------------
// One (period, pole) fit per workgroup iteration: period_search_BOINC.cpp:819-911 with mrqmin.c.
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool LDSB, bool MIX>
__global__ __launch_bounds__(kWG) void fit_kernel(DevProblem P, Slots S, const double *freq, const int *pole, int nfits,
                                                  double conw_r, ps_fit_result *out) {
    __shared__ Shared sh;
    const int tid = threadIdx.x;
    const int nc = P.nc, ma = P.ma, mfit = P.mfit;
    double *J = S.J + S.j_stride * blockIdx.x;
    double *ytemp = S.ytemp + S.y_stride * blockIdx.x;
    double *dave = S.dave + S.dave_stride * blockIdx.x;
    double *const alpha0 = S.alpha + S.a_stride * (2 * blockIdx.x), *const alpha1 = alpha0 + S.a_stride;
    double *const beta0 = S.beta + S.b_stride * (2 * blockIdx.x), *const beta1 = beta0 + S.b_stride;
    load_facets(P, sh);

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
        int cur = 0;
        double alamda = -1, ochisq = 0, chisq = 0, rchisq = -1;
        double iter_diff = 1e40, dev_old = 1e30, dev_new = 0;
        double chck0 = 0, chck1 = 0, chck2 = 0;
        int niter = 0, nfail = 0;
        while (((niter < P.n_iter_max) && (iter_diff > P.iter_diff_max)) || (niter < P.n_iter_min)) {
            if (alamda < 0) {
                alamda = P.alamda_start;
                if (tid < ma) sh.a[tid] = sh.cg[tid];
                __syncthreads();
                ochisq = mrqcof<NT, R, LDSB, MIX>(P, J, ytemp, dave, cur ? alpha1 : alpha0, cur ? beta1 : beta0, sh, nullptr,
                                             nullptr);
                if (tid < ma) sh.atry[tid] = sh.cg[tid];
            }
            for (int it = tid; it < mfit * mfit; it += kWG) {
                const int j = it / mfit, k = it - j * mfit;
                const double v = (cur ? alpha1 : alpha0)[it];
                const double d = (j == k) ? v * (1 + alamda) : v;
                sh.u.g.cov[it] = d;
                if (MIX) sh.u.g.cov0[it] = d;
            }
            for (int j = tid; j < mfit; j += kWG) {
                sh.da[j] = (cur ? beta1 : beta0)[j];
                if (MIX) sh.b0[j] = sh.da[j];
            }
            __syncthreads();
            const int err = gauss_jordan(sh, mfit);
            __syncthreads();
            if (MIX && err == 0) refine_solution(sh, mfit);
            if (err == 0) {
                if (tid < mfit) {
                    const int l = P.fit_idx[tid];
                    sh.atry[l] = sh.cg[l] + sh.da[tid];
                }
                __syncthreads();
                if (tid < ma) sh.a[tid] = sh.atry[tid];
                __syncthreads();
                chisq = mrqcof<NT, R, LDSB, MIX>(P, J, ytemp, dave, cur ? alpha0 : alpha1, cur ? beta0 : beta1, sh, nullptr,
                                            nullptr);
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
            if (niter == 1 || chisq < ochisq) {
                ochisq = chisq;
                if (tid < ma) sh.a[tid] = sh.cg[tid];
                __syncthreads();
                curv_lds(P, sh.a, sh);
                __syncthreads();
                if (tid < 3) {
                    const double *nn = (tid == 0) ? sh.nx : (tid == 1) ? sh.ny : sh.nz;
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
template <int NT, int R, bool LDSB, bool MIX>
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
    load_facets(P, sh);
    for (int v = blockIdx.x; v < nvec; v += gridDim.x) {
        if (tid < P.ma) sh.a[tid] = avec[size_t(v) * P.ma + tid];
        __syncthreads();
        const double c = mrqcof<NT, R, LDSB, MIX>(P, J, ytemp, dave, al, be, sh,
                                             raw_y ? raw_y + size_t(v) * P.npts : nullptr,
                                             raw_dyda ? raw_dyda + size_t(v) * P.npts * P.ma : nullptr);
        for (int it = tid; it < mfit * mfit; it += kWG)
            if (alpha) alpha[size_t(v) * mfit * mfit + it] = al[it];
        for (int it = tid; it < mfit; it += kWG)
            if (beta) beta[size_t(v) * mfit + it] = be[it];
        if (tid == 0 && chisq) chisq[v] = c;
        __syncthreads();
    }
}

}  // namespace psk
