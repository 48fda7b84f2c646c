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
constexpr int kRes = kMap - 1;  // J column holding the residual after phase C (ma <= 63)
constexpr int kMaxFac = 512;    // nrows <= 8
constexpr int kTP = 2;          // 16-point tiles per wave per pass of the facet loop
constexpr int kDsphLds = 288 * 49;  // FP64 Dsph held in LDS (facets x pitch: nrows <= 6, Lmax <= 6), else read from L2
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
    const double *t, *ex, *ey, *ez, *sx, *sy, *sz, *y, *sig;
    const double *wq;     // weight/sig^2 of each point
    const double *alpha;  // phase angle of each point, acos(ee . ee0)
    const int *lc_start, *lc_inrel, *lc_of;
    // phase B tiles (16 observed points each) split into runs of one lightcurve ("segments"):
    // tile t has segments seg_first[t] .. seg_first[t+1]-1, segment s covers tile-relative points
    // seg_range[s] & 255 .. (seg_range[s] >> 8) - 1 of lightcurve seg_lc[s]; lightcurve L owns the
    // consecutive segments lc_seg[L] .. lc_seg[L+1]-1
    int nseg;
    const int *seg_first, *seg_lc, *seg_range, *lc_seg;
    const double *nx, *ny, *nz, *darea;
    const double *dsph;  // [nf][dpitch], zero-padded beyond nc
    const float *dsphf;  // the same in FP32 (mixed precision)
    const double *pleg, *fc, *fs;
    const int *pair_m;
    const int *fit_idx;
    const double *cg_first;
    const double *pole_beta, *pole_lambda;
    // Pass schedule of the J-free kernels (ps_fit2.h).  There the point arrays above are in schedule
    // order: positions grouped into passes of at most kPassTiles 16-point tiles, every relative
    // lightcurve inside one pass, the three convexity rows at positions conv_q..conv_q+2 of the last
    // pass, padding positions all zero.  Pass k covers tiles pass_tile0[k] .. pass_tile0[k+1]-1;
    // its relative lightcurves are units pass_unit0[k] .. pass_unit0[k+1]-1, unit u covering
    // pass-relative rows unit_rows[u] & 255 .. (unit_rows[u] >> 8) - 1 of lightcurve unit_lc[u].
    int npass, nsched, conv_q, c00col;
    const int *pass_tile0, *pass_unit0, *unit_rows, *unit_lc;
    const int *qrel;   // [nsched] 1: the position holds a point of a relative lightcurve
    const int *qperm;  // [nsched] the point's index in input order (-1: padding)
    const int *fpos;   // [ma] the row column of a parameter (-1: fixed); column mfit = the residual
    const double *qsl; // [nsched] sig*Lpoints of a relative lightcurve's point (mrqcof.c:94), else 0
};

struct Slots {
    double *J, *ytemp, *dave, *alpha, *beta;
    size_t j_stride, y_stride, dave_stride, a_stride, b_stride;
};

struct Shared {
    double a[kMaxMa];     // parameters under evaluation
    double cg[kMaxMa];    // accepted parameters
    double atry[kMaxMa];
    double nx[kMaxFac], ny[kMaxFac], nz[kMaxFac];
    double area[kMaxFac], acl[kMaxFac], acls[kMaxFac];  // Area, Area*cl, Area*cls of the evaluation
    float nxf[kMaxFac], nyf[kMaxFac], nzf[kMaxFac];     // FP32 copies (mixed precision)
    float areaf[kMaxFac], aclf[kMaxFac], aclsf[kMaxFac];
    double da[64], col[64], b0[64];
    int indxr[64], indxc[64];  // Gauss-Jordan: the pivots' rows and columns (for the refinement)
    int ipiv[64];
    double scal[8];
    int iscal[4];
    int job;                // the queue kernel: the job claimed by the workgroup
    int next_tp;            // phase B: the next unclaimed tile pair
    double bl[9], dbl[18];  // pole rotation and its beta/lambda derivatives (blmatrix)
    double hbl[18];         // dbl_j * bl^T: the beta/lambda derivative generators in the pole frame
    union {
        double dsph[kDsphLds];       // phase B, FP64
        float dsphf[2 * kDsphLds];   // phase B, FP32
        struct {
            double cov[64 * 64];     // Gauss-Jordan: the damped matrix, then its inverse
            double cov0[64 * 64];    // the damped matrix kept for iterative refinement
            double cand[2][kWaves][64];  // each wave's pivot candidate row (two pivots in flight)
            double red_v[2][kWaves];     // each wave's candidate |a| and index
            int red_q[2][kWaves];
        } g;
        struct {
            double red[64 * 64];     // phase D: cross-wave sum of the normal equations
        } d;
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

// A load from a pointer held in the problem description, marked as global memory.  Those pointers
// are read from memory, so the compiler cannot infer their address space and would issue flat loads,
// which count against the LDS wait counter too: an LDS read after them waits for the global data.
template <class T>
__device__ inline T gload(const T *p, size_t i) {
    return reinterpret_cast<const __attribute__((address_space(1))) T *>(reinterpret_cast<unsigned long long>(p))[i];
}

// ---------------------------------------------------------------------------------------------
// curv (period_search/curv.c:10-39): g = exp(sum_{m,l} Pleg*(c*Fc + s*Fs)), Area = Darea*g, and the
// facet loop's Area*cl and Area*cls.
// ---------------------------------------------------------------------------------------------
// A wave-uniform read of read-only problem data through the scalar cache (s_load): unlike a vector
// load it does not wait behind the wave's outstanding vector stores.
__device__ inline int uload(const int *p, int i) {
    return reinterpret_cast<const __attribute__((address_space(4))) int *>(reinterpret_cast<unsigned long long>(p))[i];
}

// The exponent of facet i's area factor, sum over the (l, m) pairs of Pleg (a_n Fc_m + a_n+1 Fs_m) in
// the order of curv.c:19-31.  The facet's table entries are loaded four pairs ahead of the sums (the
// loads do not depend on them), so their latencies overlap instead of adding up pair by pair.
__device__ inline double facet_shape_sum(const DevProblem &P, const double *a, int i) {
    const int np = P.npairs;
    const double *pl = P.pleg + size_t(i) * np;
    const double *fc = P.fc + size_t(i) * P.mmax1, *fs = P.fs + size_t(i) * P.mmax1;
    double gsum = 0;
    int n = 0;
    for (int q0 = 0; q0 < np; q0 += 4) {
        double vp[4], vc[4], vs[4];
        int mm[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = min(q0 + j, np - 1);  // clamped: the loads past the last pair are not used
            mm[j] = uload(P.pair_m, q);
            vp[j] = gload(pl, q);
            vc[j] = gload(fc, mm[j]);
            vs[j] = gload(fs, mm[j]);  // m = 0 reads Fs_0, unused
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (q0 + j < np) {
                double fsum = a[n] * vc[j];
                ++n;
                if (mm[j] != 0) {
                    fsum = fsum + a[n] * vs[j];
                    ++n;
                }
                gsum = gsum + vp[j] * fsum;
            }
        }
    }
    return gsum;
}

__device__ inline void curv_lds(const DevProblem &P, const double *a, Shared &sh) {
    const double cl = exp(a[P.nc + 6]), cls = a[P.nc + 7];
    for (int i = threadIdx.x; i < P.nf; i += kWG) {
        const double ar = gload(P.darea, i) * exp(facet_shape_sum(P, a, i));
        sh.area[i] = ar;
        sh.acl[i] = ar * cl;
        sh.acls[i] = ar * cls;
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
template <class SH>
__device__ inline void blmatrix_lds(double bet, double lam, SH &sh) {
    if (threadIdx.x == 0) {
        double bl[3][3], dbl[2][3][3];
        blmatrix(bet, lam, bl, dbl);
        for (int i = 0; i < 9; ++i) sh.bl[i] = bl[i / 3][i % 3];
        for (int i = 0; i < 18; ++i) sh.dbl[i] = dbl[i / 9][(i / 3) % 3][i % 3];
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k)
                    sh.hbl[9 * j + 3 * i + k] = dbl[j][i][0] * bl[k][0] + dbl[j][i][1] * bl[k][1] + dbl[j][i][2] * bl[k][2];
    }
}

// Rotation phase of epoch t: an unfused FP64 multiply-add reduced modulo 2 pi exactly (matrix.c:18-19).
__device__ inline void rotation_phase(double omg, double t, double phi0, double &sf, double &cf) {
    double f = __dadd_rn(__dmul_rn(omg, t), phi0);
    f = fmod(f, 2 * kPi);
    sincos(f, &sf, &cf);
}

// Earth and Sun directions of point p in the body frame (bright.c:41-57 with matrix.c).
template <class SH>
__device__ inline void point_dirs(const DevProblem &P, const double *a, int p, const SH &sh, double e[3],
                                  double e0[3], double rot[3]) {
    const double ee[3] = {gload(P.ex, p), gload(P.ey, p), gload(P.ez, p)};
    const double ee0[3] = {gload(P.sx, p), gload(P.sy, p), gload(P.sz, p)};
    double sf, cf;
    rot[2] = gload(P.t, p);
    rotation_phase(a[P.nc + 2], rot[2], P.phi0, sf, cf);
    rot[0] = sf;
    rot[1] = cf;
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

// Phase function of a point and its derivatives w.r.t. the three phase parameters (phasec.c:11-24,
// bright.c:33-35).  The phase angle alpha = acos(ee . ee0) (bright.c:30-31) depends on the point only
// and is computed once per problem on the host.  The two terms that need exp and a division,
// exp(-alpha/p2) and p1*exp(-alpha/p2)*alpha/p2^2, are formed once per evaluation by phase_terms; the
// facet-loop epilogue only combines them.
__device__ inline void phase_terms(const DevProblem &P, const double *a, double *ephi, double *dph1, size_t pitch) {
    const double p1 = a[P.nc + 3], p2 = a[P.nc + 4];
    for (int p = threadIdx.x; p < P.nobs; p += kWG) {
        const double alpha = gload(P.alpha, p);
        const double e = exp(-alpha / p2);
        ephi[p] = e;
        dph1[size_t(p) * pitch] = p1 * e * alpha / (p2 * p2);
    }
}
__device__ inline double phase_scale(const double *a, int nc, double alpha, double ephi, double dph1, double dphp[3]) {
    dphp[0] = ephi;
    dphp[1] = dph1;
    dphp[2] = alpha;
    return 1 + a[nc + 3] * ephi + a[nc + 5] * alpha;
}

// Brightness derivatives w.r.t. beta, lambda, omega before the phase factor (bright.c:88-104 with
// matrix.c:11-58): js[j] = S.de_j + U.de0_j, S and U the facet sums of dbr/de and dbr/de0.  With the
// body-frame direction e = Rz(phi) B ee: de_j = Rz H_j Rz^T e for beta and lambda (H_j = dB_j B^T) and
// de_omega = t (e_y, -e_x, 0) (dRz/dphi Rz^T is that generator); rot = (sin phi, cos phi, t).
template <class SH>
__device__ inline void spin_columns(const SH &sh, const double e[3], const double e0[3], const double rot[3],
                                    const double S[3], const double U[3], double js[3]) {
    const double sf = rot[0], cf = rot[1], t = rot[2];
    const double ep[3] = {cf * e[0] - sf * e[1], sf * e[0] + cf * e[1], e[2]};  // Rz^T e
    const double e0p[3] = {cf * e0[0] - sf * e0[1], sf * e0[0] + cf * e0[1], e0[2]};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double *H = sh.hbl + 9 * j;
        double h[3], h0[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            h[i] = H[3 * i] * ep[0] + H[3 * i + 1] * ep[1] + H[3 * i + 2] * ep[2];
            h0[i] = H[3 * i] * e0p[0] + H[3 * i + 1] * e0p[1] + H[3 * i + 2] * e0p[2];
        }
        const double d[3] = {cf * h[0] + sf * h[1], -sf * h[0] + cf * h[1], h[2]};
        const double d0[3] = {cf * h0[0] + sf * h0[1], -sf * h0[0] + cf * h0[1], h0[2]};
        js[j] = (S[0] * d[0] + S[1] * d[1] + S[2] * d[2]) + (U[0] * d0[0] + U[1] * d0[1] + U[2] * d0[2]);
    }
    js[2] = t * ((S[0] * e[1] - S[1] * e[0]) + (U[0] * e0[1] - U[1] * e0[0]));
}

// Cross-lane sums on the VALU (no LDS round trip, no index registers).  v_permlane16/32_swap give each
// lane the pair (x[l & ~16], x[l | 16]) (resp. 32), added in that fixed order, so both lanes of a
// pair get the same bits; DPP quad permutes and row (half) mirrors pair lanes within a 16-lane row.
__device__ inline double pair_sum(double x, bool rows32) {
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = unsigned(b), hi = unsigned(b >> 32);
    unsigned a0, a1, h0, h1;
    if (rows32) {
        auto l2 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        auto h2 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        a0 = l2[0]; a1 = l2[1]; h0 = h2[0]; h1 = h2[1];
    } else {
        auto l2 = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        auto h2 = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        a0 = l2[0]; a1 = l2[1]; h0 = h2[0]; h1 = h2[1];
    }
    return __longlong_as_double((long long)((unsigned long long)h0 << 32 | a0)) +
           __longlong_as_double((long long)((unsigned long long)h1 << 32 | a1));
}
__device__ inline float pair_sum(float x, bool rows32) {
    const unsigned u = __float_as_uint(x);
    auto r = rows32 ? __builtin_amdgcn_permlane32_swap(u, u, false, false) : __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

template <class T>
__device__ inline T sum_groups(T v) {  // sum over the four 16-lane groups of a wave
    return pair_sum(pair_sum(v, false), true);
}

template <int CTRL>
__device__ inline double dpp(double x) {
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_mov_dpp(unsigned(b), CTRL, 0xF, 0xF, false);
    const unsigned hi = __builtin_amdgcn_mov_dpp(unsigned(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

__device__ inline double sum_row16(double x) {  // sum over the lane's 16-lane row
    x += dpp<0xB1>(x);   // quad_perm [1,0,3,2]: lane ^ 1
    x += dpp<0x4E>(x);   // quad_perm [2,3,0,1]: lane ^ 2
    x += dpp<0x141>(x);  // row_half_mirror: the other quad of the 8
    x += dpp<0x140>(x);  // row_mirror: the other 8 of the row
    return x;
}

// ---------------------------------------------------------------------------------------------
// Phase B: bright (period_search/bright.c:14-126) for every observed point, conv (conv.c:12-29) for
// the three convexity rows.  NT full 16-column MFMA tiles + R leftover coefficients (NT*16 + R >= nc).
// Each wave takes two 16-point tiles per pass: both share every facet and Dsph load and give the
// lane two independent dependency chains.  Visibility is a 0/1 factor, not a branch: the reciprocal
// is taken of max(mu + mu0, 1e-8) -- a no-op for visible facets (mu, mu0 > 1e-8) and, for the rest,
// small enough that (mu0/(mu+mu0))^2 <= 1e16 stays finite in FP32 -- so invisible facets add exact
// zeros.  With A = Area, Ac = Area*cl, As = Area*cls (per facet, in LDS):
//   X   = A*s            = Ac*mm + As*mm/(mu+mu0),       mm = vis*mu*mu0
//   ad  = A*ds/dmu       = vis*(As*(mu0/(mu+mu0))^2 + Ac*mu0),  ad0 likewise with mu
// (bright.c:69-76).  MIX: facet terms, rotation sums and the shape contraction in FP32; the
// body-frame directions and the brightness accumulation stay FP64.  LS: also the Lambert and
// Lommel-Seeliger columns, which only a fit of cl (the Lommel-Seeliger weight is always fixed,
// period_search_BOINC.cpp:570-572) and the raw rows of ps_eval need.
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool MIX>
struct TileSums {
    using Pr = Prec<MIX>;
    using T = typename Pr::T;
    typename Pr::V4 acc[NT];
    T racc[R > 0 ? R : 1];
    T S[3], U[3], d1, d2;
    double br;
};

// The next tile pair of phase B for a wave, first come first served.  The two waves that share a SIMD
// do not progress at the same rate (the older one issues first); with a static split the slower one
// finishes its last tiles alone, at well under half the pair's combined rate.
__device__ inline int next_tile_pair(Shared &sh, int lane) {
    int t = 0;
    if (lane == 0) t = atomicAdd(&sh.next_tp, 1);
    return __builtin_amdgcn_readfirstlane(t);
}

template <int NT, int R, bool LDSB, bool MIX, bool LS>
__device__ __forceinline__ void model_rows(const DevProblem &P, Shared &sh, double *J, double *ytemp, double *part) {
    using Pr = Prec<MIX>;
    using T = typename Pr::T;
    // the wave index, provably uniform: loops over it hold cross-lane shuffles
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pl = lane & 15, kg = lane >> 4;
    const int nc = P.nc, pitch = P.dpitch;
    // With leftover columns (R > 0) the layout is used for nc = 16 NT + R exactly, so every column of
    // it is a coefficient; with R == 0 (nc <= 16 NT) the columns from nc on are padding.
    auto is_coef = [&](int col) { return R > 0 || col < nc; };
    const double *a = sh.a;
    const T *dsph, *snx, *sny, *snz, *sar, *sac, *sas;
    if constexpr (MIX) {
        dsph = LDSB ? sh.u.dsphf : P.dsphf;
        snx = sh.nxf; sny = sh.nyf; snz = sh.nzf; sar = sh.areaf; sac = sh.aclf; sas = sh.aclsf;
    } else {
        dsph = LDSB ? sh.u.dsph : P.dsph;
        snx = sh.nx; sny = sh.ny; snz = sh.nz; sar = sh.area; sac = sh.acl; sas = sh.acls;
    }
    const double cl_d = exp(a[nc + 6]);
    const int ntiles = (P.nobs + 15) >> 4;
    const int nchunk = P.nf >> 2;  // nf = 8 nrows^2
    for (int tp = wave; kTP * tp < ntiles; tp = next_tile_pair(sh, lane)) {
        T e[kTP][3], e0[kTP][3];
        double rot[kTP][3];                 // sin phi, cos phi, t: all the epilogue needs of the point
        double ed[kTP][MIX ? 3 : 1], e0d[kTP][MIX ? 3 : 1];  // FP64 directions when e, e0 are FP32
        bool valid[kTP];
        TileSums<NT, R, MIX> ts[kTP];
#pragma unroll
        for (int h = 0; h < kTP; ++h) {
            const int p = (kTP * tp + h) * 16 + pl;
            valid[h] = p < P.nobs;
            double e64[3], e064[3];
            point_dirs(P, a, valid[h] ? p : 0, sh, e64, e064, rot[h]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[h][k] = T(e64[k]);
                e0[h][k] = T(e064[k]);
                if constexpr (MIX) {
                    ed[h][k] = e64[k];
                    e0d[h][k] = e064[k];
                }
                ts[h].S[k] = 0;
                ts[h].U[k] = 0;
            }
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) ts[h].acc[tt] = Pr::zero();
#pragma unroll
            for (int r = 0; r < R; ++r) ts[h].racc[r] = 0;
            ts[h].d1 = ts[h].d2 = 0;
            ts[h].br = 0;
        }
        for (int c = 0; c < nchunk; ++c) {
            const int j = 4 * c + kg;
            const T n0 = snx[j], n1 = sny[j], n2 = snz[j], Ac = sac[j], As = sas[j];
            const T A = LS ? sar[j] : T(0);
            T B[NT], Br[R > 0 ? R : 1];
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) B[tt] = LDSB ? dsph[j * pitch + 16 * tt + pl] : gload(dsph, j * pitch + 16 * tt + pl);
#pragma unroll
            for (int r = 0; r < R; ++r) Br[r] = LDSB ? dsph[j * pitch + 16 * NT + r] : gload(dsph, j * pitch + 16 * NT + r);
#pragma unroll
            for (int h = 0; h < kTP; ++h) {
                TileSums<NT, R, MIX> &s = ts[h];
                const T mu = e[h][0] * n0 + e[h][1] * n1 + e[h][2] * n2;
                const T mu0 = e0[h][0] * n0 + e0[h][1] * n1 + e0[h][2] * n2;
                const T vis = (valid[h] && mu > T(kTiny) && mu0 > T(kTiny)) ? T(1) : T(0);
                const T rd = Pr::rcp(fmax(mu + mu0, T(kTiny)));
                const T muv = vis * mu, mu0v = vis * mu0;
                const T mm = muv * mu0;              // mu*mu0 of a visible facet, else 0
                const T X = mm * fma(As, rd, Ac);    // Area*s = Area*mu*mu0*(cl + cls/(mu+mu0))
                s.br += double(X);
                if constexpr (LS) {
                    s.d1 += A * mm;
                    s.d2 += A * (mm * rd);
                }
                // Area*dsmu = Area*(cls*(mu0/dnom)^2 + cl*mu0) = mu0*(As*mu0/dnom^2 + Ac), dsmu0 likewise
                const T h2 = As * (rd * rd);
                const T ad = mu0v * fma(h2, mu0, Ac);
                const T ad0 = muv * fma(h2, mu, Ac);
                s.S[0] += ad * n0; s.S[1] += ad * n1; s.S[2] += ad * n2;
                s.U[0] += ad0 * n0; s.U[1] += ad0 * n1; s.U[2] += ad0 * n2;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) s.acc[tt] = Pr::mma(X, B[tt], s.acc[tt]);
#pragma unroll
                for (int r = 0; r < R; ++r) s.racc[r] += X * Br[r];
            }
        }
        // Epilogue.  The phase factors of both tiles first, so that the MFMA accumulators (the shape
        // columns: register r of lane l holds point Pr::row(r, kg), coefficient 16tt + pl) are written
        // and released before the spin columns are formed.
        // the body-frame directions in FP64
        auto dir = [&](int h, double v[3], double v0[3]) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (MIX) {
                    v[k] = ed[h][k];
                    v0[k] = e0d[h][k];
                } else {
                    v[k] = e[h][k];
                    v0[k] = e0[h][k];
                }
            }
        };
        double scale[kTP], dphp[kTP][3];
#pragma unroll
        for (int h = 0; h < kTP; ++h) {
            const int p = valid[h] ? (kTP * tp + h) * 16 + pl : 0;
            scale[h] = phase_scale(a, nc, gload(P.alpha, p), gload(ytemp, p), gload(J, size_t(p) * kMap + kRes), dphp[h]);
        }
#pragma unroll
        for (int h = 0; h < kTP; ++h) {
            const int tile = kTP * tp + h;
            const int s0 = uload(P.seg_first, tile), s1 = uload(P.seg_first, tile + 1);
            double val[NT][4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = Pr::row(r, kg);
                const double sc = __shfl(scale[h], i);
                const int pi = tile * 16 + i;
                double *jr = J + size_t(pi < P.nobs ? pi : 0) * kMap;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) {
                    val[tt][r] = (pi < P.nobs) ? sc * double(ts[h].acc[tt][r]) : 0.0;
                    if (pi < P.nobs && is_coef(16 * tt + pl)) jr[16 * tt + pl] = val[tt][r];
                }
            }
            // per-lightcurve column sums of this tile's rows, one partial row per segment
            for (int sg = s0; sg < s1; ++sg) {
                if (!uload(P.lc_inrel, uload(P.seg_lc, sg))) continue;
                const int rg = uload(P.seg_range, sg), q0 = rg & 255, q1 = rg >> 8;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) {
                    double x = 0;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = Pr::row(r, kg);
                        x += (i >= q0 && i < q1) ? val[tt][r] : 0.0;
                    }
                    x = sum_groups(x);
                    if (kg == 0 && is_coef(16 * tt + pl)) part[size_t(sg) * kMap + 16 * tt + pl] = x;
                }
            }
        }
#pragma unroll
        for (int h = 0; h < kTP; ++h) {
            TileSums<NT, R, MIX> &s = ts[h];
            const int tile = kTP * tp + h;
            const int p = tile * 16 + pl;
            const double br = sum_groups(s.br);
            double S[3], U[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                S[k] = double(sum_groups(s.S[k]));
                U[k] = double(sum_groups(s.U[k]));
            }
            double d1 = 0, d2 = 0;  // the Lambert and Lommel-Seeliger columns, when asked for
            if constexpr (LS) {
                d1 = double(sum_groups(s.d1));
                d2 = double(sum_groups(s.d2));
            }
            T racc[R > 0 ? R : 1];
#pragma unroll
            for (int r = 0; r < R; ++r) racc[r] = sum_groups(s.racc[r]);
            // the remaining columns of point p (every lane group holds the same point-p values)
            constexpr int kOther = R + 9;  // R leftover shape columns, beta/lambda/omega, 3 phase, cl, cls, ymod
            double ov[kOther];
            {
                double js[3], v[3], v0[3];
                dir(h, v, v0);
                spin_columns(sh, v, v0, rot[h], S, U, js);
#pragma unroll
                for (int r = 0; r < R; ++r) ov[r] = scale[h] * double(racc[r]);
#pragma unroll
                for (int k = 0; k < 3; ++k) ov[R + k] = scale[h] * js[k];
#pragma unroll
                for (int i = 0; i < 3; ++i) ov[R + 3 + i] = br * dphp[h][i];
                ov[R + 6] = LS ? scale[h] * d1 * cl_d : 0.0;
                ov[R + 7] = LS ? scale[h] * d2 : 0.0;
                ov[R + 8] = br * scale[h];
            }
            auto ocol = [&](int k) { return (k < R) ? 16 * NT + k : (k < R + 8) ? nc + (k - R) : P.ma; };
            if (kg == 0 && valid[h]) {
                double *jr = J + size_t(p) * kMap;
#pragma unroll
                for (int k = 0; k < kOther - 1; ++k)
                    if (k >= R || is_coef(16 * NT + k)) jr[ocol(k)] = ov[k];
                ytemp[p] = ov[kOther - 1];
            }
            // per-segment column sums: the four 16-lane rows hold the same values, so row kg sums
            // column 4j + kg, four columns per reduction
            for (int sg = uload(P.seg_first, tile); sg < uload(P.seg_first, tile + 1); ++sg) {
                if (!uload(P.lc_inrel, uload(P.seg_lc, sg))) continue;
                const int rg = uload(P.seg_range, sg), q0 = rg & 255, q1 = rg >> 8;
                const bool in = valid[h] && pl >= q0 && pl < q1;
#pragma unroll
                for (int j = 0; 4 * j < kOther; ++j) {
                    double v = 0;
#pragma unroll
                    for (int q = 0; q < 4 && 4 * j + q < kOther; ++q) v = (kg == q) ? ov[4 * j + q] : v;
                    const double x = sum_row16(in ? v : 0.0);
                    const int k = 4 * j + kg;
                    if (pl == 0 && k < kOther && (k >= R || is_coef(16 * NT + k))) part[size_t(sg) * kMap + ocol(k)] = x;
                }
            }
        }
    }
    // convexity rows: res_c = sum Area*n_c, dres_k = sum Area*Dsph_k*n_c (conv.c:12-29; Area = Darea*g),
    // FP64, on the last four waves: lane (k, q) sums the three components for coefficient k (k == nc:
    // the plain sums of Area*n) over facets q, q+4, ..., and the four quarters are added across the lane
    // quad.  Dsph comes from LDS where its FP64 copy is resident.
    constexpr int kConvWaves = 4;  // 16 coefficients per wave: nc <= 63
    if (wave >= kWaves - kConvWaves) {
        const double *cds;
        if constexpr (LDSB && !MIX)
            cds = sh.u.dsph;
        else
            cds = P.dsph;
        const int q = lane & 3, k = (wave - (kWaves - kConvWaves)) * 16 + (lane >> 2);
        double s0 = 0, s1 = 0, s2 = 0;
        if (k <= nc) {
#pragma unroll 4
            for (int i = q; i < P.nf; i += 4) {
                const double ds = (LDSB && !MIX) ? cds[i * pitch + k] : gload(cds, i * pitch + k);
                const double t = sh.area[i] * (k < nc ? ds : 1.0);
                s0 += t * sh.nx[i];
                s1 += t * sh.ny[i];
                s2 += t * sh.nz[i];
            }
        }
        s0 += dpp<0xB1>(s0);  // lane ^ 1
        s1 += dpp<0xB1>(s1);
        s2 += dpp<0xB1>(s2);
        s0 += dpp<0x4E>(s0);  // lane ^ 2
        s1 += dpp<0x4E>(s1);
        s2 += dpp<0x4E>(s2);
        if (q == 0 && k < nc) {
            J[size_t(P.nobs) * kMap + k] = s0;
            J[size_t(P.nobs + 1) * kMap + k] = s1;
            J[size_t(P.nobs + 2) * kMap + k] = s2;
        } else if (q == 0 && k == nc) {
            ytemp[P.nobs] = s0;
            ytemp[P.nobs + 1] = s1;
            ytemp[P.nobs + 2] = s2;
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
// The normal equations are accumulated in FP64 in both precisions: in MIX the Jacobian rows carry
// the rounding of the FP32 facet loop, but J^T W J is summed exactly enough to stay the Gram matrix of
// those rows.  (Accumulated in FP32 it is not, and on these systems -- condition numbers 1e8-1e10
// after damping -- the LM steps along the ill-conditioned directions become noise.)
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool LDSB, bool MIX, bool LS>
__device__ __forceinline__ double mrqcof(const DevProblem &P, double *J, double *ytemp, double *dave, double *alpha, double *beta,
                         Shared &sh, double *raw_y, double *raw_dyda) {
    using Pd = Prec<false>;  // phase D arithmetic
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int pl = lane & 15, kg = lane >> 4;
    const int ma = P.ma, npts = P.npts, mfit = P.mfit;
    curv_lds(P, sh.a, sh);
    blmatrix_lds(sh.a[P.nc], sh.a[P.nc + 1], sh);
    // exp(-alpha/p2) into ytemp (phase B replaces it by the model brightness of the point) and the p2
    // derivative of the phase function into the residual column of J (free until phase D)
    phase_terms(P, sh.a, ytemp, J + kRes, kMap);
    if (tid == 0) sh.next_tp = kWaves;  // tile pairs 0..7 go to waves 0..7
    if (LDSB) {
        // a loop of uniform trip count: its exit runs with every lane active (register spills placed
        // at the exit of a divergent loop run before the lanes are restored)
        const int n = P.nf * P.dpitch;
        for (int i0 = 0; i0 < n; i0 += kWG) {
            const int i = i0 + tid;
            if (i < n) {
                if constexpr (MIX)
                    sh.u.dsphf[i] = gload(P.dsphf, i);
                else
                    sh.u.dsph[i] = gload(P.dsph, i);
            }
        }
    }
    if constexpr (MIX)
        for (int i = tid; i < P.nf; i += kWG) {
            sh.areaf[i] = float(sh.area[i]);
            sh.aclf[i] = float(sh.acl[i]);
            sh.aclsf[i] = float(sh.acls[i]);
        }
    __syncthreads();
    model_rows<NT, R, LDSB, MIX, LS>(P, sh, J, ytemp, dave);
    __syncthreads();
    if (raw_y) {  // raw bright()/conv() rows, before the lightcurve normalisation below
        for (int p = tid; p < npts; p += kWG) raw_y[p] = ytemp[p];
        for (int it = tid; it < npts * ma; it += kWG) {
            const int p = it / ma, k = it - p * ma;
            raw_dyda[it] = J[size_t(p) * kMap + k];
        }
    }
    // Phase C: per relative lightcurve L the column sums dave_l = sum dyda_l and ave = sum ymod over
    // its points (mrqcof.c:53-87), added up from the per-segment partial rows of phase B, into an LDS
    // table in the union (Dsph is no longer needed): tab[L][l] = dave_l for l < ma, tab[L][ma] = ave,
    // cfl[L] = Lpoints / ave, rav[L] = 1 / ave and rell[L] = 1 for relative curves (ave, and so cfl,
    // can have either sign: a trial step may turn the phase function negative).
    const int tpitch = ma + 1;
    double *tab = sh.u.dsph, *cfl = sh.u.dsph + P.nlc * tpitch, *rav = cfl + P.nlc;
    int *rell = reinterpret_cast<int *>(rav + P.nlc);
    for (int it = tid; it < P.nlc * tpitch; it += kWG) {
        const int L = it / tpitch, l = it - L * tpitch;
        double s = 0;
        if (P.lc_inrel[L])
            for (int q = P.lc_seg[L]; q < P.lc_seg[L + 1]; ++q) s += dave[size_t(q) * kMap + (l < ma ? l : ma)];
        tab[it] = s;
    }
    __syncthreads();
    for (int L = tid; L < P.nlc; L += kWG) {
        const bool rel = P.lc_inrel[L];
        const double ave = tab[L * tpitch + ma];
        cfl[L] = rel ? double(P.lc_start[L + 1] - P.lc_start[L]) / ave : 0.0;
        rav[L] = rel ? 1 / ave : 0.0;
        rell[L] = rel;
    }
    __syncthreads();
    // Phase D: alpha/beta/chi2 (mrqcof.c:89-145).  Each wave takes 4-point K-steps; lane (pl, kg) holds
    // row 4s + kg, column 16u + pl of the normalised fitted columns plus the residual as column mfit,
    // which is at once the MFMA A operand (column x point) and, times w/sig^2, the B operand.  A row
    // of relative lightcurve L is normalised as it is loaded: coef*(dyda_l - (ymod*dave_l)/ave) with
    // coef = sig*Lpoints/ave, dyda for c00 = 0, residual y - coef*ymod; other rows keep dyda and have
    // residual y - ymod.  ymod*dave_l is formed first, as the reference does (mrqcof.c:99): where it
    // overflows (a runaway Lambert weight), the reference's normal equations hold inf and its
    // Gauss-Jordan fails, and so do these.
    const int nb = (mfit + 16) >> 4;  // 16-column blocks covering mfit + 1 columns
    int colu[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int cc = pl + 16 * u;
        colu[u] = (cc < mfit) ? P.fit_idx[cc] : (cc == mfit) ? kRes : -1;
    }
    Pd::V4 acc[10];
#pragma unroll
    for (int b = 0; b < 10; ++b) acc[b] = Pd::zero();
    const int nsteps = (npts + 3) >> 2;
    // Software pipeline over the K-steps: stage1 issues the global loads of a step (raw columns and the
    // point's weight, model value, observation and lightcurve), stage2 normalises it from the LDS table.
    struct RowG {
        double raw[4], wq, yt, sig, y;
        int L;
        bool in;
    };
    struct RowV {
        double v[4], wq;
    };
    // loads are unconditional (unused columns and rows past the end read column 0 / row 0) so that no
    // branch and no wait separates them from the LDS reads of the normalisation
    int coll[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) coll[u] = (u < nb && colu[u] >= 0) ? colu[u] : 0;
    auto stage1 = [&](int s, RowG &g) {
        const int p = 4 * s + kg;
        g.in = s < nsteps && p < npts;
        const size_t pp = g.in ? p : 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) g.raw[u] = gload(J, pp * kMap + coll[u]);
        g.wq = gload(P.wq, pp);
        g.yt = gload(ytemp, pp);
        g.sig = gload(P.sig, pp);
        g.y = gload(P.y, pp);
        g.L = gload(P.lc_of, pp);
    };
    // per-lane column kinds, fixed for the evaluation: 0 none, 1 residual, 2 c00 (zero in relative
    // curves), 3 any other; stage2 reads the table unconditionally (column 0 where unused) and selects,
    // so it has no branches
    int kind[4], colt[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int col = colu[u];
        kind[u] = !(u < nb && col >= 0) ? 0 : (col == kRes) ? 1 : (col == 0) ? 2 : 3;
        colt[u] = kind[u] == 3 ? col : 0;
    }
    auto stage2 = [&](const RowG &g, RowV &r) {
        // The table reads are issued together and pinned by an empty asm that takes their values: left
        // alone, the compiler moves each read into a branch over relative curves and waits for them one
        // at a time.
        int reli = rell[g.L];
        double cf = cfl[g.L], ra = rav[g.L], mv[4];
        const double *m = tab + g.L * tpitch;
#pragma unroll
        for (int u = 0; u < 4; ++u) mv[u] = m[colt[u]];
        asm volatile("" : "+v"(reli), "+v"(cf), "+v"(ra), "+v"(mv[0]), "+v"(mv[1]), "+v"(mv[2]), "+v"(mv[3]));
        const bool rel = reli != 0;
        const double coef = rel ? g.sig * cf : 1.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double xr = coef * (g.raw[u] - (g.yt * mv[u]) * ra);
            const double xo = rel ? ((kind[u] == 2) ? 0.0 : xr) : g.raw[u];
            const double x = (kind[u] == 1) ? g.y - coef * g.yt : xo;
            r.v[u] = (g.in && kind[u] != 0) ? x : 0.0;
        }
        r.wq = g.in ? g.wq : 0.0;
    };
    auto accumulate = [&](const RowV &r) {
        int idx = 0;
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj <= bi; ++bj, ++idx)  // blocks beyond nb add zeros
                acc[idx] = Pd::mma(r.v[bi], r.wq * r.v[bj], acc[idx]);
    };
    // Four load buffers, unrolled by four so that none is copied (a register copy of fresh load data
    // would wait for it): the loads of step t + 4 are issued while step t + 1 is normalised and step t
    // runs on the MFMA.  Steps past the end load row 0 and contribute zeros.
    constexpr int K = kWaves;
    RowG ga, gb, gc, gd;
    RowV cur, nxt;
    stage1(wave, ga);
    stage1(wave + K, gb);
    stage1(wave + 2 * K, gc);
    stage1(wave + 3 * K, gd);
    stage2(ga, cur);
    for (int s = wave; s < nsteps; s += 4 * K) {
        stage1(s + 4 * K, ga);
        stage2(gb, nxt);
        accumulate(cur);
        cur = nxt;
        stage1(s + 5 * K, gb);
        stage2(gc, nxt);
        accumulate(cur);
        cur = nxt;
        stage1(s + 6 * K, gc);
        stage2(gd, nxt);
        accumulate(cur);
        cur = nxt;
        stage1(s + 7 * K, gd);
        stage2(ga, nxt);
        accumulate(cur);
        cur = nxt;
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
                            const int o = (16 * bi + Pd::row(r, kg)) * 64 + 16 * bj + pl;
                            const double v = acc[idx][r];
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
    const double chisq = sh.u.d.red[mfit * 64 + mfit];
    __syncthreads();
    return chisq;
}

// ---------------------------------------------------------------------------------------------
// Full-pivot Gauss-Jordan on sh.u.cov (n x n, n <= 63) with right-hand side sh.da (gauss_errc.c:9-73).
// Pivot: the largest |a| over unpivoted rows and columns, ties to the last one in row-major order
// (the reference's `>=` scan); NaN entries are never chosen.  Returns 0, 1 (no valid pivot) or
// 2 (zero pivot).  Wave w keeps rows w, w + 8, ... in registers for the whole solve, lane l holding
// column l (column n = the right-hand side).  The reference swaps the pivot row into row icol
// (:38-42); here rows stay in place and carry a label, the index the row has in the reference's
// swapped matrix (the pivot row takes label icol, the row labelled icol takes label irow), so every
// row receives the reference's arithmetic and candidates are ordered as the reference scans them.
// Each wave publishes its best candidate with the row that holds it (double-buffered LDS slots): one
// barrier per pivot, after which every wave reads the pivot row from the winner's slot.  The
// pivot-column entry of a row is read from its lane; the search for the next pivot runs inside the
// elimination.  On success the inverse (columns not unscrambled, gauss_errc.c:63-67) and the solution
// are written back to sh.u.cov / sh.da by label.
// ---------------------------------------------------------------------------------------------
struct PivotCand {
    double v;  // |a|
    int q;     // row * n + column, -1 = none
};

// A lane visits its candidates in increasing index order (rows ascending, its column fixed), so
// "larger |a|, ties to the larger index" is |a| >= the best so far.
// (Selects, not branches: the choice is per lane, and a branch would switch the execution mask.)
__device__ inline void pivot_take(PivotCand &c, double s, int q) {
    const double v = fabs(s);
    const bool t = v >= c.v;
    c.v = t ? v : c.v;
    c.q = t ? q : c.q;
}

// c = the larger of c and o, if e
__device__ inline void pivot_merge(PivotCand &c, const PivotCand &o, bool e = true) {
    const bool t = e & ((o.v > c.v) | ((o.v == c.v) & (o.q > c.q)));
    c.v = t ? o.v : c.v;
    c.q = t ? o.q : c.q;
}

template <int CTRL>
__device__ inline PivotCand pivot_dpp(const PivotCand &c) {
    return {dpp<CTRL>(c.v), int(__builtin_amdgcn_mov_dpp(unsigned(c.q), CTRL, 0xF, 0xF, false))};
}

__device__ inline double swap_half(double x, bool rows32, int k) {  // element k of the permlane swap pair
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = unsigned(b), hi = unsigned(b >> 32);
    auto l2 = rows32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false) : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto h2 = rows32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false) : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __longlong_as_double((long long)((unsigned long long)h2[k] << 32 | l2[k]));
}

// The larger of lanes l and l ^ 16 (rows32: l ^ 32), the same bits in both lanes.
__device__ inline double pair_max(double x, bool rows32) {
    return fmax(swap_half(x, rows32, 0), swap_half(x, rows32, 1));
}
__device__ inline int pair_max(int x, bool rows32) {
    const unsigned u = unsigned(x);
    auto r = rows32 ? __builtin_amdgcn_permlane32_swap(u, u, false, false) : __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return max(int(r[0]), int(r[1]));
}
template <int CTRL>
__device__ inline int dpp_i(int x) {
    return int(__builtin_amdgcn_mov_dpp(unsigned(x), CTRL, 0xF, 0xF, false));
}

// Wave-wide argmax of the pivot candidates (a total order: larger |a|, then larger index), on DPP and
// permlane swaps; every lane ends with the same candidate.  Two passes: the largest |a|, then the
// largest index among the lanes that hold it.  A NaN |a| is never the maximum (fmax returns the other
// operand), as a NaN never passes gauss_errc's |a| >= big; all NaN leaves no candidate.
__device__ inline PivotCand pivot_wave_max(PivotCand c) {
    double m = c.v;
    m = fmax(m, dpp<0xB1>(m));
    m = fmax(m, dpp<0x4E>(m));
    m = fmax(m, dpp<0x141>(m));
    m = fmax(m, dpp<0x140>(m));
    m = pair_max(m, false);
    m = pair_max(m, true);
    int q = (c.v == m) ? c.q : -1;
    q = max(q, dpp_i<0xB1>(q));
    q = max(q, dpp_i<0x4E>(q));
    q = max(q, dpp_i<0x141>(q));
    q = max(q, dpp_i<0x140>(q));
    q = pair_max(q, false);
    q = pair_max(q, true);
    return {q < 0 ? -1.0 : m, q};  // no candidate (or only NaN): {-1, -1}, as a lane starts
}

// A double every lane holds (an LDS broadcast), made provably wave-uniform: the LM control flow that
// depends on it then branches on scalar conditions and keeps its state in scalar registers.
__device__ inline double uniform_double(double x) {
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(b))));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(b >> 32))));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ inline double read_lane(double x, int l) {  // x of lane l (l wave-uniform)
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(b), l), hi = __builtin_amdgcn_readlane(unsigned(b >> 32), l);
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
}

constexpr int kGJRows = (kMaxFit + kWaves - 1) / kWaves;  // rows held per wave

__device__ inline int gauss_jordan(Shared &sh, int n) {
    const int tid = threadIdx.x, l = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform row ownership
    n = __builtin_amdgcn_readfirstlane(n);
    double *a = sh.u.g.cov, *b = sh.da;
    double row[kGJRows];
    int lab[kGJRows];  // wave-uniform
    PivotCand c{-1.0, -1};
#pragma unroll
    for (int k = 0; k < kGJRows; ++k) {
        const int ll = wave + kWaves * k;
        lab[k] = ll;
        row[k] = (ll < n && l <= n) ? ((l < n) ? a[ll * n + l] : b[ll]) : 0.0;
        if (ll < n && l < n) pivot_take(c, row[k], ll * n + l);
    }
    unsigned long long done = 0;  // pivoted columns (= labels of pivot rows)
    for (int i = 0; i < n; ++i) {
        const int buf = i & 1;
        // workgroup argmax: each wave on the VALU, then every lane combines the eight wave results
        c = pivot_wave_max(c);
        const int wq = __builtin_amdgcn_readfirstlane(c.q);
        if (l == 0) {
            sh.u.g.red_v[buf][wave] = c.v;
            sh.u.g.red_q[buf][wave] = wq;
        }
        if (wq >= 0) {
            const int vr = wq / n;
            double v = 0;
#pragma unroll
            for (int k = 0; k < kGJRows; ++k)
                if (lab[k] == vr) v = row[k];
            sh.u.g.cand[buf][wave][l] = v;
        }
        __syncthreads();
        // the workgroup's candidate: a tree over the eight wave candidates (the order is total, so the
        // result does not depend on the grouping); candidate indices of different waves are distinct
        PivotCand cw[kWaves];
#pragma unroll
        for (int w = 0; w < kWaves; ++w) cw[w] = {sh.u.g.red_v[buf][w], sh.u.g.red_q[buf][w]};
#pragma unroll
        for (int st = 1; st < kWaves; st *= 2)
#pragma unroll
            for (int w = 0; w < kWaves; w += 2 * st) pivot_merge(cw[w], cw[w + st]);
        const int cq = __builtin_amdgcn_readfirstlane(cw[0].q);
        if (cq < 0) return 1;
        int ww = 0;
#pragma unroll
        for (int w = 1; w < kWaves; ++w)
            if (__builtin_amdgcn_readfirstlane(sh.u.g.red_q[buf][w]) == cq) ww = w;
        const int irow = cq / n, icol = cq - irow * n;
        if (tid == 0) {
            sh.indxr[i] = irow;
            sh.indxc[i] = icol;
        }
        done |= 1ull << icol;
        const double *prow = sh.u.g.cand[buf][ww];
        const double pv = prow[l];  // one read of the pivot row: the pivot is its lane icol
        const double s = read_lane(pv, icol);
        if (s == 0.0) return 2;  // the same in every wave
        const double pivinv = 1.0 / s;
        // the new pivot row (gauss_errc.c:44-53; lanes past n are not read)
        const double pr = ((l == icol) ? 1.0 : pv) * pivinv;
        // eliminate column icol from every other row (gauss_errc.c:55-62), searching the next pivot
        c = {-1.0, -1};
        const bool cand_col = l < n && !((done >> l) & 1);
        const bool piv_lane = l == icol;
#pragma unroll
        for (int k = 0; k < kGJRows; ++k) {
            const int lb = lab[k];
            if (lb >= n) continue;
            if (lb == irow) {
                row[k] = pr;
                lab[k] = icol;
                continue;
            }
            if (lb == icol) lab[k] = irow;                  // the row the swap moves (:38-42)
            const double cur = row[k];
            const double dum = read_lane(cur, icol);        // the pivot-column entry of the row
            const double v = (piv_lane ? 0.0 : cur) - pr * dum;
            row[k] = v;
            if (!((done >> lab[k]) & 1)) pivot_merge(c, {fabs(v), lab[k] * n + l}, cand_col);  // a scalar branch: the row test is wave-uniform
        }
    }
#pragma unroll
    for (int k = 0; k < kGJRows; ++k) {
        const int lb = lab[k];
        if (lb < n && l <= n) {
            if (l < n)
                a[lb * n + l] = row[k];
            else
                b[lb] = row[k];
        }
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
// One (period, pole) fit: period_search_BOINC.cpp:819-911 with mrqmin.c.  The slot's buffers: the
// Jacobian rows, the model values, the per-segment column sums, and alpha/beta double-buffered
// (accepted / trial).  sh.nx/ny/nz hold the problem's normals.
// ---------------------------------------------------------------------------------------------
struct SlotPtrsL {
    double *J, *ytemp, *dave, *alpha0, *alpha1, *beta0, *beta1;
};
__device__ inline SlotPtrsL slot_ptrs_legacy(const Slots &S, int slot) {
    SlotPtrsL p;
    p.J = S.J + S.j_stride * slot;
    p.ytemp = S.ytemp + S.y_stride * slot;
    p.dave = S.dave + S.dave_stride * slot;
    p.alpha0 = S.alpha + S.a_stride * (2 * size_t(slot));
    p.alpha1 = p.alpha0 + S.a_stride;
    p.beta0 = S.beta + S.b_stride * (2 * size_t(slot));
    p.beta1 = p.beta0 + S.b_stride;
    return p;
}

template <int NT, int R, bool LDSB, bool MIX, bool LS>
__device__ __forceinline__ void fit_one_legacy(const DevProblem &P, Shared &sh, const SlotPtrsL &sp, double freq,
                                               int pole, double conw_r, ps_fit_result *out) {
    const int tid = threadIdx.x;
    const int nc = P.nc, ma = P.ma, mfit = P.mfit;
    double *const J = sp.J, *const ytemp = sp.ytemp, *const dave = sp.dave;
    double *const alpha0 = sp.alpha0, *const alpha1 = sp.alpha1, *const beta0 = sp.beta0, *const beta1 = sp.beta1;
    // initial parameters (:819-844)
    if (tid < ma) {
        double v;
        const int m = pole - 1;
        if (tid < nc)
            v = P.cg_first[tid];
        else if (tid == nc)
            v = P.pole_beta[m];
        else if (tid == nc + 1)
            v = P.pole_lambda[m];
        else if (tid == nc + 2)
            v = 24 * 2 * kPi / (1 / freq);
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
    // the LM state is wave-uniform; its values meet at the joins of lane-predicated stores
    // (tid < ma), so settle() re-asserts it after each such region: the loop exit and the
    // accept / reject decisions then branch on scalars
    auto settle = [&] {
        alamda = uniform_double(alamda);
        ochisq = uniform_double(ochisq);
        chisq = uniform_double(chisq);
        rchisq = uniform_double(rchisq);
        iter_diff = uniform_double(iter_diff);
        dev_old = uniform_double(dev_old);
        dev_new = uniform_double(dev_new);
        chck0 = uniform_double(chck0);
        chck1 = uniform_double(chck1);
        chck2 = uniform_double(chck2);
        niter = __builtin_amdgcn_readfirstlane(niter);
        nfail = __builtin_amdgcn_readfirstlane(nfail);
        cur = __builtin_amdgcn_readfirstlane(cur);
    };
    for (;;) {
        settle();
        if (!(((niter < P.n_iter_max) && (iter_diff > P.iter_diff_max)) || (niter < P.n_iter_min))) break;
        // One mrqcof per pass (a single inlined copy): the first pass of a fit evaluates the
        // starting point (mrqmin.c:34-52) and goes straight on to the step; later passes solve the
        // damped system and evaluate the trial point (mrqmin.c:53-79).
        const bool init = alamda < 0;
        int err = 0;
        if (init) {
            alamda = P.alamda_start;
            if (tid < ma) sh.a[tid] = sh.cg[tid];
        } else {
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
            err = __builtin_amdgcn_readfirstlane(gauss_jordan(sh, mfit));
            __syncthreads();
            if (MIX && err == 0) refine_solution(sh, mfit);
            if (err == 0) {
                if (tid < mfit) {
                    const int l = P.fit_idx[tid];
                    sh.atry[l] = sh.cg[l] + sh.da[tid];
                }
                __syncthreads();
                if (tid < ma) sh.a[tid] = sh.atry[tid];
            }
        }
        if (err == 0) {
            __syncthreads();
            // the starting point fills the current buffers, a trial point the other ones
            const bool other = !init;
            const double chi = uniform_double(mrqcof<NT, R, LDSB, MIX, LS>(
                P, J, ytemp, dave, (cur ^ other) ? alpha1 : alpha0, (cur ^ other) ? beta1 : beta0, sh, nullptr, nullptr));
            if (init) {
                ochisq = chi;
                chisq = chi;  // a failed first step then reports this fit's own starting chi2
                if (tid < ma) sh.atry[tid] = sh.cg[tid];
            } else {
                chisq = chi;
                if (chisq < ochisq) {
                    alamda = alamda / P.alamda_incr;
                    cur ^= 1;
                    if (tid < ma) sh.cg[tid] = sh.atry[tid];
                } else {
                    alamda = P.alamda_incr * alamda;
                    chisq = ochisq;
                }
            }
            __syncthreads();
        } else {
            ++nfail;
        }
        settle();
        if (init) continue;  // the starting point is evaluated; the first step follows
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
            chck0 = uniform_double(sh.scal[1]);
            chck1 = uniform_double(sh.scal[2]);
            chck2 = uniform_double(sh.scal[3]);
            rchisq = chisq - (chck0 * chck0 + chck1 * chck1 + chck2 * chck2) * (conw_r * conw_r);
            __syncthreads();
        }
        settle();
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
        *out = r;
    }
    __syncthreads();
}

// The fits of one problem, a grid-stride loop over them.
template <int NT, int R, bool LDSB, bool MIX, bool LS>
__global__ __launch_bounds__(kWG) void fit_kernel(const DevProblem *__restrict__ Pp, Slots S, const double *freq,
                                                  const int *pole, int nfits, double conw_r, ps_fit_result *out) {
    // The problem description is read through a pointer (scalar loads on demand) rather than passed
    // by value: ~40 kernel-argument pointers held in SGPRs would spill into VGPRs.
    const DevProblem &P = *Pp;
    __shared__ Shared sh;
    const SlotPtrsL sp = slot_ptrs_legacy(S, blockIdx.x);
    load_facets(P, sh);
    for (int fi = blockIdx.x; fi < nfits; fi += gridDim.x)
        fit_one_legacy<NT, R, LDSB, MIX, LS>(P, sh, sp, freq[fi], pole[fi], conw_r, out + fi);
}

// mrqcof for externally supplied parameter vectors (ps_eval).
template <int NT, int R, bool LDSB, bool MIX>
__global__ __launch_bounds__(kWG) void eval_kernel(const DevProblem *__restrict__ Pp, Slots S, const double *avec,
                                                   int nvec, double *alpha, double *beta, double *chisq, double *raw_y,
                                                   double *raw_dyda) {
    const DevProblem &P = *Pp;
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
        const double c = mrqcof<NT, R, LDSB, MIX, true>(P, J, ytemp, dave, al, be, sh,
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
