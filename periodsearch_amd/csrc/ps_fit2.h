// ps_fit2.h -- the J-free fit and evaluation kernels (gfx950), included by ps_device.hip after
// ps_kernels.h.
//
// One 256-lane workgroup (4 wave64) owns one (period, pole) fit, and two such workgroups share a CU
// (LDS <= 80 KB and 256 VGPRs each): while one waits in the Gauss-Jordan pivot chain or at a
// barrier, the other's facet loop issues.
//
// The Jacobian is never written to memory.  An evaluation (mrqcof, period_search/mrqcof.c:16-148)
// walks the points in passes of at most 8 tiles of 16 points, built on the host so that every
// relative lightcurve lies inside one pass (the reference, too, holds one lightcurve's rows at a
// time: dytemp, mrqcof.c:23, 51-104).  Per pass:
//  B  model rows (bright.c:14-126, conv.c:12-29): each wave runs the MFMA facet loop of ps_kernels.h
//     over its two tiles and writes the fitted columns of the raw rows into LDS, plus the model value
//     of a relative point, or the residual y - ymod of a calibrated one;
//  C  for every relative lightcurve of the pass (one wave each): the column sums dave and ave, then
//     the normalisation of its rows in place (mrqcof.c:89-101) and the residual y - coef*ymod;
//  D  the weighted Gram matrix of the pass's normalised rows plus the residual column on
//     v_mfma_f64_16x16x4f64, each wave owning fixed 16x16 blocks of its lower triangle (at most
//     three), accumulated over all passes; between passes a wave parks its accumulators in a small
//     per-slot buffer.
// After the last pass the blocks are alpha (J^T W J), beta (J^T W r) and chi2.
#pragma once

#include "ps_kernels.h"

namespace psk {

constexpr int kWG2 = 256;
constexpr int kWaves2 = kWG2 / 64;
constexpr int kPassTiles = 8;                 // 16-point tiles per pass: two per wave
constexpr int kPassRows = 16 * kPassTiles;
constexpr int kJP = 57;                       // LDS row pitch (doubles): mfit + 1 <= 57 for Lmax <= 6
constexpr int kGJRows2 = (kMaxFit + kWaves2 - 1) / kWaves2;  // Gauss-Jordan rows per wave
constexpr int kOwn = 3;                       // Gram blocks owned per wave

template <int NFM>
struct Shared2 {
    double a[64];     // parameters under evaluation
    double cg[64];    // accepted parameters
    double atry[64];
    double nx[NFM], ny[NFM], nz[NFM], acl[NFM], acls[NFM];  // normals, Area*cl, Area*cls
    union {
        double da[64];       // the LM step: the Gauss-Jordan right-hand side, then its solution
        double csum[2][64];  // phase C: the column sums of two consecutive lightcurves
    };
    double bl[9], dbl[18], hbl[18];  // pole rotation, its derivatives, the pole-frame generators
    double scal[4];
    int job;  // the queue kernel: the job claimed by the workgroup
    union {
        double rows[kPassRows * kJP];  // the pass's Jacobian rows (fitted columns + residual)
        struct {
            double cov[64 * 64];             // Gauss-Jordan: the damped matrix
            double cand[2][kWaves2][64];     // each wave's pivot candidate row (two pivots in flight)
            double red_v[2][kWaves2];        // each wave's candidate |a| and index
            int red_q[2][kWaves2];
        } g;
    } u;
};

// Per-slot device buffers of the J-free kernels.
struct Slots2 {
    double *alpha, *beta;  // [slot][2][mfit*mfit], [slot][2][mfit]: accepted / trial normal equations
    double *stash;         // [slot][wave][kOwn][64 lanes][4]: Gram accumulators between passes
    double *area;          // [slot][nf]: Area of the last curv
    size_t a_stride, b_stride, f_stride;
};
constexpr size_t kStashPerSlot = size_t(kWaves2) * kOwn * 64 * 4;

// curv (curv.c:10-39): Area to the slot buffer, Area*cl and Area*cls to LDS.
template <int NFM>
__device__ inline void curv2(const DevProblem &P, const double *a, Shared2<NFM> &sh, double *area) {
    const double cl = exp(a[P.nc + 6]), cls = a[P.nc + 7];
    for (int i = threadIdx.x; i < P.nf; i += kWG2) {
        const double ar = gload(P.darea, i) * exp(facet_shape_sum(P, a, i));
        area[i] = ar;
        sh.acl[i] = ar * cl;
        sh.acls[i] = ar * cls;
    }
}

__device__ inline double wave_sum(double x) {  // sum over the 64 lanes, the same bits in every lane
    return sum_groups(sum_row16(x));
}

// The sum over the four 16-lane groups, complete in the lanes of group 0 only (the others hold partial
// sums), in the association of sum_groups, ((g0 + g1) + (g2 + g3)): the lanes are exchanged through the
// LDS crossbar (ds_bpermute), which leaves the VALU issue slots to the other waves on the SIMD.
__device__ inline double sum_groups_lo(double x) {
    const int l = threadIdx.x & 63;
    x = x + __shfl(x, (l + 16) & 63);
    return x + __shfl(x, (l + 32) & 63);
}

// sum_i Area_i * v_i over the facets (the convexity check, period_search_BOINC.cpp:875-880, and the
// total area), by wave 0: lane l adds facets l, l + 64, ..., then the lanes are added in a fixed tree.
__device__ inline double facet_sum(const double *area, const double *v, int nf) {
    const int l = threadIdx.x & 63;
    double s = 0;
    for (int i = l; i < nf; i += 64) s += gload(area, i) * (v ? v[i] : 1.0);
    return wave_sum(s);
}

// The three sums of the convexity check in one pass over the facets, each in facet_sum's order.
__device__ inline void facet_sums_n(const double *area, const double *nx, const double *ny, const double *nz, int nf,
                                    double s[3]) {
    const int l = threadIdx.x & 63;
    double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll 4
    for (int i = l; i < nf; i += 64) {
        const double ar = gload(area, i);
        s0 += ar * nx[i];
        s1 += ar * ny[i];
        s2 += ar * nz[i];
    }
    s[0] = wave_sum(s0);
    s[1] = wave_sum(s1);
    s[2] = wave_sum(s2);
}

// ---------------------------------------------------------------------------------------------
// Phase B of one pass: the model rows of the wave's two tiles (pass-relative tiles 2w, 2w+1) into
// sh.u.rows, and, in the last pass, the convexity rows.  The facet arithmetic is that of model_rows
// (ps_kernels.h), with Dsph read through L2, the two tiles' steps in lockstep, and the brightness taken
// from the MFMA's coefficient-0 column instead of a sum of its own.  raw_y / raw_dyda (ps_eval) receive
// the raw rows, all ma columns, in input order.
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool LS, int NFM>
__device__ __forceinline__ void model_rows2(const DevProblem &P, Shared2<NFM> &sh, int tile0, int ntl, int rot,
                                            bool last, const double *area, double *raw_y, double *raw_dyda) {
    using Pr = Prec<false>;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pl = lane & 15, kg = lane >> 4;
    const int nc = P.nc, pitch = P.dpitch, mfit = P.mfit, ma = P.ma;
    const double *a = sh.a;
    const double cl_d = exp(a[nc + 6]);
    const double p1 = a[nc + 3], p2 = a[nc + 4];
    const int nchunk = P.nf >> 2;
    // the wave's first tile, pass-relative: waves take tile pairs in an order that rotates with the
    // pass, so the short pair of a pass with an odd tile count falls on each SIMD in turn
    const int tl = 2 * ((wave + rot) & 3);
    // the row column of this lane's shape column of each MFMA tile; a coefficient that is not fitted
    // (or past nc) goes to column mfit, which the same wave rewrites afterwards with the model value
    // or residual of the point (its LDS writes are done in program order)
    int fcol[NT];
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
        const int k = 16 * tt + pl;
        const int c = (k < nc) ? P.fpos[k] : -1;
        fcol[tt] = c >= 0 ? c : mfit;
    }
    // the wave's tiles: a pair, or the last tile of a pass alone (its facet loop then carries one
    // tile, and the SIMD time of the missing one goes to the other workgroup on the CU)
    auto tiles = [&]<int KT>() {
        double e[KT][3], e0[KT][3], rot[KT][3], alp[KT], eph[KT], dph[KT];
        TileSums<NT, R, false> ts[KT];
#pragma unroll
        for (int h = 0; h < KT; ++h) {
            const int q = (tile0 + tl + h) * 16 + pl;
            point_dirs(P, a, q, sh, e[h], e0[h], rot[h]);
            alp[h] = gload(P.alpha, q);
            eph[h] = exp(-alp[h] / p2);
            dph[h] = p1 * eph[h] * alp[h] / (p2 * p2);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                ts[h].S[k] = 0;
                ts[h].U[k] = 0;
            }
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) ts[h].acc[tt] = Pr::zero();
#pragma unroll
            for (int r = 0; r < R; ++r) ts[h].racc[r] = 0;
            ts[h].d1 = ts[h].d2 = 0;
        }
        // Dsph through L2: the operands of chunk c + 1 are loaded while chunk c is computed (the table
        // has a zero chunk of facets past its end, so the last chunk's load needs no branch)
        const __attribute__((address_space(1))) double *dpl =
            reinterpret_cast<const __attribute__((address_space(1))) double *>(
                reinterpret_cast<unsigned long long>(P.dsph)) + kg * pitch + pl;
        const __attribute__((address_space(1))) double *dpr = dpl - pl + 16 * NT;
        const int dstep = 4 * pitch;
        constexpr int RP = R <= 1 ? R : 0;  // leftover columns prefetched as well
        double Bn[NT], Brn[RP > 0 ? RP : 1];
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) Bn[tt] = dpl[16 * tt];
#pragma unroll
        for (int r = 0; r < RP; ++r) Brn[r] = dpr[r];
#pragma unroll 2
        for (int c = 0; c < nchunk; ++c) {
            const int j = 4 * c + kg;
            const double n0 = sh.nx[j], n1 = sh.ny[j], n2 = sh.nz[j], Ac = sh.acl[j], As = sh.acls[j];
            const double A = LS ? gload(area, j) : 0.0;
            double B[NT], Br[R > 0 ? R : 1];
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) B[tt] = Bn[tt];
#pragma unroll
            for (int r = 0; r < R; ++r) Br[r] = (r < RP) ? Brn[r < RP ? r : 0] : dpr[r];
            dpl += dstep;
            dpr += dstep;
#pragma unroll
            for (int tt = 0; tt < NT; ++tt) Bn[tt] = dpl[16 * tt];
#pragma unroll
            for (int r = 0; r < RP; ++r) Brn[r] = dpr[r];
            // The tiles' dependency chains in lockstep, step by step (the same operations as one tile
            // at a time): a wave alone on its SIMD has two independent instructions to issue at every
            // step instead of waiting on each result.
            double mu[KT], mu0[KT], vis[KT], dn[KT], rd[KT], er[KT];
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                mu[h] = e[h][0] * n0 + e[h][1] * n1 + e[h][2] * n2;
                mu0[h] = e0[h][0] * n0 + e0[h][1] * n1 + e0[h][2] * n2;
            }
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                vis[h] = (mu[h] > kTiny && mu0[h] > kTiny) ? 1.0 : 0.0;
                dn[h] = fmax(mu[h] + mu0[h], kTiny);
                rd[h] = __builtin_amdgcn_rcp(dn[h]);  // Prec<false>::rcp, step by step
            }
#pragma unroll
            for (int it = 0; it < 2; ++it) {
#pragma unroll
                for (int h = 0; h < KT; ++h) er[h] = fma(-dn[h], rd[h], 1.0);
#pragma unroll
                for (int h = 0; h < KT; ++h) rd[h] = fma(rd[h], er[h], rd[h]);
            }
            double muv[KT], mu0v[KT], mm[KT], X[KT], h2[KT], ad[KT], ad0[KT];
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                muv[h] = vis[h] * mu[h];
                mu0v[h] = vis[h] * mu0[h];
            }
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                mm[h] = muv[h] * mu0[h];
                X[h] = mm[h] * fma(As, rd[h], Ac);
                h2[h] = As * (rd[h] * rd[h]);
            }
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                ad[h] = mu0v[h] * fma(h2[h], mu0[h], Ac);
                ad0[h] = muv[h] * fma(h2[h], mu[h], Ac);
            }
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                TileSums<NT, R, false> &s = ts[h];
                if constexpr (LS) {
                    s.d1 += A * mm[h];
                    s.d2 += A * (mm[h] * rd[h]);
                }
                s.S[0] += ad[h] * n0; s.S[1] += ad[h] * n1; s.S[2] += ad[h] * n2;
                s.U[0] += ad0[h] * n0; s.U[1] += ad0[h] * n1; s.U[2] += ad0[h] * n2;
            }
#pragma unroll
            for (int h = 0; h < KT; ++h) {
                TileSums<NT, R, false> &s = ts[h];
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) s.acc[tt] = Pr::mma(X[h], B[tt], s.acc[tt]);
#pragma unroll
                for (int r = 0; r < R; ++r) s.racc[r] += X[h] * Br[r];
            }
        }
        double scale[KT], dphp[KT][3];
#pragma unroll
        for (int h = 0; h < KT; ++h) scale[h] = phase_scale(a, nc, alp[h], eph[h], dph[h], dphp[h]);
        // shape columns: register r of lane l holds point row(r, kg), coefficient 16 tt + pl
#pragma unroll
        for (int h = 0; h < KT; ++h) {
            const int tile = tl + h;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = Pr::row(r, kg);
                const double sc = __shfl(scale[h], i);
                const int q = (tile0 + tile) * 16 + i;
                const int orig = raw_dyda ? gload(P.qperm, q) : -1;
#pragma unroll
                for (int tt = 0; tt < NT; ++tt) {
                    const double v = sc * ts[h].acc[tt][r];
                    sh.u.rows[(16 * tile + i) * kJP + fcol[tt]] = v;
                    if (orig >= 0 && 16 * tt + pl < nc) raw_dyda[size_t(orig) * ma + 16 * tt + pl] = v;
                }
            }
        }
        // the remaining columns of point pl (complete in the lanes of group 0, which write them)
#pragma unroll
        for (int h = 0; h < KT; ++h) {
            TileSums<NT, R, false> &s = ts[h];
            const int tile = tl + h;
            const int q = (tile0 + tile) * 16 + pl;
            // the brightness: the MFMA's coefficient-0 column (Dsph_0 = 1 on every facet), point 4r + g
            // in register r of lane 16g
            double b4[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) b4[r] = __shfl(s.acc[0][r], 16 * (pl & 3));
            const int rb = pl >> 2;
            const double br = rb == 0 ? b4[0] : rb == 1 ? b4[1] : rb == 2 ? b4[2] : b4[3];
            double S[3], U[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                S[k] = sum_groups_lo(s.S[k]);
                U[k] = sum_groups_lo(s.U[k]);
            }
            double d1 = 0, d2 = 0;
            if constexpr (LS) {
                d1 = sum_groups_lo(s.d1);
                d2 = sum_groups_lo(s.d2);
            }
            double racc[R > 0 ? R : 1];
#pragma unroll
            for (int r = 0; r < R; ++r) racc[r] = sum_groups_lo(s.racc[r]);
            constexpr int kOther = R + 9;  // R leftover shape columns, beta/lambda/omega, 3 phase, cl, cls, ymod
            double ov[kOther];
            {
                double js[3];
                spin_columns(sh, e[h], e0[h], rot[h], S, U, js);
#pragma unroll
                for (int r = 0; r < R; ++r) ov[r] = scale[h] * racc[r];
#pragma unroll
                for (int k = 0; k < 3; ++k) ov[R + k] = scale[h] * js[k];
#pragma unroll
                for (int i = 0; i < 3; ++i) ov[R + 3 + i] = br * dphp[h][i];
                ov[R + 6] = LS ? scale[h] * d1 * cl_d : 0.0;
                ov[R + 7] = LS ? scale[h] * d2 : 0.0;
                ov[R + 8] = br * scale[h];
            }
            auto ocol = [&](int k) { return (k < R) ? 16 * NT + k : nc + (k - R); };
            if (kg == 0) {
                double *row = sh.u.rows + (16 * tile + pl) * kJP;
#pragma unroll
                for (int k = 0; k < kOther - 1; ++k) {
                    const int c = uload(P.fpos, ocol(k));
                    if (c >= 0) row[c] = ov[k];
                }
                const double ymod = ov[kOther - 1];
                // relative point: the model value (phase C normalises it); calibrated: the residual
                row[mfit] = gload(P.qrel, q) ? ymod : gload(P.y, q) - ymod;
                if (raw_dyda) {
                    const int orig = gload(P.qperm, q);
                    if (orig >= 0) {
                        raw_y[orig] = ymod;
#pragma unroll
                        for (int k = 0; k < kOther - 1; ++k) raw_dyda[size_t(orig) * ma + ocol(k)] = ov[k];
                    }
                }
            }
        }
        };
    if (tl + 1 < ntl)
        tiles.template operator()<2>();
    else if (tl < ntl)
        tiles.template operator()<1>();
    if (last) {
        // convexity rows (conv.c:12-29): res_c = sum Area*n_c, dres_k = sum Area*Dsph_k*n_c; lane (k, q)
        // sums coefficient k (k == nc: the plain sums) over facets q, q+4, ..., the quarters are added
        // across the lane quad.  Their residual is 0 - res_c (y = 0).  The tile that holds them has
        // written model rows there too: every column is rewritten here, after all waves are done.
        __syncthreads();
        const int q4 = lane & 3, k = wave * 16 + (lane >> 2);
        double s0 = 0, s1 = 0, s2 = 0;
        if (k <= nc) {
#pragma unroll 4
            for (int i = q4; i < P.nf; i += 4) {
                const double t = gload(area, i) * (k < nc ? gload(P.dsph, i * pitch + k) : 1.0);
                s0 += t * sh.nx[i];
                s1 += t * sh.ny[i];
                s2 += t * sh.nz[i];
            }
        }
        s0 += dpp<0xB1>(s0);
        s1 += dpp<0xB1>(s1);
        s2 += dpp<0xB1>(s2);
        s0 += dpp<0x4E>(s0);
        s1 += dpp<0x4E>(s1);
        s2 += dpp<0x4E>(s2);
        const int r0 = P.conv_q - tile0 * 16;
        const double sv[3] = {s0, s1, s2};
        if (q4 == 0 && k <= nc) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double *row = sh.u.rows + (r0 + c) * kJP;
                if (k < nc) {
                    const int col = P.fpos[k];
                    if (col >= 0) row[col] = sv[c];
                    if (raw_dyda) raw_dyda[size_t(P.npts - 3 + c) * ma + k] = sv[c];
                } else {
                    row[mfit] = 0.0 - sv[c];
                    if (raw_dyda) raw_y[P.npts - 3 + c] = sv[c];
                }
            }
        }
        // the other parameters' columns of the convexity rows are zero
        for (int it = threadIdx.x; it < 3 * (ma - nc); it += kWG2) {
            const int c = it / (ma - nc), kk = nc + it - c * (ma - nc);
            const int col = P.fpos[kk];
            if (col >= 0) sh.u.rows[(r0 + c) * kJP + col] = 0.0;
            if (raw_dyda) raw_dyda[size_t(P.npts - 3 + c) * ma + kk] = 0.0;
        }
    }
}

// Per-row data of a pass for phases C and D, loaded once after phase B (the loads overlap the barrier):
// lane i holds pass rows i and 64 + i.  y: the observation, sl: sig*Lpoints of a relative point
// (coef = sl/ave, mrqcof.c:94), w: the weight w/sig^2.
struct PassRowData {
    double y_a, y_b, sl_a, sl_b, w_a, w_b;
};
// wqc: the weight of the three convexity rows, 1/sig^2 with sig = 1/conw_r (period_search_BOINC.cpp:721-726),
// formed by the kernel from the call's conw_r.
__device__ inline PassRowData load_pass_rows(const DevProblem &P, int tile0, double wqc) {
    const int c = threadIdx.x & 63;
    const int q = tile0 * 16 + c;  // the point arrays are padded by a whole pass
    const bool ca = q >= P.conv_q && q < P.conv_q + 3, cb = q + 64 >= P.conv_q && q + 64 < P.conv_q + 3;
    return {gload(P.y, q), gload(P.y, q + 64), gload(P.qsl, q), gload(P.qsl, q + 64), ca ? wqc : gload(P.wq, q),
            cb ? wqc : gload(P.wq, q + 64)};
}
__device__ inline double conv_weight(double conw_r) {
    const double sig = 1 / conw_r;
    return 1 / (sig * sig);
}

// ---------------------------------------------------------------------------------------------
// Phase C of one pass: every relative lightcurve (unit) of the pass is normalised in place
// (mrqcof.c:53-101): dave_c = the sum of its raw column c, ave = the sum of its model values,
// coef = sig*Lpoints/ave; dyda_c <- coef*(dyda_c - (ymod*dave_c)*(1/ave)), the c00 column 0, and the
// residual column y - coef*ymod.  The column sums are split over the waves: lane (pl, kg) of wave w
// adds column 16w + pl over the rows r0 + kg, r0 + kg + 4, ..., the four row classes are added across
// the lane groups in a fixed order, ((0 + 1) + (2 + 3)), and the 64 sums go through LDS (sh.csum, two
// units in flight, so one barrier per unit suffices).  Each wave then normalises rows r0 + w, r0 + w + 4, ... with lane c
// on column c, taking the per-row factors from lane i (pass rows i and 64 + i).
// ---------------------------------------------------------------------------------------------
template <int NFM>
__device__ inline void normalise_pass(const DevProblem &P, Shared2<NFM> &sh, const PassRowData &rd, int u0, int u1) {
    const int c = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pl = c & 15, kg = c >> 4;
    const int mfit = P.mfit, c00col = P.c00col;
    const int *unit_rows = P.unit_rows;
    // the columns the row loop writes: the residual and c00 columns are written separately
    const bool wcol = c < mfit && c != c00col;
    const int cl = min(c, mfit);  // the column a lane reads (lanes past mfit read a valid one, unused)
    // the summed column, clamped: lanes past mfit add a valid column and their sums are not used
    const int scol = min(16 * wave + pl, mfit);
    const double ym_a = sh.u.rows[c * kJP + mfit], ym_b = sh.u.rows[(64 + c) * kJP + mfit];
    for (int u = u0; u < u1; ++u) {
        const int rg = uload(unit_rows, u), r0 = rg & 255, r1 = rg >> 8;
        // class kg: rows r0 + kg, r0 + kg + 4, ... of the whole groups of four, in order; the rows past
        // the last whole group go to class 0, in order
        const int nfull = (r1 - r0) >> 2;
        const double *cp = sh.u.rows + (r0 + kg) * kJP + scol;  // the lane's class, row by row
        double sk = 0;
        int i = 0;
        for (; i + 4 <= nfull; i += 4) {  // four reads in flight, added in row order
            double x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = cp[(4 * (i + j)) * kJP];
#pragma unroll
            for (int j = 0; j < 4; ++j) sk += x[j];
        }
        for (; i < nfull; ++i) sk += cp[(4 * i) * kJP];
        for (int r = r0 + 4 * nfull; r < r1; ++r) {
            const double x = sh.u.rows[r * kJP + scol];
            sk += kg == 0 ? x : 0.0;
        }
        double *cs = sh.csum[(u - u0) & 1];
        const double sw = sum_groups(sk);  // (class 0 + class 1) + (class 2 + class 3)
        if (kg == 0) cs[16 * wave + pl] = sw;
        // every wave has read the rows of the unit and written its sums
        __syncthreads();
        const double s = cs[c], ave = cs[mfit];
        const double rave = 1 / ave;
        const double coef_a = rd.sl_a / ave, coef_b = rd.sl_b / ave;
        const double res_a = rd.y_a - coef_a * ym_a, res_b = rd.y_b - coef_b * ym_b;
        // The residual column and the c00 column (0) of row r: by lane r (or r - 64) of the wave that
        // normalises the row, so that every row is written by one wave.
        auto fixed_cols = [&](int r, double res) {
            if (r >= r0 && r < r1 && (r - r0) % kWaves2 == wave) {
                double *row = sh.u.rows + r * kJP;
                row[mfit] = res;
                if (c00col >= 0) row[c00col] = 0.0;
            }
        };
        fixed_cols(c, res_a);
        fixed_cols(64 + c, res_b);
        // the other columns of row r, with the row's factors from lane r (rows below 64) or r - 64: the
        // two halves as two loops, so the lane reads need no selects
        auto update = [&](int r, double coef, double ym) {
            double *row = sh.u.rows + r * kJP;
            const double v = coef * (row[cl] - (ym * s) * rave);
            if (wcol) row[c] = v;
        };
        int r = r0 + wave;
        const int rm = min(r1, 64);
#pragma unroll 4
        for (; r < rm; r += kWaves2) update(r, read_lane(coef_a, r), read_lane(ym_a, r));
#pragma unroll 4
        for (; r < r1; r += kWaves2) update(r, read_lane(coef_b, r - 64), read_lane(ym_b, r - 64));
    }
}

// ---------------------------------------------------------------------------------------------
// Phase D: the Gram blocks a wave owns, G[16bi + i][16bj + j] = sum_p w_p v_p[16bi + i] v_p[16bj + j]
// over the rows of the pass (v = the normalised fitted columns plus the residual), on
// v_mfma_f64_16x16x4f64: A[i][k] = v[row 4s + k][16bi + i] in lane i + 16k, B[k][j] = w v[row][16bj
// + j] in lane j + 16k, result G[16bi + 4r + kg][16bj + pl] in register r of lane l.  Ownership for
// NB = 4 column blocks (lower triangle: 10 blocks): wave 0 {(0,0),(1,1),(1,0)}, wave 1
// {(2,2),(3,3),(3,2)}, wave 2 {(2,0),(3,0)}, wave 3 {(2,1),(3,1)}: at most three MFMAs and three
// column reads per K-step.
// ---------------------------------------------------------------------------------------------
struct OwnList {
    int n;
    int bi[kOwn], bj[kOwn];
};
__host__ __device__ constexpr OwnList own_blocks(int nb, int w) {
    if (nb == 4) {
        constexpr int bi[4][3] = {{0, 1, 1}, {2, 3, 3}, {2, 3, 0}, {2, 3, 0}};
        constexpr int bj[4][3] = {{0, 1, 0}, {2, 3, 2}, {0, 0, 0}, {1, 1, 0}};
        return {w < 2 ? 3 : 2, {bi[w][0], bi[w][1], bi[w][2]}, {bj[w][0], bj[w][1], bj[w][2]}};
    }
    if (nb == 3) {
        constexpr int bi[4][2] = {{0, 1}, {1, 2}, {2, 0}, {2, 0}};
        constexpr int bj[4][2] = {{0, 0}, {1, 1}, {2, 0}, {0, 0}};
        return {w < 2 ? 2 : 1, {bi[w][0], bi[w][1], 0}, {bj[w][0], bj[w][1], 0}};
    }
    if (nb == 2) {
        constexpr int bi[4] = {0, 1, 1, 0};
        constexpr int bj[4] = {0, 1, 0, 0};
        return {w < 3 ? 1 : 0, {bi[w], 0, 0}, {bj[w], 0, 0}};
    }
    return {w == 0 ? 1 : 0, {0, 0, 0}, {0, 0, 0}};
}

__host__ __device__ constexpr int own_used(int nb, int w) {  // bit cb: the wave reads column block cb
    const OwnList o = own_blocks(nb, w);
    int m = 0;
    for (int b = 0; b < o.n; ++b) m |= (1 << o.bi[b]) | (1 << o.bj[b]);
    return m;
}

template <int NB, int W, int NFM>
__device__ __forceinline__ void gram_pass_w(const DevProblem &P, const Shared2<NFM> &sh, const PassRowData &rd, int ntl,
                                            Prec<false>::V4 acc[kOwn]) {
    using Pd = Prec<false>;
    constexpr OwnList o = own_blocks(NB, W);
    constexpr int used = own_used(NB, W);
    const int lane = threadIdx.x & 63, pl = lane & 15, kg = lane >> 4;
    const int ncol = P.mfit + 1;
    // the lane's column of each block, clamped: a lane past the last column reads column 0, and its
    // products land only in rows / columns of G that are not stored
    int col[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) col[cb] = (16 * cb + pl < ncol) ? 16 * cb + pl : 0;
    const double *base = sh.u.rows + kg * kJP;
    for (int t = 0; t < ntl; ++t) {  // one tile: four K-steps of four rows
        const double wt = t < 4 ? rd.w_a : rd.w_b;
        // the weights of the lane's rows in the four K-steps, fetched together (one LDS-crossbar
        // latency per tile instead of one per K-step)
        double w4[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) w4[s4] = __shfl(wt, (16 * t + 4 * s4 + kg) & 63);
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int r = 16 * t + 4 * s4;  // K-step rows r .. r + 3, the lane's row r + kg
            const double w = w4[s4];
            double v[4];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) v[cb] = (used & (1 << cb)) ? base[r * kJP + col[cb]] : 0.0;
#pragma unroll
            for (int b = 0; b < o.n; ++b) acc[b] = Pd::mma(v[o.bi[b]], w * v[o.bj[b]], acc[b]);
        }
    }
}

template <int NB, int NFM>
__device__ __forceinline__ void gram_pass_nb(const DevProblem &P, const Shared2<NFM> &sh, const PassRowData &rd, int ntl,
                                             int ow, Prec<false>::V4 acc[kOwn]) {
    if (ow == 0)
        gram_pass_w<NB, 0>(P, sh, rd, ntl, acc);
    else if (ow == 1)
        gram_pass_w<NB, 1>(P, sh, rd, ntl, acc);
    else if (ow == 2)
        gram_pass_w<NB, 2>(P, sh, rd, ntl, acc);
    else
        gram_pass_w<NB, 3>(P, sh, rd, ntl, acc);
}

// nb = the number of 16-column blocks covering mfit + 1 columns, ow = the ownership set the wave
// holds in this pass (both wave-uniform)
template <int NFM>
__device__ __forceinline__ void gram_pass(int nb, const DevProblem &P, const Shared2<NFM> &sh, const PassRowData &rd,
                                          int ntl, int ow, Prec<false>::V4 acc[kOwn]) {
    if (nb == 4)
        gram_pass_nb<4>(P, sh, rd, ntl, ow, acc);
    else if (nb == 3)
        gram_pass_nb<3>(P, sh, rd, ntl, ow, acc);
    else if (nb == 2)
        gram_pass_nb<2>(P, sh, rd, ntl, ow, acc);
    else
        gram_pass_nb<1>(P, sh, rd, ntl, ow, acc);
}

// The owned blocks into alpha (full, the lower triangle mirrored as mrqcof.c:143-145 does), beta and
// (into sh.scal[0]) chi2.
template <int NFM>
__device__ __forceinline__ void gram_store(int nb, const DevProblem &P, Shared2<NFM> &sh, int ow,
                                           const Prec<false>::V4 acc[kOwn], double *alpha, double *beta) {
    const int lane = threadIdx.x & 63, pl = lane & 15, kg = lane >> 4;
    const int mfit = P.mfit;
    const OwnList o = own_blocks(nb, ow);
#pragma unroll
    for (int b = 0; b < kOwn; ++b) {
        if (b < o.n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 16 * o.bi[b] + Prec<false>::row(r, kg), j = 16 * o.bj[b] + pl;
                const double v = acc[b][r];
                if (i >= j) {  // diagonal blocks: the lower triangle
                    if (i < mfit) {
                        alpha[i * mfit + j] = v;
                        alpha[j * mfit + i] = v;
                    } else if (i == mfit && j < mfit) {
                        beta[j] = v;
                    } else if (i == mfit && j == mfit) {
                        sh.scal[0] = v;
                    }
                }
            }
        }
    }
}

// The accumulators of ownership set ow to / from the slot's stash (between passes; the set changes
// hands between passes, the barrier at the end of each pass orders the store and the load).
__device__ __forceinline__ void stash_io(int nb, int ow, Prec<false>::V4 acc[kOwn], double *stash, bool store) {
    const int lane = threadIdx.x & 63;
    const int n = own_blocks(nb, ow).n;
    d4 *p = reinterpret_cast<d4 *>(stash) + size_t(ow) * kOwn * 64 + lane;
#pragma unroll
    for (int b = 0; b < kOwn; ++b) {
        if (b < n) {
            if (store)
                p[b * 64] = acc[b];
            else
                acc[b] = p[b * 64];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// mrqcof for the parameters in sh.a: alpha[mfit][mfit], beta[mfit] to the given buffers, chi2
// returned.  raw_y / raw_dyda: the raw rows (ps_eval).
// ---------------------------------------------------------------------------------------------
template <int NT, int R, bool LS, int NFM>
__device__ __forceinline__ double mrqcof2(const DevProblem &P, Shared2<NFM> &sh, double wqc, double *area,
                                          double *stash, double *alpha, double *beta, double *raw_y,
                                          double *raw_dyda) {
    using Pd = Prec<false>;
    curv2(P, sh.a, sh, area);
    blmatrix_lds(sh.a[P.nc], sh.a[P.nc + 1], sh);
    __syncthreads();
    const int npass = P.npass;
    const int nb = (P.mfit + 16) >> 4;  // 16-column blocks covering mfit + 1 columns
    for (int k = 0; k < npass; ++k) {
        // the schedule through the scalar cache: wave-uniform trip counts and row ranges
        const int tile0 = uload(P.pass_tile0, k), ntl = uload(P.pass_tile0, k + 1) - tile0;
        const int u0 = uload(P.pass_unit0, k), u1 = uload(P.pass_unit0, k + 1);
        const bool last = k == npass - 1;
        // the facet loop issues at the lowest wave priority: when the other workgroup on the CU is in a
        // latency-bound phase (lightcurve sums, Gram, Gauss-Jordan), its instructions go first and the
        // facet loop fills the gaps
        __builtin_amdgcn_s_setprio(0);
        model_rows2<NT, R, LS>(P, sh, tile0, ntl, k & 3, last, area, raw_y, raw_dyda);
        __builtin_amdgcn_s_setprio(2);
        const PassRowData rd = load_pass_rows(P, tile0, wqc);
        __syncthreads();
        normalise_pass(P, sh, rd, u0, u1);
        __syncthreads();
        // the accumulators live only through phase D: parked in the slot between passes.  The
        // ownership sets with three blocks go to waves 0, 1 in even passes and 2, 3 in odd ones, so the
        // Gram work falls evenly on the four SIMDs
        const int ow = (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) + 2 * (k & 1)) & 3;
        Pd::V4 acc[kOwn];
#pragma unroll
        for (int b = 0; b < kOwn; ++b) acc[b] = Pd::zero();
        if (k > 0) stash_io(nb, ow, acc, stash, false);
        gram_pass(nb, P, sh, rd, ntl, ow, acc);
        if (last)
            gram_store(nb, P, sh, ow, acc, alpha, beta);
        else
            stash_io(nb, ow, acc, stash, true);
        __syncthreads();
    }
    return sh.scal[0];
}

// ---------------------------------------------------------------------------------------------
// Full-pivot Gauss-Jordan (gauss_errc.c:9-73) on sh.u.g.cov with right-hand side sh.da, four waves each
// holding rows w, w + 4, ... in registers.  Returns 0, 1 (no valid pivot) or 2 (zero pivot); only the
// solution is written back.
//
// The reference swaps the pivot row into row icol (:38-42).  Here rows stay in place and carry a label,
// the index the row has in the reference's swapped matrix: the pivot row takes label icol and the row
// labelled icol takes label irow, so every row receives the reference's arithmetic, candidates are
// ordered by (|a|, label * n + column) exactly as the reference scans them (rows ascending, `>=`), and
// the solution is read back by label.  Each wave publishes its best candidate together with the row
// that holds it, so one barrier per pivot suffices: after it every wave knows the pivot and reads the
// pivot row from the winner's slot.  The slots are double-buffered, since a fast wave publishes the
// next pivot's candidate while a slow one may still read this pivot's row.
// ---------------------------------------------------------------------------------------------
template <int NFM>
__device__ inline int gauss_jordan2(Shared2<NFM> &sh, int n) {
    const int tid = threadIdx.x, l = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    n = __builtin_amdgcn_readfirstlane(n);
    double *a = sh.u.g.cov, *b = sh.da;
    double row[kGJRows2];
    int lab[kGJRows2];  // wave-uniform
    PivotCand c{-1.0, -1};
#pragma unroll
    for (int k = 0; k < kGJRows2; ++k) {
        const int ll = wave + kWaves2 * k;
        lab[k] = ll;
        row[k] = (ll < n && l <= n) ? ((l < n) ? a[ll * n + l] : b[ll]) : 0.0;
        if (ll < n && l < n) pivot_take(c, row[k], ll * n + l);
    }
    unsigned long long done = 0;
    for (int i = 0; i < n; ++i) {
        const int buf = i & 1;
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
            for (int k = 0; k < kGJRows2; ++k)
                if (lab[k] == vr) v = row[k];
            sh.u.g.cand[buf][wave][l] = v;
        }
        __syncthreads();
        PivotCand cw[kWaves2];
#pragma unroll
        for (int w = 0; w < kWaves2; ++w) cw[w] = {sh.u.g.red_v[buf][w], sh.u.g.red_q[buf][w]};
        int ww = 0;  // the winning wave: candidate indices of different waves are distinct
#pragma unroll
        for (int st = 1; st < kWaves2; st *= 2)
#pragma unroll
            for (int w = 0; w < kWaves2; w += 2 * st) pivot_merge(cw[w], cw[w + st]);
        const int cq = __builtin_amdgcn_readfirstlane(cw[0].q);
        if (cq < 0) return 1;
#pragma unroll
        for (int w = 1; w < kWaves2; ++w)
            if (__builtin_amdgcn_readfirstlane(sh.u.g.red_q[buf][w]) == cq) ww = w;
        const int irow = cq / n, icol = cq - irow * n;
        done |= 1ull << icol;
        const double *prow = sh.u.g.cand[buf][ww];
        const double pv = prow[l];  // one read of the pivot row: the pivot is its lane icol
        const double s = read_lane(pv, icol);
        if (s == 0.0) return 2;
        const double pivinv = 1.0 / s;
        // the new pivot row, gauss_errc.c:44-53 (lanes past n are not read)
        const double pr = ((l == icol) ? 1.0 : pv) * pivinv;
        c = {-1.0, -1};
        const bool cand_col = l < n && !((done >> l) & 1);
        const bool piv_lane = l == icol;
        // eliminate column icol from every other row (:55-62), searching the next pivot
#pragma unroll
        for (int k = 0; k < kGJRows2; ++k) {
            const int lb = lab[k];
            if (lb >= n) continue;
            if (lb == irow) {
                row[k] = pr;
                lab[k] = icol;
                continue;
            }
            if (lb == icol) lab[k] = irow;
            const double cur = row[k];
            const double dum = read_lane(cur, icol);
            const double v = (piv_lane ? 0.0 : cur) - pr * dum;
            row[k] = v;
            if (!((done >> lab[k]) & 1)) pivot_merge(c, {fabs(v), lab[k] * n + l}, cand_col);  // a scalar branch: the row test is wave-uniform
        }
    }
#pragma unroll
    for (int k = 0; k < kGJRows2; ++k)
        if (lab[k] < n && l == n) b[lab[k]] = row[k];
    return 0;
}

template <int NFM>
__device__ inline void load_facets2(const DevProblem &P, Shared2<NFM> &sh) {
    for (int i = threadIdx.x; i < P.nf; i += kWG2) {
        sh.nx[i] = P.nx[i];
        sh.ny[i] = P.ny[i];
        sh.nz[i] = P.nz[i];
    }
}

// ---------------------------------------------------------------------------------------------
// One (period, pole) fit: period_search_BOINC.cpp:819-911 with mrqmin.c.  The slot's buffers: area,
// stash, and alpha/beta double-buffered (accepted / trial).  sh.nx/ny/nz hold the problem's normals.
// ---------------------------------------------------------------------------------------------
struct SlotPtrs {
    double *area, *stash, *alpha0, *alpha1, *beta0, *beta1;
};
__device__ inline SlotPtrs slot_ptrs(const Slots2 &S, int slot) {
    SlotPtrs p;
    p.area = S.area + S.f_stride * slot;
    p.stash = S.stash + kStashPerSlot * slot;
    p.alpha0 = S.alpha + S.a_stride * (2 * size_t(slot));
    p.alpha1 = p.alpha0 + S.a_stride;
    p.beta0 = S.beta + S.b_stride * (2 * size_t(slot));
    p.beta1 = p.beta0 + S.b_stride;
    return p;
}

template <int NT, int R, bool LS, int NFM>
__device__ __forceinline__ void fit_one(const DevProblem &P, Shared2<NFM> &sh, const SlotPtrs &sp, double freq,
                                        int pole, double conw_r, ps_fit_result *out) {
    const int tid = threadIdx.x;
    const int nc = P.nc, ma = P.ma, mfit = P.mfit;
    double *const area = sp.area;
    const double wqc = conv_weight(conw_r);
    if (pole < 1 || pole > PS_N_POLES || !(freq > 0) || !(conw_r > 0)) {
        // a job the host did not check (ps_fit_batch_noalloc reads device arrays): no fit, niter = -1
        if (tid == 0) {
            const double nan = __builtin_nan("");
            *out = {nan, nan, nan, nan, nan, nan, nan, -1, 0};
        }
        __syncthreads();
        return;
    }
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
    // the LM controls and state are wave-uniform.  The compiler cannot see that for the state: its
    // values meet at the joins of lane-predicated stores (tid < ma), so settle() re-asserts them after
    // each such region; the loop exit and the accept / reject decisions then branch on scalars.
    const int n_iter_max = __builtin_amdgcn_readfirstlane(P.n_iter_max);
    const int n_iter_min = __builtin_amdgcn_readfirstlane(P.n_iter_min);
    const double iter_diff_max = uniform_double(P.iter_diff_max), alamda_start = uniform_double(P.alamda_start);
    const double alamda_incr = uniform_double(P.alamda_incr);
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
        if (!(((niter < n_iter_max) && (iter_diff > iter_diff_max)) || (niter < n_iter_min))) break;
        // One mrqcof per pass: the first pass evaluates the starting point (mrqmin.c:34-52), later
        // passes solve the damped system and evaluate the trial point (mrqmin.c:53-79).
        const bool init = alamda < 0;
        int err = 0;
        bool fresh = false;  // the step was accepted: cg is the trial point just evaluated
        if (init) {
            alamda = alamda_start;
            if (tid < ma) sh.a[tid] = sh.cg[tid];
        } else {
            // the damped matrix: four reads in flight per lane; diagonal entries are every (mfit + 1)-th
            const double *al = cur ? sp.alpha1 : sp.alpha0;
            const int nn = mfit * mfit;
            for (int i0 = tid; i0 < nn; i0 += 4 * kWG2) {
                double v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = i0 + u * kWG2 < nn ? al[i0 + u * kWG2] : 0.0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int it = i0 + u * kWG2;
                    if (it < nn) sh.u.g.cov[it] = (it % (mfit + 1) == 0) ? v[u] * (1 + alamda) : v[u];
                }
            }
            for (int j = tid; j < mfit; j += kWG2) sh.da[j] = (cur ? sp.beta1 : sp.beta0)[j];
            __syncthreads();
            err = __builtin_amdgcn_readfirstlane(gauss_jordan2(sh, mfit));
            __syncthreads();
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
            const double chi = uniform_double(mrqcof2<NT, R, LS>(P, sh, wqc, area, sp.stash,
                                                                 (cur ^ other) ? sp.alpha1 : sp.alpha0,
                                                                 (cur ^ other) ? sp.beta1 : sp.beta0, nullptr, nullptr));
            if (init) {
                ochisq = chi;
                chisq = chi;  // a failed first step then reports this fit's own starting chi2
                if (tid < ma) sh.atry[tid] = sh.cg[tid];
            } else {
                chisq = chi;
                if (chisq < ochisq) {
                    alamda = alamda / alamda_incr;
                    cur ^= 1;
                    fresh = true;
                    if (tid < ma) sh.cg[tid] = sh.atry[tid];
                } else {
                    alamda = alamda_incr * alamda;
                    chisq = ochisq;
                }
            }
            __syncthreads();
        } else {
            ++nfail;
        }
        settle();
        if (init) continue;
        ++niter;
        if (niter == 1 || chisq < ochisq) {
            ochisq = chisq;
            if (tid < ma) sh.a[tid] = sh.cg[tid];
            __syncthreads();
            // curv(cg) (period_search_BOINC.cpp:874); after an accepted step cg is the trial point, whose
            // areas the evaluation has just computed from the same bits
            if (!__builtin_amdgcn_readfirstlane(fresh)) curv2(P, sh.a, sh, area);
            __syncthreads();
            if (__builtin_amdgcn_readfirstlane(tid >> 6) == 0) {
                double sn[3];
                facet_sums_n(area, sh.nx, sh.ny, sh.nz, P.nf, sn);
                if (tid == 0) {
                    sh.scal[1] = sn[0];
                    sh.scal[2] = sn[1];
                    sh.scal[3] = sn[2];
                }
            }
            __syncthreads();
            chck0 = uniform_double(sh.scal[1]);
            chck1 = uniform_double(sh.scal[2]);
            chck2 = uniform_double(sh.scal[3]);
            rchisq = chisq - (chck0 * chck0 + chck1 * chck1 + chck2 * chck2) * (conw_r * conw_r);
            __syncthreads();
        }
        settle();
        dev_new = sqrt(rchisq / __builtin_amdgcn_readfirstlane(P.nobs));
        if (dev_old - dev_new > 1e-10) {
            iter_diff = dev_old - dev_new;
            dev_old = dev_new;
        }
    }
    if (__builtin_amdgcn_readfirstlane(tid >> 6) == 0) {
        const double totarea = facet_sum(area, nullptr, P.nf);
        if (tid == 0) {
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
    }
    __syncthreads();
}

// The fits of one problem, a grid-stride loop over them.
template <int NT, int R, bool LS, int NFM>
__global__ __launch_bounds__(kWG2, 2) void fit_kernel2(const DevProblem *__restrict__ Pp, Slots2 S, const double *freq,
                                                       const int *pole, int nfits, double conw_r, ps_fit_result *out) {
    const DevProblem &P = *Pp;
    __shared__ Shared2<NFM> sh;
    const SlotPtrs sp = slot_ptrs(S, blockIdx.x);
    load_facets2(P, sh);
    for (int fi = blockIdx.x; fi < nfits; fi += gridDim.x) fit_one<NT, R, LS>(P, sh, sp, freq[fi], pole[fi], conw_r, out + fi);
}

// ---------------------------------------------------------------------------------------------
// The fits of many problems (asteroids) in one persistent launch: every workgroup claims the next job
// of a device work queue (one atomic counter, jobs ordered longest first on the host) until the queue
// is empty.  A job names its problem; the problem's coefficient layout selects the fit body.  The
// slot buffers are sized for the largest problem.
// ---------------------------------------------------------------------------------------------
// A pointer to global memory the whole wave holds, made provably uniform.  readfirstlane works on
// 32-bit ints: both halves go through unsigned, so the low half is not sign-extended into the high
// one.  The result is cast from the global address space, so the compiler reads its fields with
// global (or scalar) loads; through a generic pointer it would treat every field it loads as
// varying across lanes (a generic address may name per-lane scratch), and all LM control flow
// that depends on the problem's fields would branch per lane.
template <class T>
__device__ inline T *uniform_ptr(T *p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(v));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(v >> 32));
    auto *g = reinterpret_cast<__attribute__((address_space(1))) T *>((static_cast<unsigned long long>(hi) << 32) | lo);
    return (T *)g;
}

struct QueueJob {
    double freq;  // trial frequency [1/h]
    int prob;     // problem index
    int pole;     // initial pole 1..10
    int out;      // result index
    int pad;
};

template <bool LS>
__global__ __launch_bounds__(kWG2, 2) void queue_kernel2(const DevProblem *const *probs, const int *layout,
                                                         const double *conw, Slots2 S, const QueueJob *jobs, int njobs,
                                                         int *counter, ps_fit_result *out) {
    __shared__ Shared2<kMaxFac> sh;
    const SlotPtrs sp = slot_ptrs(S, blockIdx.x);
    int prev = -1;
    for (;;) {
        if (threadIdx.x == 0) sh.job = atomicAdd(counter, 1);
        __syncthreads();
        const int j = __builtin_amdgcn_readfirstlane(sh.job);
        __syncthreads();
        if (j >= njobs) break;
        const int prob = __builtin_amdgcn_readfirstlane(jobs[j].prob);
        const int pole = __builtin_amdgcn_readfirstlane(jobs[j].pole);
        const int oi = __builtin_amdgcn_readfirstlane(jobs[j].out);
        const double freq = jobs[j].freq;
        const DevProblem &P = *uniform_ptr(probs[prob]);
        if (prob != prev) {  // the next fit_one's first barrier publishes the normals
            load_facets2(P, sh);
            prev = prob;
        }
        const double cw = conw[prob];
        switch (__builtin_amdgcn_readfirstlane(layout[prob])) {
            case 0: fit_one<1, 0, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            case 1: fit_one<1, 9, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            case 2: fit_one<2, 4, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            case 3: fit_one<3, 1, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            default: fit_one<4, 0, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
        }
    }
}

// The same persistent queue in mixed precision: the materialised-Jacobian fit (FP32 facet loop on
// FP32 MFMA with the FP32 Dsph table in LDS, FP64 normal equations and solve, ps_kernels.h), one
// 512-lane workgroup per CU, each slot's Jacobian buffer sized for the largest problem.  Coefficient
// layouts 0-3 (Lmax <= 6).
template <bool LS>
__global__ __launch_bounds__(kWG) void queue_kernel_mixed(const DevProblem *const *probs, const int *layout,
                                                          const double *conw, Slots S, const QueueJob *jobs, int njobs,
                                                          int *counter, ps_fit_result *out) {
    __shared__ Shared sh;
    const SlotPtrsL sp = slot_ptrs_legacy(S, blockIdx.x);
    int prev = -1;
    for (;;) {
        if (threadIdx.x == 0) sh.job = atomicAdd(counter, 1);
        __syncthreads();
        const int j = __builtin_amdgcn_readfirstlane(sh.job);
        __syncthreads();
        if (j >= njobs) break;
        const int prob = __builtin_amdgcn_readfirstlane(jobs[j].prob);
        const int pole = __builtin_amdgcn_readfirstlane(jobs[j].pole);
        const int oi = __builtin_amdgcn_readfirstlane(jobs[j].out);
        const double freq = jobs[j].freq;
        const DevProblem &P = *uniform_ptr(probs[prob]);
        if (prob != prev) {  // the next fit's first barrier publishes the normals
            load_facets(P, sh);
            prev = prob;
        }
        const double cw = conw[prob];
        switch (__builtin_amdgcn_readfirstlane(layout[prob])) {
            case 0: fit_one_legacy<1, 0, true, true, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            case 1: fit_one_legacy<1, 9, true, true, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            case 2: fit_one_legacy<2, 4, true, true, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
            default: fit_one_legacy<3, 1, true, true, LS>(P, sh, sp, freq, pole, cw, out + oi); break;
        }
    }
}

// Resets the queue counters of a launch (a kernel rather than a memset: the launch sequence is
// captured in a graph).
__global__ void queue_reset(int *counter, int n) {
    if (int(threadIdx.x) < n) counter[threadIdx.x] = 0;
}

// mrqcof for externally supplied parameter vectors (ps_eval), J-free.
template <int NT, int R, int NFM>
__global__ __launch_bounds__(kWG2, 2) void eval_kernel2(const DevProblem *__restrict__ Pp, Slots2 S, const double *avec,
                                                     int nvec, double conw_r, double *alpha, double *beta,
                                                     double *chisq, double *raw_y, double *raw_dyda) {
    const DevProblem &P = *Pp;
    __shared__ Shared2<NFM> sh;
    const int tid = threadIdx.x;
    const int mfit = P.mfit;
    const SlotPtrs sp = slot_ptrs(S, blockIdx.x);
    double *area = sp.area, *stash = sp.stash, *al = sp.alpha0, *be = sp.beta0;
    load_facets2(P, sh);
    for (int v = blockIdx.x; v < nvec; v += gridDim.x) {
        if (tid < P.ma) sh.a[tid] = avec[size_t(v) * P.ma + tid];
        __syncthreads();
        const double c = mrqcof2<NT, R, true>(P, sh, conv_weight(conw_r), area, stash, al, be,
                                                  raw_y ? raw_y + size_t(v) * P.npts : nullptr,
                                                  raw_dyda ? raw_dyda + size_t(v) * P.npts * P.ma : nullptr);
        for (int it = tid; it < mfit * mfit; it += kWG2)
            if (alpha) alpha[size_t(v) * mfit * mfit + it] = al[it];
        for (int it = tid; it < mfit; it += kWG2)
            if (beta) beta[size_t(v) * mfit + it] = be[it];
        if (tid == 0 && chisq) chisq[v] = c;
        __syncthreads();
    }
}

}  // namespace psk
