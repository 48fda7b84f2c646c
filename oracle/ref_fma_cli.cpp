/*
 * ref_fma_cli.cpp -- TEST INFRASTRUCTURE ONLY.  Times the reference's AVX-FMA CPU core
 * (period_search_optimization_avx_fma/period_search: mrqmin, mrqcof_fma, bright_fma, conv_fma,
 * gauss_errc_fma, curv, ... compiled straight from the reference checkout by oracle/Makefile with
 * -O3 -march=corei7-avx -mfma -DFMA, the flags that variant's Makefile means) as a second CPU baseline
 * beside the plain-C core.
 *
 *   ps_ref_fma_cli fits <period_search_in> <n_from> <n_to> <conw_r> [npoles]
 *       runs the (period, pole) fits of periods n_from..n_to (poles 1..npoles, default 10) and prints
 *       "fits <count> seconds <wall>"; with "lines" instead of "fits" it prints each fit's
 *       rms, period, chi2 and iteration count (%.17g) instead of the timing.
 *
 * The variant's driver (period_search_BOINC.cpp of that directory) needs the BOINC client library,
 * which this image lacks; this file defines the globals it owns (:138-161, zero-indexed Nor/Area/Darea/
 * Dg) and restates its setup (:520-660) and fit loop (:686-793) around the variant's own core.  The
 * input is parsed with the oracle's parser (a restatement of the same grammar, :277-485).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "constants.h"
#include "declarations.h"
#include "ps_oracle.h"

/* Global state of the variant (period_search_BOINC.cpp:138-161 of the AVX-FMA directory). */
int Lmax, Mmax, Niter, Lastcall, Ncoef, Numfac, Lcurves, Nphpar, Lpoints[MAX_LC + 1], Inrel[MAX_LC + 1], Deallocate,
    n_iter;
double Ochisq, Chisq, Alamda, Alamda_incr, Alamda_start, Phi_0, Scale, Sclnw[MAX_LC + 1], Yout[MAX_N_OBS + 1],
    Fc[MAX_N_FAC + 1][MAX_LM + 1], Fs[MAX_N_FAC + 1][MAX_LM + 1], Tc[MAX_N_FAC + 1][MAX_LM + 1],
    Ts[MAX_N_FAC + 1][MAX_LM + 1], Dsph[MAX_N_FAC + 1][MAX_N_PAR + 1], Blmat[4][4],
    Pleg[MAX_N_FAC + 1][MAX_LM + 1][MAX_LM + 1], Dblm[3][4][4], Weight[MAX_N_OBS + 1];
double Nor[3][MAX_N_FAC + 4] __attribute__((aligned(32))), Area[MAX_N_FAC + 4] __attribute__((aligned(32))),
    Darea[MAX_N_FAC + 4] __attribute__((aligned(32))), Dg[MAX_N_FAC + 8][MAX_N_PAR + 4] __attribute__((aligned(32)));

static const double beta_pole[N_POLES + 1] = {0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
static const double lambda_pole[N_POLES + 1] = {0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300};

static struct {
    int ndata;  // including the three convexity rows once the weight is set
    double **ee, **ee0, **covar, **aalpha, *tim, *brightness, *sig, *cg_first, *at, *af, *cg;
    int *ia;    // zero-indexed, as in the variant's driver
    double par[4], cl, stop_condition, conw_r, freq_start, freq_step;
    int ia_par[4], n_iter_min;
} H;

static double now() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int setup(const pso_problem *p, double conw_r) {
    H.ee = matrix_double(MAX_N_OBS, 3);
    H.ee0 = matrix_double(MAX_N_OBS, 3);
    H.covar = aligned_matrix_double(MAX_N_PAR, MAX_N_PAR);
    H.aalpha = aligned_matrix_double(MAX_N_PAR, MAX_N_PAR + 4);
    H.tim = vector_double(MAX_N_OBS);
    H.brightness = vector_double(MAX_N_OBS);
    H.sig = vector_double(MAX_N_OBS);
    H.cg_first = vector_double(MAX_N_PAR);
    H.cg = vector_double(MAX_N_PAR);
    H.at = vector_double(MAX_N_FAC);
    H.af = vector_double(MAX_N_FAC);
    H.ia = vector_int(MAX_N_PAR);
    Lmax = p->lmax;
    Mmax = p->mmax;
    Phi_0 = p->phi0;
    Alamda_incr = p->alamda_incr;
    Alamda_start = p->alamda_start;
    Lcurves = p->lcurves + 1;  // :526-528
    for (int i = 1; i <= p->lcurves; i++) {
        Lpoints[i] = p->lpoints[i];
        Inrel[i] = p->inrel[i];
    }
    Lpoints[Lcurves] = 3;
    Inrel[Lcurves] = 0;
    for (int i = 1; i <= p->ndata; i++) {
        H.tim[i] = p->tim[i - 1];
        H.brightness[i] = p->brightness[i - 1];
        H.sig[i] = p->sig[i - 1];
        for (int k = 1; k <= 3; k++) {
            H.ee[i][k] = p->ee[(i - 1) * 3 + k - 1];
            H.ee0[i][k] = p->ee0[(i - 1) * 3 + k - 1];
        }
    }
    for (int i = 1; i <= p->ndata + 3; i++) Weight[i] = p->weight[i - 1];
    // convexity rows (:554-561)
    H.ndata = p->ndata;
    for (int j = 1; j <= 3; j++) {
        H.ndata++;
        H.brightness[H.ndata] = 0;
        H.sig[H.ndata] = 1 / conw_r;
    }
    H.conw_r = conw_r;
    Ncoef = pso_ncoef(Lmax, Mmax);
    const int nrows = p->nrows;
    double *t = vector_double(4 * nrows * nrows + 4), *f = vector_double(4 * nrows * nrows + 4);
    const int ndir = pso_vertex_dirs(nrows, t, f);
    Numfac = 8 * nrows * nrows;
    int **ifp = matrix_int(MAX_N_FAC, 4);
    trifac(nrows, ifp);
    areanorm(t, f, ndir, Numfac, ifp, H.at, H.af);
    sphfunc(Numfac, H.at, H.af);
    ellfit(H.cg_first, p->a_axis, p->b_axis, p->c_axis, Numfac, Ncoef, H.at, H.af);
    H.freq_start = 1 / p->per_start;
    H.freq_step = 0.5 / (p->jd_max - p->jd_min) / 24 * p->per_step_coef;
    Nphpar = 3;  // :636-650, ia zero-indexed
    H.ia[Ncoef + 1 - 1] = 1;
    H.ia[Ncoef + 2 - 1] = 1;
    H.ia[Ncoef + 3 - 1] = p->ia_prd;
    for (int i = 0; i < Ncoef; i++) H.ia[i] = 1;
    if (p->onlyrel == 1) H.ia[0] = 0;
    H.ia[Ncoef + 3 + Nphpar + 1 - 1] = p->ia_cl;
    H.ia[Ncoef + 3 + Nphpar + 2 - 1] = 0;
    for (int i = 1; i <= 3; i++) {
        H.par[i] = p->par[i];
        H.ia_par[i] = p->ia_par[i];
    }
    H.cl = p->cl;
    H.stop_condition = p->stop_condition;
    H.n_iter_min = p->n_iter_min;
    Lastcall = 0;
    Deallocate = 0;
    return 0;
}

/* One (period, pole) fit: :686-793 of the variant's driver. */
static int fit(double freq, int m, double *dev_out, double *prd_out) {
    double *cg = H.cg;
    const double prd = 1 / freq;
    for (int i = 1; i <= Ncoef; i++) cg[i] = H.cg_first[i];
    cg[Ncoef + 1] = beta_pole[m];
    cg[Ncoef + 2] = lambda_pole[m];
    cg[Ncoef + 1] = 90 - cg[Ncoef + 1];
    cg[Ncoef + 1] = DEG2RAD * cg[Ncoef + 1];
    cg[Ncoef + 2] = DEG2RAD * cg[Ncoef + 2];
    cg[Ncoef + 3] = 24 * 2 * PI / prd;
    for (int i = 1; i <= Nphpar; i++) {
        cg[Ncoef + 3 + i] = H.par[i];
        H.ia[Ncoef + 3 + i - 1] = H.ia_par[i];
    }
    cg[Ncoef + 3 + Nphpar + 2] = 1;
    cg[Ncoef + 3 + Nphpar + 1] = log(H.cl);
    int n_iter_max = 0, n_iter_min = H.n_iter_min;
    double iter_diff_max = -1, rchisq = -1;
    if (H.stop_condition > 1) {
        n_iter_max = (int)H.stop_condition;
        iter_diff_max = 0;
        n_iter_min = 0;
    }
    if (H.stop_condition < 1) {
        n_iter_max = MAX_N_ITER;
        iter_diff_max = H.stop_condition;
    }
    Alamda = -1;
    Niter = 0;
    double iter_diff = 1e40, dev_old = 1e30, dev_new = 0, chck[4] = {0, 0, 0, 0};
    Lastcall = 0;
    while (((Niter < n_iter_max) && (iter_diff > iter_diff_max)) || (Niter < n_iter_min)) {
        mrqmin(H.ee, H.ee0, H.tim, H.brightness, H.sig, cg, H.ia, Ncoef + 5 + Nphpar, H.covar, H.aalpha);
        Niter++;
        if ((Niter == 1) || (Chisq < Ochisq)) {
            Ochisq = Chisq;
            curv(cg);
            for (int i = 1; i <= 3; i++) {
                chck[i] = 0;
                for (int j = 1; j <= Numfac; j++) chck[i] = chck[i] + Area[j - 1] * Nor[i - 1][j - 1];
            }
            rchisq = Chisq - (pow(chck[1], 2) + pow(chck[2], 2) + pow(chck[3], 2)) * pow(H.conw_r, 2);
        }
        dev_new = sqrt(rchisq / (H.ndata - 3));
        if (dev_old - dev_new > 1e-10) {
            iter_diff = dev_old - dev_new;
            dev_old = dev_new;
        }
    }
    Deallocate = 1;
    mrqmin(H.ee, H.ee0, H.tim, H.brightness, H.sig, cg, H.ia, Ncoef + 5 + Nphpar, H.covar, H.aalpha);
    Deallocate = 0;
    *dev_out = dev_new;
    *prd_out = 2 * PI / cg[Ncoef + 3];
    return Niter;
}

int main(int argc, char **argv) {
    if (argc < 6 || (strcmp(argv[1], "fits") != 0 && strcmp(argv[1], "lines") != 0)) {
        fprintf(stderr, "usage: %s fits|lines <period_search_in> <n_from> <n_to> <conw_r> [npoles]\n", argv[0]);
        return 2;
    }
    pso_problem p;
    if (pso_parse_file(argv[2], &p) != 0) {
        fprintf(stderr, "cannot load %s\n", argv[2]);
        return 2;
    }
    const bool timing = strcmp(argv[1], "fits") == 0;
    const int n_from = atoi(argv[3]), n_to = atoi(argv[4]);
    int npoles = argc > 6 ? atoi(argv[6]) : N_POLES;
    if (npoles < 1 || npoles > N_POLES) npoles = N_POLES;
    setup(&p, atof(argv[5]));
    pso_problem_free(&p);
    const double t0 = now();
    int nfits = 0;
    for (int n = n_from; n <= n_to; n++)
        for (int m = 1; m <= npoles; m++) {
            double dev, prd;
            const int it = fit(H.freq_start - (n - 1) * H.freq_step, m, &dev, &prd);
            if (!timing) printf("%d %d %.17g %.17g %.17g %d\n", n, m, dev, prd, Chisq, it);
            nfits++;
        }
    if (timing) printf("fits %d seconds %.6f\n", nfits, now() - t0);
    return 0;
}
