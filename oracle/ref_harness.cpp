/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.  Links the reference's unmodified plain-C core
 * (period_search/{bright,mrqcof,mrqmin,gauss_errc,curv,conv,matrix,blmatrix,phasec,dot_product,
 * trifac,areanorm,sphfunc,ellfit,ludcmp,lubksb,memory}.c, compiled straight from the reference
 * checkout by oracle/Makefile with the reference Makefile's compiler and flags, g++ -O3) into
 * oracle/_ref/libps_ref.so.  The psr_* entry points have C linkage for ctypes.
 *
 * The reference driver period_search_BOINC.cpp needs the BOINC client headers and library, which
 * this image lacks, so it is not built.  This harness instead defines the global state the driver
 * owns (period_search_BOINC.cpp:127-143) and restates the driver loops (:807-945, :487-715) on top
 * of the reference's own mrqmin/mrqcof/curv/bright.  Input parsing reuses the oracle's parser (it
 * restates :277-485; the parsed arrays are copied into the reference's 1-based globals).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ps_oracle.h"

#define MAX_N_OBS 20000
#define MAX_LC 200
#define MAX_N_FAC 1000
#define MAX_N_PAR 200
#define MAX_LM 10
#define N_POLES 10

/* Global state of the reference (period_search_BOINC.cpp:127-143, declared in globals.h). */
int Lmax, Mmax, Niter, Lastcall, Ncoef, Numfac, Lcurves, Nphpar, Lpoints[MAX_LC + 1], Inrel[MAX_LC + 1], Deallocate,
    n_iter;
double Ochisq, Chisq, Alamda, Alamda_incr, Alamda_start, Phi_0, Scale, Area[MAX_N_FAC + 1], Darea[MAX_N_FAC + 1],
    Sclnw[MAX_LC + 1], Yout[MAX_N_OBS + 1], Fc[MAX_N_FAC + 1][MAX_LM + 1], Fs[MAX_N_FAC + 1][MAX_LM + 1],
    Tc[MAX_N_FAC + 1][MAX_LM + 1], Ts[MAX_N_FAC + 1][MAX_LM + 1], Dsph[MAX_N_FAC + 1][MAX_N_PAR + 1],
    Dg[MAX_N_FAC + 1][MAX_N_PAR + 1], Nor[MAX_N_FAC + 1][4], Blmat[4][4], Pleg[MAX_N_FAC + 1][MAX_LM + 1][MAX_LM + 1],
    Dblm[3][4][4], Weight[MAX_N_OBS + 1];

/* Reference core entry points (declarations.h). */
void trifac(int nrows, int **ifp);
void areanorm(double t[], double f[], int ndir, int nfac, int **ifp, double at[], double af[]);
void sphfunc(int ndir, double at[], double af[]);
void ellfit(double r[], double a, double b, double c, int ndir, int ncoef, double at[], double af[]);
int mrqmin(double **x1, double **x2, double x3[], double y[], double sig[], double a[], int ia[], int ma,
           double **covar, double **alpha);
double mrqcof(double **x1, double **x2, double x3[], double y[], double sig[], double a[], int ia[], int ma,
              double **alpha, double beta[]);
void curv(double cg[]);
void blmatrix(double bet, double lam);
double conv(int nc, double dres[], int ma);
double bright(double ee[], double ee0[], double t, double cg[], double dyda[], int ncoef);
int gauss_errc(double **aa, int n, double b[]);
double *vector_double(int length);
int *vector_int(int length);
double **matrix_double(int rows, int columns);
int **matrix_int(int rows, int columns);

extern "C" {
int psr_setup(const pso_problem *p);
void psr_set_conw(double conw_r);
int psr_ma(void);
int psr_ncoef(void);
void psr_cg_first(double *cg);
void psr_ia(int *ia);
void psr_geometry(double *nor, double *darea, double *dsph, double *at, double *af);
void psr_curv(const double *cg, double *area, double *dg);
void psr_blmatrix(double bet, double lam);
double psr_bright(const double *ee3, const double *ee03, double t, const double *cg, double *dyda, int ma,
                  double *scale);
double psr_conv(int nc, double *dres, int ma);
double psr_mrqcof(const double *cg, double *alpha_out, double *beta_out);
int psr_gauss_errc(double *a, int n, double *b);
int psr_fit(double freq, int m, double *out7);
int psr_run(int n_from, int n_to, double *lines, double *conw_out);
int psr_load_file(const char *path);
int psr_freq(double *f0, double *df, int *nmax);
}

static struct {
    int ready, ndata_obs, ndata, ma, nrows;
    double **ee, **ee0, **covar, **aalpha;
    double *tim, *brightness, *sig, *cg_first, *at, *af;
    int *ia;
    double freq_start, freq_end, freq_step, escl, conw_r, stop_condition, cl, par[4];
    int ia_par[4], n_iter_min;
    double chck[4];
} H;

static const double beta_pole[N_POLES + 1] = {0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
static const double lambda_pole[N_POLES + 1] = {0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300};

/* Copies a parsed problem into the reference globals and runs the reference's geometry setup
   (period_search_BOINC.cpp:728-802). */
int psr_setup(const pso_problem *p) {
    if (!H.ready) {
        H.ee = matrix_double(MAX_N_OBS, 3);
        H.ee0 = matrix_double(MAX_N_OBS, 3);
        H.covar = matrix_double(MAX_N_PAR, MAX_N_PAR);
        H.aalpha = matrix_double(MAX_N_PAR, MAX_N_PAR);
        H.tim = vector_double(MAX_N_OBS);
        H.brightness = vector_double(MAX_N_OBS);
        H.sig = vector_double(MAX_N_OBS);
        H.cg_first = vector_double(MAX_N_PAR);
        H.at = vector_double(MAX_N_FAC);
        H.af = vector_double(MAX_N_FAC);
        H.ia = vector_int(MAX_N_PAR);
        H.ready = 1;
    }
    Lmax = p->lmax;
    Mmax = p->mmax;
    Phi_0 = p->phi0;
    Alamda_incr = p->alamda_incr;
    Alamda_start = p->alamda_start;
    Lcurves = p->lcurves + 1;
    for (int i = 1; i <= p->lcurves; i++) {
        Lpoints[i] = p->lpoints[i];
        Inrel[i] = p->inrel[i];
    }
    Lpoints[Lcurves] = 3;
    Inrel[Lcurves] = 0;
    H.ndata_obs = p->ndata;
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

    Ncoef = pso_ncoef(Lmax, Mmax);
    int nrows = p->nrows;
    double *t = vector_double(4 * nrows * nrows + 4), *f = vector_double(4 * nrows * nrows + 4);
    int ndir = pso_vertex_dirs(nrows, t, f);
    Numfac = 8 * nrows * nrows;
    int **ifp = matrix_int(MAX_N_FAC, 4);
    trifac(nrows, ifp);
    areanorm(t, f, ndir, Numfac, ifp, H.at, H.af);
    sphfunc(Numfac, H.at, H.af);
    ellfit(H.cg_first, p->a_axis, p->b_axis, p->c_axis, Numfac, Ncoef, H.at, H.af);
    free(t);
    free(f);

    H.freq_start = 1 / p->per_start;
    H.freq_end = 1 / p->per_end;
    H.freq_step = 0.5 / (p->jd_max - p->jd_min) / 24 * p->per_step_coef;
    Nphpar = 3;
    H.ia[Ncoef + 1] = 1;
    H.ia[Ncoef + 2] = 1;
    H.ia[Ncoef + 3] = p->ia_prd;
    for (int i = 1; i <= Ncoef; i++) H.ia[i] = 1;
    if (p->onlyrel == 1) H.ia[1] = 0;
    for (int i = 1; i <= Nphpar; i++) H.ia[Ncoef + 3 + i] = p->ia_par[i];
    H.ia[Ncoef + 3 + Nphpar + 1] = p->ia_cl;
    H.ia[Ncoef + 3 + Nphpar + 2] = 0;
    H.ma = Ncoef + 5 + Nphpar;
    H.escl = p->escl;
    H.conw_r = p->conw / p->escl / p->escl;
    H.stop_condition = p->stop_condition;
    H.n_iter_min = p->n_iter_min;
    H.cl = p->cl;
    for (int i = 1; i <= 3; i++) {
        H.par[i] = p->par[i];
        H.ia_par[i] = p->ia_par[i];
    }
    H.ndata = H.ndata_obs;
    Lastcall = 0;
    Deallocate = 0;
    return 0;
}

void psr_set_conw(double conw_r) {
    H.conw_r = conw_r;
    for (int j = 1; j <= 3; j++) {
        H.brightness[H.ndata_obs + j] = 0;
        H.sig[H.ndata_obs + j] = 1 / conw_r;
    }
    H.ndata = H.ndata_obs + 3;
}

int psr_ma(void) { return H.ma; }
int psr_ncoef(void) { return Ncoef; }
void psr_cg_first(double *cg) {
    for (int i = 1; i <= Ncoef; i++) cg[i] = H.cg_first[i];
}
void psr_ia(int *ia) {
    for (int i = 1; i <= H.ma; i++) ia[i] = H.ia[i];
}
/* Geometry tables in the oracle's 0-based layouts. */
void psr_geometry(double *nor, double *darea, double *dsph, double *at, double *af) {
    for (int i = 1; i <= Numfac; i++) {
        for (int c = 1; c <= 3; c++)
            if (nor) nor[(i - 1) * 3 + c - 1] = Nor[i][c];
        if (darea) darea[i - 1] = Darea[i];
        if (dsph)
            for (int k = 1; k <= Ncoef; k++) dsph[(i - 1) * Ncoef + k - 1] = Dsph[i][k];
        if (at) at[i - 1] = H.at[i];
        if (af) af[i - 1] = H.af[i];
    }
}

/* Kernel-level entry points for parity tests. */
void psr_curv(const double *cg, double *area, double *dg) {
    curv((double *)cg);
    for (int i = 1; i <= Numfac; i++) {
        if (area) area[i - 1] = Area[i];
        if (dg)
            for (int k = 1; k <= Ncoef; k++) dg[(i - 1) * Ncoef + k - 1] = Dg[i][k];
    }
}
void psr_blmatrix(double bet, double lam) { blmatrix(bet, lam); }
double psr_bright(const double *ee3, const double *ee03, double t, const double *cg, double *dyda, int ma,
                  double *scale) {
    double ee[4], ee0[4];
    for (int k = 1; k <= 3; k++) {
        ee[k] = ee3[k - 1];
        ee0[k] = ee03[k - 1];
    }
    double br = bright(ee, ee0, t, (double *)cg, dyda, ma);
    if (scale) *scale = Scale;
    return br;
}
double psr_conv(int nc, double *dres, int ma) { return conv(nc, dres, ma); }
/* alpha_out: mfit*mfit row-major (0-based), beta_out: mfit */
double psr_mrqcof(const double *cg, double *alpha_out, double *beta_out) {
    double *beta = vector_double(MAX_N_PAR);
    double chisq = mrqcof(H.ee, H.ee0, H.tim, H.brightness, H.sig, (double *)cg, H.ia, H.ma, H.aalpha, beta);
    int mfit = 0;
    for (int j = 1; j <= H.ma; j++)
        if (H.ia[j]) mfit++;
    for (int j = 1; j <= mfit; j++) {
        for (int k = 1; k <= mfit; k++)
            if (alpha_out) alpha_out[(j - 1) * mfit + k - 1] = H.aalpha[j][k];
        if (beta_out) beta_out[j - 1] = beta[j];
    }
    free(beta);
    return chisq;
}
/* a: n*n row-major (0-based), overwritten by the inverse; b: n, overwritten by the solution */
int psr_gauss_errc(double *a, int n, double *b) {
    double **m = matrix_double(n, n);
    double *v = vector_double(n);
    for (int i = 1; i <= n; i++) {
        for (int j = 1; j <= n; j++) m[i][j] = a[(i - 1) * n + j - 1];
        v[i] = b[i - 1];
    }
    int rc = gauss_errc(m, n, v);
    for (int i = 1; i <= n; i++) {
        for (int j = 1; j <= n; j++) a[(i - 1) * n + j - 1] = m[i][j];
        b[i - 1] = v[i];
    }
    for (int i = 0; i <= n; i++) free(m[i]);
    free(m);
    free(v);
    return rc;
}

/* One (period, pole) fit on the reference core; restates period_search_BOINC.cpp:819-911. */
int psr_fit(double freq, int m, double *out7) {
    double cg[MAX_N_PAR + 2];
    double prd = 1 / freq;
    for (int i = 1; i <= Ncoef; i++) cg[i] = H.cg_first[i];
    cg[Ncoef + 1] = beta_pole[m];
    cg[Ncoef + 2] = lambda_pole[m];
    cg[Ncoef + 1] = 90 - cg[Ncoef + 1];
    cg[Ncoef + 1] = (M_PI / 180) * cg[Ncoef + 1];
    cg[Ncoef + 2] = (M_PI / 180) * cg[Ncoef + 2];
    cg[Ncoef + 3] = 24 * 2 * M_PI / prd;
    for (int i = 1; i <= Nphpar; i++) {
        cg[Ncoef + 3 + i] = H.par[i];
        H.ia[Ncoef + 3 + i] = H.ia_par[i];
    }
    cg[Ncoef + 3 + Nphpar + 2] = 1;
    cg[Ncoef + 3 + Nphpar + 1] = log(H.cl);
    int n_iter_max = 0;
    double iter_diff_max = -1, rchisq = -1;
    if (H.stop_condition > 1) {
        n_iter_max = (int)H.stop_condition;
        iter_diff_max = 0;
        H.n_iter_min = 0;
    }
    if (H.stop_condition < 1) {
        n_iter_max = 100;
        iter_diff_max = H.stop_condition;
    }
    Alamda = -1;
    Niter = 0;
    double iter_diff = 1e40, dev_old = 1e30, dev_new = 0;
    Lastcall = 0;
    while (((Niter < n_iter_max) && (iter_diff > iter_diff_max)) || (Niter < H.n_iter_min)) {
        mrqmin(H.ee, H.ee0, H.tim, H.brightness, H.sig, cg, H.ia, Ncoef + 5 + Nphpar, H.covar, H.aalpha);
        Niter++;
        if ((Niter == 1) || (Chisq < Ochisq)) {
            Ochisq = Chisq;
            curv(cg);
            for (int i = 1; i <= 3; i++) {
                H.chck[i] = 0;
                for (int j = 1; j <= Numfac; j++) H.chck[i] = H.chck[i] + Area[j] * Nor[j][i];
            }
            rchisq = Chisq - (pow(H.chck[1], 2) + pow(H.chck[2], 2) + pow(H.chck[3], 2)) * pow(H.conw_r, 2);
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
    double totarea = 0;
    for (int i = 1; i <= Numfac; i++) totarea = totarea + Area[i];
    double sum = pow(H.chck[1], 2) + pow(H.chck[2], 2) + pow(H.chck[3], 2);
    out7[0] = dev_new;
    out7[1] = 2 * M_PI / cg[Ncoef + 3];
    out7[2] = sqrt(sum);
    out7[3] = totarea;
    out7[4] = (180 / M_PI) * cg[Ncoef + 2];
    out7[5] = 90 - (180 / M_PI) * cg[Ncoef + 1];
    out7[6] = Chisq;
    return Niter;
}

static void psr_period(double freq, double *best5) {
    double dev_best = 1e40, per_best = 0, dark_best = 0, la_best = 0, be_best = 0;
    for (int m = 1; m <= N_POLES; m++) {
        double r[7];
        psr_fit(freq, m, r);
        if (r[0] < dev_best) {
            dev_best = r[0];
            per_best = r[1];
            dark_best = r[2] / r[3] * 100;
            la_best = r[4];
            be_best = r[5];
        }
    }
    if (la_best < 0) la_best += 360;
    best5[0] = dev_best;
    best5[1] = per_best;
    best5[2] = dark_best;
    best5[3] = la_best;
    best5[4] = be_best;
}

/* Convexity-weight search (:487-715) and main scan (:807-945) on the reference core.
   lines: 6 doubles per line (24P, rms, chi2, dark or conw, lambda, beta). */
int psr_run(int n_from, int n_to, double *lines, double *conw_out) {
    int new_conw = 0;
    double conw_r = H.conw_r, escl = H.escl;
    int nmax = (int)((H.freq_start - H.freq_end) / H.freq_step) + 1;
    int max_test = nmax < 10 ? nmax : 10;
    while ((new_conw != 1) && ((conw_r * escl * escl) < 10.0)) {
        psr_set_conw(conw_r);
        double sum_dark = 0.0;
        for (int n = 1; n <= max_test; n++) {
            double b[5];
            psr_period(H.freq_start - (n - 1) * H.freq_step, b);
            if (isnan(b[2])) b[2] = 1.0;
            sum_dark = sum_dark + b[2];
        }
        double ave = sum_dark / max_test;
        if (ave < 1.0) new_conw = 1;
        if (ave >= 1.0) conw_r = conw_r * 2;
        H.ndata = H.ndata_obs;
    }
    psr_set_conw(conw_r);
    if (n_to <= 0 || n_to > nmax) n_to = nmax;
    int cnt = 0;
    for (int n = n_from; n <= n_to; n++) {
        double b[5];
        psr_period(H.freq_start - (n - 1) * H.freq_step, b);
        double *l = &lines[6 * cnt++];
        l[0] = 24 * b[1];
        l[1] = b[0];
        l[2] = b[0] * b[0] * (H.ndata - 3);
        l[3] = (n == 1) ? conw_r * escl * escl : b[2];
        l[4] = b[3];
        l[5] = b[4];
    }
    if (conw_out) *conw_out = conw_r;
    return cnt;
}

/* Parses a period_search_in file with the oracle's parser and installs it in the reference globals. */
int psr_load_file(const char *path) {
    pso_problem p;
    int rc = pso_parse_file(path, &p);
    if (rc == 0) rc = psr_setup(&p);
    pso_problem_free(&p);
    return rc;
}

/* Frequency grid of the loaded problem (period_search_BOINC.cpp:784-806). */
int psr_freq(double *f0, double *df, int *nmax) {
    if (f0) *f0 = H.freq_start;
    if (df) *df = H.freq_step;
    if (nmax) *nmax = (int)((H.freq_start - H.freq_end) / H.freq_step) + 1;
    return 0;
}
