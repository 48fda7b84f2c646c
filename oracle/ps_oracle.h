/*
 * ps_oracle.h -- CPU restatement of the reference convex-inversion period scan (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.  The product library
 * (libpsamd.so) never links, loads or calls it.
 *
 * Every routine restates the arithmetic of the plain-C reference in period_search/ of
 * ahorek/PeriodSearch, in the same operation order, so that compiled with the same flags as the
 * reference core it reproduces the reference's results bit for bit (pinned against the reference's
 * own core sources compiled into oracle/_ref and against the reference unit-test golden vectors).
 *
 * Conventions: parameter vectors are 1-based (cg[1..ma], index 0 unused) exactly like the reference;
 * points, facets and lightcurves are 0-based.  No globals: all state lives in pso_ctx.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define PSO_POINTS_MAX 1000 /* period_search/constants.h:1 */
#define PSO_MAX_N_OBS 20000 /* constants.h:2 */
#define PSO_MAX_LC 200      /* constants.h:3 */
#define PSO_MAX_N_FAC 1000  /* constants.h:5 */
#define PSO_MAX_N_ITER 100  /* constants.h:6 */
#define PSO_MAX_N_PAR 200   /* constants.h:7 */
#define PSO_MAX_LM 10       /* constants.h:8 */
#define PSO_TINY 1e-8       /* constants.h:10 */
#define PSO_N_POLES 10      /* constants.h:11 */

/* Parsed period_search_in (period_search_BOINC.cpp:277-485). */
typedef struct pso_problem {
    double per_start, per_step_coef, per_end;
    int ia_prd;
    double jd_00, phi0; /* phi0 in radians after parsing (:442) */
    double conw;
    int lmax, mmax, nrows;
    double par[4];
    int ia_par[4]; /* 1-based, [1..3] */
    double cl;
    int ia_cl;
    double stop_condition;
    int n_iter_min;
    double alamda_incr, alamda_start;

    int lcurves;   /* observed lightcurves (the convexity pseudo-lc is not counted) */
    int lpoints[PSO_MAX_LC + 1];
    int inrel[PSO_MAX_LC + 1];
    int ndata;     /* observed points */
    double *tim;   /* [ndata + 3], JD - jd_0 */
    double *brightness, *sig, *weight; /* [ndata + 3] */
    double *ee, *ee0;                  /* [(ndata + 3) * 3] unit Earth / Sun directions */
    double jd_min, jd_max, jd_0;
    int onlyrel;
    double al0;
    int ial0;
    double escl, a_axis, b_axis, c_axis;
} pso_problem;

/* One period_search_out line: 24*P [h], rms, chi2, dark% (or conw on line 1), lambda, beta. */
typedef struct pso_line {
    double per, dev, chi2, dark, la, be;
} pso_line;

/* Result of one (period, pole) LM fit (period_search_BOINC.cpp:847-911). */
typedef struct pso_fit_result {
    double dev_new, prd, dark, totarea, la, be, chisq;
    int niter;
} pso_fit_result;

typedef struct pso_ctx pso_ctx;

/* ---- input ---- */
int pso_parse_file(const char *path, pso_problem *p);     /* 0 ok, <0 error */
int pso_parse_text(const char *text, pso_problem *p);
void pso_problem_free(pso_problem *p);

/* ---- setup (period_search_BOINC.cpp:498-556, trifac.c, areanorm.c, sphfunc.c, ellfit.c) ---- */
int pso_ncoef(int lmax, int mmax);
int pso_vertex_dirs(int nrows, double *t, double *f);  /* 1-based t[1..ndir]; returns ndir */
void pso_trifac(int nrows, int *ifp);                   /* ifp[(i-1)*3 + (v-1)], 1-based vertex ids */
void pso_areanorm(const double *t, const double *f, int ndir, int nfac, const int *ifp,
                  double *nor, double *darea, double *at, double *af); /* nor[i*3+c], 0-based i */
void pso_sphfunc(int nfac, const double *at, const double *af, int lmax, int mmax,
                 double *pleg, double *fc, double *fs, double *dsph);
/* pleg[(i*(L+1) + l)*(L+1) + m], fc/fs[i*(L+1)+m], dsph[i*ncoef + (k-1)] with L = PSO_MAX_LM */
int pso_ellfit(double *cg, double a, double b, double c, int nfac, int ncoef, const double *at,
               const double *af, const double *pleg, int lmax, int mmax); /* cg[1..ncoef] */
int pso_ludcmp(double *a, int n, int *indx, double *d);  /* a: (n+1)x(n+1) row-major, 1-based */
void pso_lubksb(const double *a, int n, const int *indx, double *b);

/* ---- context ---- */
pso_ctx *pso_ctx_create(const pso_problem *p);
pso_ctx *pso_ctx_from_file(const char *path, int *err);     /* parse + create; NULL on error */
pso_ctx *pso_ctx_from_text(const char *text, int *err);
double pso_ctx_escl(const pso_ctx *c);
double pso_ctx_conw_r(const pso_ctx *c);
int pso_ctx_ndata(const pso_ctx *c);                       /* observed points */
int pso_ctx_stop_iters(const pso_ctx *c);                  /* fixed LM iteration count or 0 */
void pso_ctx_free(pso_ctx *c);
int pso_ctx_ma(const pso_ctx *c);
int pso_ctx_ncoef(const pso_ctx *c);
int pso_ctx_nfac(const pso_ctx *c);
int pso_ctx_mfit(const pso_ctx *c);
void pso_ctx_ia(const pso_ctx *c, int *ia_out);              /* ia[1..ma] */
void pso_ctx_cg_first(const pso_ctx *c, double *cg_out);     /* cg_first[1..ncoef] */
void pso_ctx_geometry(const pso_ctx *c, double *nor, double *darea, double *dsph);
double pso_ctx_freq(const pso_ctx *c, double *f0, double *f1, double *df, int *nmax);
void pso_ctx_set_conw(pso_ctx *c, double conw_r);            /* appends/updates the 3 conv rows */
void pso_ctx_init_params(const pso_ctx *c, double freq, int pole, double *cg); /* cg[1..ma] */

/* ---- model (bright.c, curv.c, conv.c, matrix.c, blmatrix.c, phasec.c) ---- */
void pso_curv(pso_ctx *c, const double *cg);
void pso_blmatrix(pso_ctx *c, double bet, double lam);
double pso_phasec(double *dcdp, double alpha, const double *p);
void pso_matrix(const pso_ctx *c, double omg, double t, double tmat[4][4], double dtm[4][4][4]);
double pso_bright(pso_ctx *c, const double *ee, const double *ee0, double t, const double *cg,
                  double *dyda, int ma);
double pso_conv(pso_ctx *c, int nc, double *dres, int ma);
/* Direct setters for kernel-level golden tests (1-based facet/coef layout like the reference). */
void pso_ctx_set_area_dg(pso_ctx *c, const double *area, const double *dg);
void pso_ctx_set_blmat(pso_ctx *c, const double *blmat9, const double *dblm18);
void pso_ctx_get_area_dg(const pso_ctx *c, double *area, double *dg);

/* ---- LM (mrqcof.c, mrqmin.c, gauss_errc.c) ---- */
double pso_mrqcof(pso_ctx *c, const double *a, const int *ia, int ma, double *alpha, double *beta);
int pso_gauss_errc(double *a, int n, double *b); /* a: (n+1)x(n+1) row-major, 1-based */

/* ---- driver (period_search_BOINC.cpp:807-945, :488-715) ---- */
int pso_fit(pso_ctx *c, double freq, int pole, pso_fit_result *r);
int pso_conw_search(pso_ctx *c, double *conw_r, int *rounds);
int pso_scan(pso_ctx *c, int n_from, int n_to, double conw_r, pso_line *lines);
int pso_run(const pso_problem *p, int n_from, int n_to, pso_line *lines, double *conw_r_out);
int pso_format_line(const pso_line *l, char *buf, int buflen);

#ifdef __cplusplus
}
#endif
#endif
