/*
 * periodsearch_amd.h -- C ABI of the MI355X (gfx950) convex-inversion period scan.
 *
 * The drop-in boundary for the hot path of Asteroids@home period_search: for every trial period x
 * initial pole, the Levenberg-Marquardt fit of shape, spin and scattering parameters runs on the GPU
 * (HIP kernels for gfx950); the host reads period_search_in, calls this ABI and writes
 * period_search_out lines in the reference format.
 *
 * Reference interfaces replaced (paths relative to the reference repository root):
 *   ps_input_read/ps_input_parse  -> period_search/period_search_BOINC.cpp:277-485 (input parsing)
 *   ps_geometry_build             -> period_search_BOINC.cpp:498-556 + trifac.c, areanorm.c,
 *                                    sphfunc.c, ellfit.c (setup tables, initial ellipsoid)
 *   ps_prepare                    -> CUDAPrepare, period_search_cuda/period_search/start_CUDA.h:2-4
 *   ps_conw_search/ps_conw_round  -> CUDAPrecalc, start_CUDA.h:9-10 (period_search_BOINC.cpp:487-715)
 *   ps_scan                       -> CUDAStart, start_CUDA.h:6-7, without the MFILE (the host writes)
 *   ps_fit_batch                  -> the per-(period, pole) body of the scan, period_search_BOINC.cpp:817-921
 *   ps_fit_multi                  -> the same body for the fits of several asteroids in one batch (the reference
 *                                    runs one asteroid per process; no counterpart)
 *   ps_eval                       -> mrqcof (period_search/mrqcof.c:16-148) with bright()'s dyda
 *   ps_release                    -> CUDAUnprepare, start_CUDA.h:14
 *   ps_format_line                -> the out.printf of period_search_BOINC.cpp:927-931
 *   ps_run_files                  -> the process contract of period_search_BOINC.cpp main (:147-954)
 *
 * Conventions: plain C types only; the caller owns every host array (read-only during a call); the
 * library owns device memory.  Status codes: PS_OK (0) or a negative PS_ERR_*; ps_last_error()
 * returns a message for the last failure on the calling thread.  Unlike the reference, no function
 * calls exit().  Parameter vectors are 0-based: a[k] = cg[k+1] of the reference.
 */
#ifndef PERIODSEARCH_AMD_H
#define PERIODSEARCH_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define PS_OK 0
#define PS_ERR_ARG (-1)         /* invalid argument */
#define PS_ERR_PARSE (-2)       /* malformed period_search_in */
#define PS_ERR_LIMIT (-3)       /* a reference limit (constants.h) is exceeded */
#define PS_ERR_DEVICE (-4)      /* HIP runtime failure / no device */
#define PS_ERR_UNSUPPORTED (-5) /* shape not supported by the GPU kernels (see ps_prepare) */
#define PS_ERR_NOMEM (-6)
#define PS_ERR_IO (-7)

#define PS_N_POLES 10 /* period_search/constants.h:11 */

/* Kernel form of single-problem FP64 fits (ps_set_kernel, ps_autotune). */
#define PS_KERNEL_AUTO 0   /* the J-free kernels where they accept the problem, else the legacy ones */
#define PS_KERNEL_LEGACY 1 /* 512-lane workgroups, one per CU, the Jacobian rows through a per-fit buffer */
#define PS_KERNEL_JFREE 2  /* 256-lane workgroups, two per CU, lightcurve-aligned passes of 128 rows in LDS */

/* Precision of the fit kernels. */
#define PS_FP64 0  /* FP64 throughout (the reference arithmetic) */
#define PS_MIXED 1 /* FP32 facet sums / normal matrix, FP64 phase, chi2 test and solve */

/* Parsed period_search_in.  Arrays are allocated by the library (ps_input_read/parse) or by the
   caller (when the caller fills the struct itself); ps_input_free frees library-allocated ones. */
typedef struct ps_input {
    double per_start, per_step_coef, per_end; /* period interval [h] and step coefficient */
    int ia_prd;                               /* period free (1) / fixed (0) */
    double jd_00;                             /* epoch of zero time as read */
    double phi0;                              /* initial rotation angle [rad] */
    double conw;                              /* convexity weight */
    int lmax, mmax, nrows;                    /* Laplace series degree/order, triangulation rows */
    double par[3];                            /* phase-function initial values */
    int ia_par[3];
    double cl;                                /* initial Lambert coefficient */
    int ia_cl;
    double stop_condition;                    /* >1: iterations; <1: minimum rms improvement */
    int n_iter_min;
    double alamda_incr, alamda_start;
    int lcurves;                              /* observed lightcurves */
    int *lpoints;                             /* [lcurves] */
    int *inrel;                               /* [lcurves] 1 = relative, 0 = calibrated */
    int ndata;                                /* observed points */
    double *tim;                              /* [ndata] JD - jd_0 */
    double *brightness;                       /* [ndata] */
    double *sig;                              /* [ndata] mean brightness of the point's lightcurve */
    double *weight;                           /* [ndata] */
    double *ee;                               /* [ndata*3] unit asteroid->Earth vectors */
    double *ee0;                              /* [ndata*3] unit asteroid->Sun vectors */
    double jd_min, jd_max, jd_0;
    int onlyrel;                              /* 1 if every lightcurve is relative */
    double escl;                              /* initial size scale */
    double a_axis, b_axis, c_axis;            /* initial ellipsoid */
    int owned;                                /* 1 if the arrays belong to the library */
} ps_input;

/* Setup tables for one (nrows, lmax, mmax): Gaussian-image triangulation and spherical harmonics. */
typedef struct ps_geometry {
    int nrows, nfac, ndir, lmax, mmax, ncoef, npairs;
    double *nor;      /* [nfac*3] facet normals */
    double *darea;    /* [nfac] facet areas on the unit sphere */
    double *dsph;     /* [nfac*ncoef] spherical-harmonic basis per facet and coefficient */
    double *pleg;     /* [nfac*npairs] associated Legendre values per (m, l) pair, curv order */
    double *fc, *fs;  /* [nfac*(mmax+1)] cos(m phi), sin(m phi) */
    int *pair_m, *pair_l; /* [npairs] */
    double *cg_first; /* [ncoef] initial ellipsoid coefficients (ellfit) */
} ps_geometry;

/* One output line: 24*P [h], rms, chi2, dark facet % (conw*escl^2 on line 1), lambda, beta. */
typedef struct ps_line {
    double per, dev, chi2, dark, la, be;
} ps_line;

/* Result of one (period, pole) fit. */
typedef struct ps_fit_result {
    double dev;      /* rms of the best improvement, sqrt(rchisq / ndata) */
    double prd;      /* fitted period [d] */
    double dark;     /* |sum Area*n| at the last improvement */
    double totarea;  /* sum Area of the last curv evaluation */
    double la, be;   /* fitted pole [deg] (lambda not wrapped) */
    double chisq;    /* chi2 of the last iteration */
    int niter;       /* LM iterations run */
    int nsolve_fail; /* Gauss-Jordan failures (the step was skipped) */
} ps_fit_result;

typedef struct ps_handle ps_handle;

typedef struct ps_info {
    int ndata, npts, nlc, nfac, ncoef, ma, mfit, nmax;
    double freq_start, freq_end, freq_step, escl, conw_r0;
    int slots;         /* fits resident on the device at once */
    int precision;
} ps_info;

const char *ps_last_error(void);
const char *ps_version(void);

/* ---- input and setup (host) ---- */
int ps_input_read(const char *path, ps_input **out);
int ps_input_parse(const char *text, ps_input **out);
void ps_input_free(ps_input *in);
int ps_geometry_build(const ps_input *in, ps_geometry **out);
void ps_geometry_free(ps_geometry *geo);

/* ---- device ---- */
/* Uploads points, lightcurve table, facet tables and the free/fixed mask; allocates per-slot scratch.
   Returns PS_ERR_UNSUPPORTED when mfit > 63, ma > 63 or nfac > 512 (Lmax <= 6 and nrows <= 8 are
   supported; the reference allows up to MAX_N_PAR = 200 parameters and MAX_N_FAC = 1000 facets). */
int ps_prepare(const ps_input *in, const ps_geometry *geo, int device, ps_handle **out);
int ps_set_precision(ps_handle *h, int precision);
/* Sets the kernel form of single-problem FP64 fits (PS_KERNEL_*); batches and ps_fit_batch_noalloc always
   run the J-free form.  No counterpart in the reference (one CPU build). */
int ps_set_kernel(ps_handle *h, int kernel);
/* The kernel form the next single-problem FP64 fit will use (PS_KERNEL_LEGACY or PS_KERNEL_JFREE). */
int ps_get_kernel(const ps_handle *h);
/* The autotuner: times both kernel forms on a sample of the problem's own fits (the first trial periods
   x 10 poles at convexity weight conw_r, 2 x the J-free form's resident workgroups, the second of two
   launches of each form), sets the faster one and reports it in *kernel and the two times per fit
   (legacy, J-free) in ms_per_fit[2] (either may be NULL).  Results are kept per problem shape (ncoef,
   nfac, npts, mfit, Lambert column, iteration cap) for the life of the process.  Mixed precision and
   problems the J-free form does not accept get PS_KERNEL_LEGACY without timing. */
int ps_autotune(ps_handle *h, double conw_r, int *kernel, double *ms_per_fit);
int ps_get_info(const ps_handle *h, ps_info *info);
/* Runs nfits independent fits (freq[i] in 1/h, pole[i] in 1..10) with convexity weight conw_r. */
int ps_fit_batch(ps_handle *h, int nfits, const double *freq, const int *pole, double conw_r, ps_fit_result *out);
/* The fits of several prepared problems (asteroids) on one device as one batch: job j fits frequency
   freq[j] (1/h) with pole pole[j] (1..10) of problem hs[prob[j]] at convexity weight
   conw_r[prob[j]], into out[j].  When every problem runs the FP64 J-free kernels this is
   ps_batch_create + ps_batch_run + ps_batch_release (one persistent launch); otherwise each problem's
   jobs run as one launch on its own stream, all issued (largest problem first) before any result is
   waited for.  Results equal ps_fit_batch's for the same fits.  Each problem handle may appear once.
   The per-job body is that of period_search_BOINC.cpp:817-921. */
int ps_fit_multi(ps_handle *const *hs, int nh, const double *conw_r, int njobs, const int *prob, const double *freq,
                 const int *pole, ps_fit_result *out);

/* ps_fit_batch for callers that cannot allocate device memory (code under graph capture, tight serving
   loops, harnesses that forbid hipMalloc); it has no counterpart in the reference.  freq, pole and out
   are device arrays of nfits entries; the call allocates nothing and copies nothing between host and
   device.  The per-workgroup buffers come from a static device array of 40 MiB per GPU (held until the
   process ends), shared by every call on that GPU: calls that overlap on one GPU race.  Needs the
   FP64 J-free kernels (PS_ERR_UNSUPPORTED otherwise); a shape whose buffers do not fit gives
   PS_ERR_LIMIT.  A job with a pole outside 1..10 or a non-positive frequency yields niter = -1 and NaN
   fields.  Returns after the launch has finished. */
int ps_fit_batch_noalloc(ps_handle *h, int nfits, const double *freq, const int *pole, double conw_r,
                         ps_fit_result *out);

/* A batch of many problems (asteroids) on one device, for repeated runs.  The reference runs one
   asteroid per process and has no counterpart; the per-job body is period_search_BOINC.cpp:817-921.
   ps_batch_create: the problem table on the device and one pool of per-workgroup buffers sized for the
   largest problem.  All handles at PS_FP64: the J-free kernels (no relative lightcurve longer than 128
   points); all at PS_MIXED: the mixed-precision kernels (Lmax <= 6); otherwise PS_ERR_UNSUPPORTED or
   PS_ERR_ARG (precisions differ).  The handles must outlive the batch.
   ps_batch_run: njobs fits (job j: problem prob[j], frequency freq[j] in 1/h, pole pole[j] in 1..10, at
   convexity weight conw_r[prob[j]]) as one graph launch: a queue reset, then persistent workgroups
   that claim jobs from a device work queue, longest first, until it is empty.  out[j] equals
   ps_fit_batch's result for the same fit.  The graph is kept and replayed while the job counts repeat.
   ps_batch_kernel_time: device time of the graph launches since the last reset (HIP events). */
typedef struct ps_batch ps_batch;
int ps_batch_create(ps_handle *const *hs, int nh, ps_batch **out);
int ps_batch_run(ps_batch *b, const double *conw_r, int njobs, const int *prob, const double *freq, const int *pole,
                 ps_fit_result *out);
int ps_batch_kernel_time(ps_batch *b, double *ms, int *launches, int reset);
void ps_batch_release(ps_batch *b);
/* One round of the convexity-weight search over the first min(10, nmax) periods; ave_dark = mean
   best-pole dark-facet % (NaN counted as 1.0). */
int ps_conw_round(ps_handle *h, double conw_r, double *ave_dark);
/* Full convexity-weight search starting from conw/escl^2; returns new_conw (1 = converged). */
int ps_conw_search(ps_handle *h, double *conw_r, int *new_conw);
/* Scans periods n_from..n_to (1-based, inclusive; n_to <= 0 means nmax) into lines[n_to-n_from+1]. */
int ps_scan(ps_handle *h, int n_from, int n_to, double conw_r, ps_line *lines);
/* mrqcof for nvec parameter vectors a[nvec*ma]: alpha[nvec*mfit*mfit], beta[nvec*mfit], chisq[nvec]
   (normalised normal equations, fitted parameters only); when ymod/dyda are non-NULL also the raw
   bright()/conv() outputs per point, ymod[nvec*npts], dyda[nvec*npts*ma], before lightcurve
   normalisation.  Any output pointer may be NULL.  The device staging is kept by the handle between
   calls (grown when a call needs more).  The kernel time counts in ps_kernel_time. */
int ps_eval(ps_handle *h, int nvec, const double *a, double conw_r, double *alpha, double *beta, double *chisq,
            double *ymod, double *dyda);
/* Device time of the fit and evaluation kernels since the last reset, measured with HIP events on the
   library's stream: total milliseconds and number of launches. */
int ps_kernel_time(ps_handle *h, double *ms, int *launches, int reset);
void ps_release(ps_handle *h);

/* ---- output and process contract (host) ---- */
int ps_format_line(const ps_line *line, char *buf, int buflen);
/* period_search_in -> period_search_out with checkpoint/resume through state_path (reference state
   format: the number of lines written).  Lines are appended when a valid state file exists. */
int ps_run_files(const char *in_path, const char *out_path, const char *state_path, int device, int precision);

#ifdef __cplusplus
}
#endif
#endif
