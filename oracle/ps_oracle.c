/*
 * ps_oracle.c -- FP64 CPU restatement of the reference period scan.  TEST INFRASTRUCTURE ONLY:
 * the checker for the HIP product path (see ps_oracle.h).  Each function names the reference
 * file:line range (paths relative to the reference repository root) whose arithmetic it restates;
 * operation order follows the reference so that results agree bit for bit when compiled with the
 * same flags (no FMA contraction, glibc libm).
 */
#define _GNU_SOURCE
#include "ps_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PSO_PI M_PI
#define DEG2RAD (PSO_PI / 180)
#define RAD2DEG (180 / PSO_PI)
#define LM1 (PSO_MAX_LM + 1)

/* ------------------------------------------------------------------------------------------ */
/* Input parsing: period_search/period_search_BOINC.cpp:277-485.  A small reader that mimics     */
/* fscanf (whitespace-skipping numeric reads) followed by fgets (discard rest of the line).     */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    const char *s;
    size_t pos, len;
} rd_t;

static void rd_ws(rd_t *r) {
    while (r->pos < r->len && (r->s[r->pos] == ' ' || r->s[r->pos] == '\t' || r->s[r->pos] == '\n' ||
                               r->s[r->pos] == '\r' || r->s[r->pos] == '\v' || r->s[r->pos] == '\f'))
        r->pos++;
}
static int rd_double(rd_t *r, double *v) {
    rd_ws(r);
    if (r->pos >= r->len) return 0;
    char *end;
    double x = strtod(r->s + r->pos, &end);
    if (end == r->s + r->pos) return 0;
    r->pos = (size_t)(end - r->s);
    *v = x;
    return 1;
}
static int rd_int(rd_t *r, int *v) {
    rd_ws(r);
    if (r->pos >= r->len) return 0;
    char *end;
    long x = strtol(r->s + r->pos, &end, 10);
    if (end == r->s + r->pos) return 0;
    r->pos = (size_t)(end - r->s);
    *v = (int)x;
    return 1;
}
static void rd_eol(rd_t *r) {
    while (r->pos < r->len && r->s[r->pos] != '\n') r->pos++;
    if (r->pos < r->len) r->pos++;
}

int pso_parse_text(const char *text, pso_problem *p) {
    rd_t r = {text, 0, strlen(text)};
    memset(p, 0, sizeof(*p));
    /* header, period_search_BOINC.cpp:277-303 */
    if (!rd_double(&r, &p->per_start) || !rd_double(&r, &p->per_step_coef) || !rd_double(&r, &p->per_end) ||
        !rd_int(&r, &p->ia_prd))
        return -1;
    rd_eol(&r);
    if (!rd_double(&r, &p->jd_00)) return -1;
    rd_eol(&r);
    if (!rd_double(&r, &p->phi0)) return -1;
    rd_eol(&r);
    if (!rd_double(&r, &p->conw)) return -1;
    rd_eol(&r);
    if (!rd_int(&r, &p->lmax) || !rd_int(&r, &p->mmax)) return -1;
    rd_eol(&r);
    if (!rd_int(&r, &p->nrows)) return -1;
    rd_eol(&r);
    for (int i = 1; i <= 3; i++) {
        if (!rd_double(&r, &p->par[i]) || !rd_int(&r, &p->ia_par[i])) return -1;
        rd_eol(&r);
    }
    if (!rd_double(&r, &p->cl) || !rd_int(&r, &p->ia_cl)) return -1;
    rd_eol(&r);
    if (!rd_double(&r, &p->stop_condition)) return -1;
    rd_eol(&r);
    if (!rd_int(&r, &p->n_iter_min)) return -1;
    rd_eol(&r);
    if (!rd_double(&r, &p->alamda_incr)) return -1;
    rd_eol(&r);
    if (!rd_double(&r, &p->alamda_start)) return -1;
    rd_eol(&r);
    if (p->lmax < 0 || p->lmax > PSO_MAX_LM || p->mmax < 0 || p->mmax > p->lmax || p->nrows < 1) return -2;
    if (8 * p->nrows * p->nrows > PSO_MAX_N_FAC) return -2; /* :542-543 */

    /* lightcurves, :324-414 */
    if (!rd_int(&r, &p->lcurves)) return -1;
    if (p->lcurves > PSO_MAX_LC || p->lcurves < 1) return -2; /* :326-327 */
    /* first pass over the text is not possible with fscanf semantics; allocate the maximum */
    p->tim = (double *)malloc((PSO_MAX_N_OBS + 4) * sizeof(double));
    p->brightness = (double *)malloc((PSO_MAX_N_OBS + 4) * sizeof(double));
    p->sig = (double *)malloc((PSO_MAX_N_OBS + 4) * sizeof(double));
    p->weight = (double *)malloc((PSO_MAX_N_OBS + 4) * sizeof(double));
    p->ee = (double *)malloc((PSO_MAX_N_OBS + 4) * 3 * sizeof(double));
    p->ee0 = (double *)malloc((PSO_MAX_N_OBS + 4) * 3 * sizeof(double));
    if (!p->tim || !p->brightness || !p->sig || !p->weight || !p->ee || !p->ee0) return -3;

    double al0 = PSO_PI, al0_abs = PSO_PI;
    int ial0 = -1, ial0_abs = -1;
    double jd_min = 1e20, jd_max = -1e40;
    int onlyrel = 1, ndata = 0, k2 = 0;
    for (int i = 1; i <= p->lcurves; i++) {
        double ave = 0;
        int i_temp;
        if (!rd_int(&r, &p->lpoints[i]) || !rd_int(&r, &i_temp)) return -1;
        rd_eol(&r);
        p->inrel[i] = 1 - i_temp;
        if (p->inrel[i] == 0) onlyrel = 0;
        if (p->lpoints[i] > PSO_POINTS_MAX || p->lpoints[i] < 1) return -2; /* :352-353 */
        for (int j = 1; j <= p->lpoints[i]; j++) {
            double e[4], e0[4];
            ndata++;
            if (ndata > PSO_MAX_N_OBS) return -2; /* :360-361 */
            int np = ndata - 1;
            if (!rd_double(&r, &p->tim[np]) || !rd_double(&r, &p->brightness[np])) return -1;
            if (!rd_double(&r, &e0[1]) || !rd_double(&r, &e0[2]) || !rd_double(&r, &e0[3])) return -1;
            if (!rd_double(&r, &e[1]) || !rd_double(&r, &e[2]) || !rd_double(&r, &e[3])) return -1;
            if (p->tim[np] < jd_min) jd_min = p->tim[np];
            if (p->tim[np] > jd_max) jd_max = p->tim[np];
            double e0len = sqrt(e0[1] * e0[1] + e0[2] * e0[2] + e0[3] * e0[3]);
            double elen = sqrt(e[1] * e[1] + e[2] * e[2] + e[3] * e[3]);
            ave += p->brightness[np];
            for (int k = 1; k <= 3; k++) {
                p->ee[np * 3 + k - 1] = e[k] / elen;
                p->ee0[np * 3 + k - 1] = e0[k] / e0len;
            }
            if (j == 1) {
                double cos_alpha = (e[1] * e0[1] + e[2] * e0[2] + e[3] * e0[3]) / (elen * e0len);
                double al = acos(cos_alpha);
                if (al < al0) {
                    al0 = al;
                    ial0 = ndata;
                }
                if ((al < al0_abs) && (p->inrel[i] == 0)) {
                    al0_abs = al;
                    ial0_abs = ndata;
                }
            }
        }
        ave /= p->lpoints[i];
        for (int j = 1; j <= p->lpoints[i]; j++) {
            k2++;
            p->sig[k2 - 1] = ave;
        }
    }
    /* weights, :416-427: (lc, weight) pairs until end of file */
    double weight_lc[PSO_MAX_LC + 1];
    for (int i = 1; i <= p->lcurves; i++) weight_lc[i] = -1;
    for (;;) {
        int lc;
        double w;
        if (!rd_int(&r, &lc)) break;
        if (!rd_double(&r, &w)) break;
        if (lc >= 1 && lc <= p->lcurves) weight_lc[lc] = w;
    }
    /* epoch, :431-442 */
    double jd_0 = p->jd_00;
    if (jd_0 <= 0) jd_0 = (int)jd_min;
    for (int i = 0; i < ndata; i++) p->tim[i] = p->tim[i] - jd_0;
    p->phi0 = p->phi0 * DEG2RAD;
    int k = 0;
    for (int i = 1; i <= p->lcurves; i++)
        for (int j = 1; j <= p->lpoints[i]; j++) {
            k++;
            p->weight[k - 1] = (weight_lc[i] == -1) ? 1 : weight_lc[i];
        }
    for (int i = 1; i <= 3; i++) p->weight[k + i - 1] = 1;
    if (onlyrel == 0) {
        al0 = al0_abs;
        ial0 = ial0_abs;
    }
    /* initial shape guess, :466-477 */
    double a = 1.05, b = 1.00, c_axis = 0.95;
    double rfit = sqrt(2 * p->sig[ial0 - 1] / (0.5 * PSO_PI * (1 + cos(al0))));
    double escl = rfit / sqrt((a * b + b * c_axis + a * c_axis) / 3);
    if (onlyrel == 0) escl *= 0.8;
    p->a_axis = a * escl;
    p->b_axis = b * escl;
    p->c_axis = c_axis * escl;
    p->escl = escl;
    p->ndata = ndata;
    p->jd_min = jd_min;
    p->jd_max = jd_max;
    p->jd_0 = jd_0;
    p->onlyrel = onlyrel;
    p->al0 = al0;
    p->ial0 = ial0;
    return 0;
}

int pso_parse_file(const char *path, pso_problem *p) {
    FILE *f = fopen(path, "rb");
    if (!f) return -4;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    if (!buf) {
        fclose(f);
        return -3;
    }
    size_t got = fread(buf, 1, (size_t)n, f);
    fclose(f);
    buf[got] = 0;
    int rc = pso_parse_text(buf, p);
    free(buf);
    return rc;
}

void pso_problem_free(pso_problem *p) {
    free(p->tim);
    free(p->brightness);
    free(p->sig);
    free(p->weight);
    free(p->ee);
    free(p->ee0);
    p->tim = p->brightness = p->sig = p->weight = p->ee = p->ee0 = NULL;
}

/* ------------------------------------------------------------------------------------------ */
/* Setup: trifac.c:10-82, areanorm.c:10-60, sphfunc.c:11-87, ellfit.c:11-88, ludcmp.c, lubksb.c */
/* ------------------------------------------------------------------------------------------ */
int pso_ncoef(int lmax, int mmax) { /* period_search_BOINC.cpp:498-505 */
    int n = 0;
    for (int m = 0; m <= mmax; m++)
        for (int l = m; l <= lmax; l++) {
            n++;
            if (m != 0) n++;
        }
    return n;
}

int pso_vertex_dirs(int nrows, double *t, double *f) { /* period_search_BOINC.cpp:507-539 */
    double dth = PSO_PI / (2 * nrows);
    int k = 1;
    t[1] = 0;
    f[1] = 0;
    for (int i = 1; i <= nrows; i++) {
        double dph = PSO_PI / (2 * i);
        for (int j = 0; j <= 4 * i - 1; j++) {
            k++;
            t[k] = i * dth;
            f[k] = j * dph;
        }
    }
    for (int i = nrows - 1; i >= 1; i--) {
        double dph = PSO_PI / (2 * i);
        for (int j = 0; j <= 4 * i - 1; j++) {
            k++;
            t[k] = PSO_PI - i * dth;
            f[k] = j * dph;
        }
    }
    int ndir = k + 1;
    t[ndir] = PSO_PI;
    f[ndir] = 0;
    return ndir;
}

void pso_trifac(int nrows, int *ifp) { /* trifac.c:10-82 */
    int rows = 2 * nrows + 1, cols = 4 * nrows + 1;
    int *nod = (int *)calloc((size_t)rows * cols, sizeof(int));
#define NOD(i, j) nod[(i) * cols + (j)]
#define IFP(t, v) ifp[((t) - 1) * 3 + (v) - 1]
    int nnod = 1;
    NOD(0, 0) = nnod;
    for (int i = 1; i <= nrows; i++)
        for (int j = 0; j <= 4 * i - 1; j++) {
            nnod++;
            NOD(i, j) = nnod;
            if (j == 0) NOD(i, 4 * i) = nnod;
        }
    for (int i = nrows - 1; i >= 1; i--)
        for (int j = 0; j <= 4 * i - 1; j++) {
            nnod++;
            NOD(2 * nrows - i, j) = nnod;
            if (j == 0) NOD(2 * nrows - i, 4 * i) = nnod;
        }
    NOD(2 * nrows, 0) = nnod + 1;
    int ntri = 0;
    for (int j1 = 1; j1 <= nrows; j1++)
        for (int j3 = 1; j3 <= 4; j3++) {
            int j0 = (j3 - 1) * j1;
            ntri++;
            IFP(ntri, 1) = NOD(j1 - 1, j0 - (j3 - 1));
            IFP(ntri, 2) = NOD(j1, j0);
            IFP(ntri, 3) = NOD(j1, j0 + 1);
            for (int j2 = j0 + 1; j2 <= j0 + j1 - 1; j2++) {
                ntri++;
                IFP(ntri, 1) = NOD(j1, j2);
                IFP(ntri, 2) = NOD(j1 - 1, j2 - (j3 - 1));
                IFP(ntri, 3) = NOD(j1 - 1, j2 - 1 - (j3 - 1));
                ntri++;
                IFP(ntri, 1) = NOD(j1 - 1, j2 - (j3 - 1));
                IFP(ntri, 2) = NOD(j1, j2);
                IFP(ntri, 3) = NOD(j1, j2 + 1);
            }
        }
    for (int j1 = nrows + 1; j1 <= 2 * nrows; j1++)
        for (int j3 = 1; j3 <= 4; j3++) {
            int j0 = (j3 - 1) * (2 * nrows - j1);
            ntri++;
            IFP(ntri, 1) = NOD(j1, j0);
            IFP(ntri, 2) = NOD(j1 - 1, j0 + 1 + (j3 - 1));
            IFP(ntri, 3) = NOD(j1 - 1, j0 + (j3 - 1));
            for (int j2 = j0 + 1; j2 <= j0 + (2 * nrows - j1); j2++) {
                ntri++;
                IFP(ntri, 1) = NOD(j1, j2);
                IFP(ntri, 2) = NOD(j1 - 1, j2 + (j3 - 1));
                IFP(ntri, 3) = NOD(j1, j2 - 1);
                ntri++;
                IFP(ntri, 1) = NOD(j1, j2);
                IFP(ntri, 2) = NOD(j1 - 1, j2 + 1 + (j3 - 1));
                IFP(ntri, 3) = NOD(j1 - 1, j2 + (j3 - 1));
            }
        }
#undef NOD
#undef IFP
    free(nod);
}

void pso_areanorm(const double *t, const double *f, int ndir, int nfac, const int *ifp, double *nor,
                  double *darea, double *at, double *af) { /* areanorm.c:10-60 */
    double *x = (double *)malloc((ndir + 1) * sizeof(double));
    double *y = (double *)malloc((ndir + 1) * sizeof(double));
    double *z = (double *)malloc((ndir + 1) * sizeof(double));
    for (int i = 1; i <= ndir; i++) {
        double st = sin(t[i]);
        x[i] = st * cos(f[i]);
        y[i] = st * sin(f[i]);
        z[i] = cos(t[i]);
    }
    for (int i = 0; i < nfac; i++) {
        double vx[4], vy[4], vz[4], c[4];
        int v1 = ifp[i * 3 + 0];
        for (int j = 2; j <= 3; j++) {
            int vj = ifp[i * 3 + j - 1];
            vx[j] = x[vj] - x[v1];
            vy[j] = y[vj] - y[v1];
            vz[j] = z[vj] - z[v1];
        }
        c[1] = vy[2] * vz[3] - vy[3] * vz[2];
        c[2] = vz[2] * vx[3] - vz[3] * vx[2];
        c[3] = vx[2] * vy[3] - vx[3] * vy[2];
        double clen2 = c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
        double clen = sqrt(clen2);
        nor[i * 3 + 0] = c[1] / clen;
        nor[i * 3 + 1] = c[2] / clen;
        nor[i * 3 + 2] = c[3] / clen;
        at[i] = acos(nor[i * 3 + 2]);
        af[i] = atan2(nor[i * 3 + 1], nor[i * 3 + 0]);
        darea[i] = 0.5 * clen;
    }
    free(x);
    free(y);
    free(z);
}

void pso_sphfunc(int nfac, const double *at, const double *af, int lmax, int mmax, double *pleg, double *fc,
                 double *fs, double *dsph) { /* sphfunc.c:11-87 */
    static double aleg[LM1][LM1][LM1];
    double ts[LM1], tc[LM1];
    int ncoef = pso_ncoef(lmax, mmax);
    for (int m = 0; m <= lmax; m++)
        for (int l = 0; l <= lmax; l++)
            for (int n = 0; n <= lmax; n++) aleg[n][l][m] = 0;
    aleg[0][0][0] = 1;
    if (lmax >= 1) aleg[1][1][0] = 1;
    for (int l = 1; l <= lmax; l++) aleg[0][l][l] = aleg[0][l - 1][l - 1] * (2 * l - 1);
    for (int m = 0; m <= mmax; m++)
        for (int l = m + 1; l <= lmax; l++) {
            int ibot;
            if ((2 * ((l - m) / 2)) == (l - m)) {
                aleg[0][l][m] = -(l + m - 1) * aleg[0][l - 2][m] / (1 * (l - m));
                ibot = 2;
            } else
                ibot = 1;
            if (l != 1)
                for (int n = ibot; n <= l - m; n = n + 2)
                    aleg[n][l][m] = ((2 * l - 1) * aleg[n - 1][l - 1][m] - (l + m - 1) * aleg[n][l - 2][m]) / (1 * (l - m));
        }
    for (int i = 0; i < nfac; i++) {
        ts[0] = 1;
        tc[0] = 1;
        if (lmax >= 1) {
            ts[1] = sin(at[i]);
            tc[1] = cos(at[i]);
        }
        for (int j = 2; j <= lmax; j++) {
            ts[j] = ts[1] * ts[j - 1];
            tc[j] = tc[1] * tc[j - 1];
        }
        fs[i * LM1 + 0] = 0;
        fc[i * LM1 + 0] = 1;
        for (int j = 1; j <= mmax; j++) {
            fs[i * LM1 + j] = sin(j * af[i]);
            fc[i * LM1 + j] = cos(j * af[i]);
        }
        int k = 0;
        for (int m = 0; m <= mmax; m++)
            for (int l = m; l <= lmax; l++) {
                double *pl = &pleg[(i * LM1 + l) * LM1 + m];
                *pl = 0;
                int ibot = ((2 * ((l - m) / 2)) == (l - m)) ? 0 : 1;
                for (int n = ibot; n <= l - m; n = n + 2) *pl = *pl + aleg[n][l][m] * tc[n] * ts[m];
                k++;
                dsph[i * ncoef + k - 1] = fc[i * LM1 + m] * *pl;
                if (m != 0) {
                    k++;
                    dsph[i * ncoef + k - 1] = fs[i * LM1 + m] * *pl;
                }
            }
    }
}

int pso_ludcmp(double *a, int n, int *indx, double *d) { /* ludcmp.c:9-68 */
#define A(i, j) a[(i) * (n + 1) + (j)]
    int imax = -999;
    double *v = (double *)malloc((n + 1) * sizeof(double));
    *d = 1.0;
    for (int i = 1; i <= n; i++) {
        double big = 0.0, temp;
        for (int j = 1; j <= n; j++)
            if ((temp = fabs(A(i, j))) > big) big = temp;
        if (big == 0.0) {
            free(v);
            return -1; /* the reference exits with status 4 here (ludcmp.c:23) */
        }
        v[i] = 1.0 / big;
    }
    for (int j = 1; j <= n; j++) {
        for (int i = 1; i < j; i++) {
            double sum = A(i, j);
            for (int k = 1; k < i; k++) sum -= A(i, k) * A(k, j);
            A(i, j) = sum;
        }
        double big = 0.0, dum;
        for (int i = j; i <= n; i++) {
            double sum = A(i, j);
            for (int k = 1; k < j; k++) sum -= A(i, k) * A(k, j);
            A(i, j) = sum;
            if ((dum = v[i] * fabs(sum)) >= big) {
                big = dum;
                imax = i;
            }
        }
        if (j != imax) {
            for (int k = 1; k <= n; k++) {
                dum = A(imax, k);
                A(imax, k) = A(j, k);
                A(j, k) = dum;
            }
            *d = -(*d);
            v[imax] = v[j];
        }
        indx[j] = imax;
        if (A(j, j) == 0.0) A(j, j) = 1.0e-20;
        if (j != n) {
            dum = 1.0 / (A(j, j));
            for (int i = j + 1; i <= n; i++) A(i, j) *= dum;
        }
    }
    free(v);
    return 0;
#undef A
}

void pso_lubksb(const double *a, int n, const int *indx, double *b) { /* lubksb.c:3-24 */
#define A(i, j) a[(i) * (n + 1) + (j)]
    int ii = 0;
    for (int i = 1; i <= n; i++) {
        int ip = indx[i];
        double sum = b[ip];
        b[ip] = b[i];
        if (ii)
            for (int j = ii; j <= i - 1; j++) sum -= A(i, j) * b[j];
        else if (sum)
            ii = i;
        b[i] = sum;
    }
    for (int i = n; i >= 1; i--) {
        double sum = b[i];
        for (int j = i + 1; j <= n; j++) sum -= A(i, j) * b[j];
        b[i] = sum / A(i, i);
    }
#undef A
}

int pso_ellfit(double *cg, double a, double b, double c, int nfac, int ncoef, const double *at, const double *af,
               const double *pleg, int lmax, int mmax) { /* ellfit.c:11-88 */
    int *indx = (int *)malloc((ncoef + 1) * sizeof(int));
    double *fitvec = (double *)malloc((ncoef + 1) * sizeof(double));
    double *er = (double *)malloc((nfac + 1) * sizeof(double));
    double *fmat = (double *)malloc((size_t)(nfac + 1) * (ncoef + 1) * sizeof(double));
    double *fitmat = (double *)malloc((size_t)(ncoef + 1) * (ncoef + 1) * sizeof(double));
    double d;
#define FM(i, j) fmat[(i) * (ncoef + 1) + (j)]
#define FT(i, j) fitmat[(i) * (ncoef + 1) + (j)]
    for (int i = 1; i <= nfac; i++) {
        /* the reference build (g++ -O3) evaluates this loop's sin/cos pairs with sincos() */
        double st, ct, sf, cf;
        sincos(at[i - 1], &st, &ct);
        sincos(af[i - 1], &sf, &cf);
        double u = a * st * cf, v = b * st * sf, w = c * ct;
        double sum = u * u + v * v + w * w;
        er[i] = 2 * (log(a * b * c) - log(sum));
    }
    for (int i = 1; i <= nfac; i++) {
        int n = 0;
        for (int m = 0; m <= mmax; m++)
            for (int l = m; l <= lmax; l++) {
                double pl = pleg[((i - 1) * LM1 + l) * LM1 + m];
                n++;
                if (m != 0) {
                    FM(i, n) = pl * cos(m * af[i - 1]);
                    n++;
                    FM(i, n) = pl * sin(m * af[i - 1]);
                } else
                    FM(i, n) = pl;
            }
    }
    for (int i = 1; i <= ncoef; i++) {
        for (int j = 1; j <= ncoef; j++) {
            FT(i, j) = 0;
            for (int k = 1; k <= nfac; k++) FT(i, j) = FT(i, j) + FM(k, i) * FM(k, j);
        }
        fitvec[i] = 0;
        for (int j = 1; j <= nfac; j++) fitvec[i] = fitvec[i] + FM(j, i) * er[j];
    }
    int rc = pso_ludcmp(fitmat, ncoef, indx, &d);
    if (rc == 0) {
        pso_lubksb(fitmat, ncoef, indx, fitvec);
        for (int i = 1; i <= ncoef; i++) cg[i] = fitvec[i];
    }
#undef FM
#undef FT
    free(indx);
    free(fitvec);
    free(er);
    free(fmat);
    free(fitmat);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* Context                                                                                      */
/* ------------------------------------------------------------------------------------------ */
struct pso_ctx {
    int nfac, ncoef, ma, mfit, nphpar, lmax, mmax;
    double *nor, *darea, *pleg, *fc, *fs, *dsph, *area, *dg;
    double blmat[4][4], dblm[3][4][4];
    double scale, phi0;
    int ia[PSO_MAX_N_PAR + 2];
    double cg_first[PSO_MAX_N_PAR + 2];
    double freq_start, freq_end, freq_step;
    /* data; the last (convexity) lightcurve is lightcurve number lcurves_tot */
    int lcurves_tot;
    int lpoints[PSO_MAX_LC + 2], inrel[PSO_MAX_LC + 2];
    int ndata_obs, ndata; /* ndata = ndata_obs + 3 while the conv rows are present */
    double *tim, *y, *sig, *weight, *ee, *ee0;
    double conw_r, escl;
    double par[4];
    int ia_par[4];
    double cl, stop_condition, alamda_incr, alamda_start;
    int n_iter_min, n_start_max;
    /* Levenberg-Marquardt state (globals Alamda, Ochisq, Chisq of the reference) */
    double alamda, ochisq, chisq;
    int lm_mfit;
    double *alpha, *covar, *beta, *da, *atry, *dytemp, *ytemp, *dyda, *dave;
    int maxlp;
    double chck[4];
};

static const double beta_pole[PSO_N_POLES + 1] = {0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60};      /* :197-206 */
static const double lambda_pole[PSO_N_POLES + 1] = {0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300};

pso_ctx *pso_ctx_create(const pso_problem *p) {
    pso_ctx *c = (pso_ctx *)calloc(1, sizeof(pso_ctx));
    int nrows = p->nrows;
    c->lmax = p->lmax;
    c->mmax = p->mmax;
    c->ncoef = pso_ncoef(p->lmax, p->mmax);
    c->nfac = 8 * nrows * nrows;
    c->nphpar = 3;
    c->ma = c->ncoef + 5 + c->nphpar;
    if (c->ncoef + 3 + c->nphpar + 1 > PSO_MAX_N_PAR) {
        free(c);
        return NULL;
    }
    int nfac = c->nfac;
    c->nor = (double *)malloc(nfac * 3 * sizeof(double));
    c->darea = (double *)malloc(nfac * sizeof(double));
    c->pleg = (double *)calloc((size_t)nfac * LM1 * LM1, sizeof(double));
    c->fc = (double *)calloc((size_t)nfac * LM1, sizeof(double));
    c->fs = (double *)calloc((size_t)nfac * LM1, sizeof(double));
    c->dsph = (double *)calloc((size_t)nfac * c->ncoef, sizeof(double));
    c->area = (double *)calloc(nfac, sizeof(double));
    c->dg = (double *)calloc((size_t)nfac * c->ncoef, sizeof(double));
    double *t = (double *)malloc((4 * nrows * nrows + 4) * sizeof(double));
    double *f = (double *)malloc((4 * nrows * nrows + 4) * sizeof(double));
    int *ifp = (int *)malloc(nfac * 3 * sizeof(int));
    double *at = (double *)malloc(nfac * sizeof(double));
    double *af = (double *)malloc(nfac * sizeof(double));
    int ndir = pso_vertex_dirs(nrows, t, f);
    pso_trifac(nrows, ifp);
    pso_areanorm(t, f, ndir, nfac, ifp, c->nor, c->darea, at, af);
    pso_sphfunc(nfac, at, af, p->lmax, p->mmax, c->pleg, c->fc, c->fs, c->dsph);
    int rc = pso_ellfit(c->cg_first, p->a_axis, p->b_axis, p->c_axis, nfac, c->ncoef, at, af, c->pleg, p->lmax,
                        p->mmax);
    free(t);
    free(f);
    free(ifp);
    free(at);
    free(af);
    if (rc != 0) {
        pso_ctx_free(c);
        return NULL;
    }
    /* frequency grid, :784-786 */
    c->freq_start = 1 / p->per_start;
    c->freq_end = 1 / p->per_end;
    c->freq_step = 0.5 / (p->jd_max - p->jd_min) / 24 * p->per_step_coef;
    /* free/fixed mask, :788-802 */
    int ncoef = c->ncoef, nph = c->nphpar;
    c->ia[ncoef + 1] = 1;
    c->ia[ncoef + 2] = 1;
    c->ia[ncoef + 3] = p->ia_prd;
    for (int i = 1; i <= ncoef; i++) c->ia[i] = 1;
    if (p->onlyrel == 1) c->ia[1] = 0;
    for (int i = 1; i <= nph; i++) c->ia[ncoef + 3 + i] = p->ia_par[i];
    c->ia[ncoef + 3 + nph + 1] = p->ia_cl;
    c->ia[ncoef + 3 + nph + 2] = 0;
    c->mfit = 0;
    for (int j = 1; j <= c->ma; j++)
        if (c->ia[j]) c->mfit++;
    for (int i = 1; i <= 3; i++) {
        c->par[i] = p->par[i];
        c->ia_par[i] = p->ia_par[i];
    }
    c->cl = p->cl;
    c->stop_condition = p->stop_condition;
    c->n_iter_min = p->n_iter_min;
    c->alamda_incr = p->alamda_incr;
    c->alamda_start = p->alamda_start;
    c->phi0 = p->phi0;
    c->escl = p->escl;
    /* data with three extra rows for the convexity pseudo-lightcurve, :479-485 */
    c->ndata_obs = p->ndata;
    c->ndata = p->ndata;
    int nd = p->ndata + 3;
    c->tim = (double *)calloc(nd, sizeof(double));
    c->y = (double *)calloc(nd, sizeof(double));
    c->sig = (double *)calloc(nd, sizeof(double));
    c->weight = (double *)calloc(nd, sizeof(double));
    c->ee = (double *)calloc(nd * 3, sizeof(double));
    c->ee0 = (double *)calloc(nd * 3, sizeof(double));
    memcpy(c->tim, p->tim, p->ndata * sizeof(double));
    memcpy(c->y, p->brightness, p->ndata * sizeof(double));
    memcpy(c->sig, p->sig, p->ndata * sizeof(double));
    memcpy(c->weight, p->weight, nd * sizeof(double));
    memcpy(c->ee, p->ee, p->ndata * 3 * sizeof(double));
    memcpy(c->ee0, p->ee0, p->ndata * 3 * sizeof(double));
    c->lcurves_tot = p->lcurves + 1;
    c->maxlp = 3;
    for (int i = 1; i <= p->lcurves; i++) {
        c->lpoints[i] = p->lpoints[i];
        c->inrel[i] = p->inrel[i];
        if (p->lpoints[i] > c->maxlp) c->maxlp = p->lpoints[i];
    }
    c->lpoints[c->lcurves_tot] = 3;
    c->inrel[c->lcurves_tot] = 0;
    c->conw_r = p->conw / p->escl / p->escl;
    /* LM scratch */
    int mf = c->mfit, ma = c->ma;
    c->alpha = (double *)calloc((size_t)(mf + 1) * (mf + 1), sizeof(double));
    c->covar = (double *)calloc((size_t)(mf + 1) * (mf + 1), sizeof(double));
    c->beta = (double *)calloc(ma + 1, sizeof(double));
    c->da = (double *)calloc(ma + 1, sizeof(double));
    c->atry = (double *)calloc(ma + 1, sizeof(double));
    c->dyda = (double *)calloc(ma + 1, sizeof(double));
    c->dave = (double *)calloc(ma + 1, sizeof(double));
    c->ytemp = (double *)calloc(c->maxlp + 1, sizeof(double));
    c->dytemp = (double *)calloc((size_t)(c->maxlp + 1) * (ma + 1), sizeof(double));
    return c;
}

static pso_ctx *ctx_from_problem(pso_problem *p, int rc, int *err) {
    if (err) *err = rc;
    if (rc != 0) {
        pso_problem_free(p);
        return NULL;
    }
    pso_ctx *c = pso_ctx_create(p);
    pso_problem_free(p);
    if (!c && err) *err = -5;
    return c;
}
pso_ctx *pso_ctx_from_file(const char *path, int *err) {
    pso_problem p;
    int rc = pso_parse_file(path, &p);
    return ctx_from_problem(&p, rc, err);
}
pso_ctx *pso_ctx_from_text(const char *text, int *err) {
    pso_problem p;
    int rc = pso_parse_text(text, &p);
    return ctx_from_problem(&p, rc, err);
}
double pso_ctx_escl(const pso_ctx *c) { return c->escl; }
double pso_ctx_conw_r(const pso_ctx *c) { return c->conw_r; }
int pso_ctx_ndata(const pso_ctx *c) { return c->ndata_obs; }
int pso_ctx_stop_iters(const pso_ctx *c) { return c->stop_condition > 1 ? (int)c->stop_condition : 0; }

void pso_ctx_free(pso_ctx *c) {
    if (!c) return;
    free(c->nor);
    free(c->darea);
    free(c->pleg);
    free(c->fc);
    free(c->fs);
    free(c->dsph);
    free(c->area);
    free(c->dg);
    free(c->tim);
    free(c->y);
    free(c->sig);
    free(c->weight);
    free(c->ee);
    free(c->ee0);
    free(c->alpha);
    free(c->covar);
    free(c->beta);
    free(c->da);
    free(c->atry);
    free(c->dyda);
    free(c->dave);
    free(c->ytemp);
    free(c->dytemp);
    free(c);
}

int pso_ctx_ma(const pso_ctx *c) { return c->ma; }
int pso_ctx_ncoef(const pso_ctx *c) { return c->ncoef; }
int pso_ctx_nfac(const pso_ctx *c) { return c->nfac; }
int pso_ctx_mfit(const pso_ctx *c) { return c->mfit; }
void pso_ctx_ia(const pso_ctx *c, int *ia) {
    for (int i = 1; i <= c->ma; i++) ia[i] = c->ia[i];
}
void pso_ctx_cg_first(const pso_ctx *c, double *cg) {
    for (int i = 1; i <= c->ncoef; i++) cg[i] = c->cg_first[i];
}
void pso_ctx_geometry(const pso_ctx *c, double *nor, double *darea, double *dsph) {
    if (nor) memcpy(nor, c->nor, c->nfac * 3 * sizeof(double));
    if (darea) memcpy(darea, c->darea, c->nfac * sizeof(double));
    if (dsph) memcpy(dsph, c->dsph, (size_t)c->nfac * c->ncoef * sizeof(double));
}
double pso_ctx_freq(const pso_ctx *c, double *f0, double *f1, double *df, int *nmax) {
    if (f0) *f0 = c->freq_start;
    if (f1) *f1 = c->freq_end;
    if (df) *df = c->freq_step;
    if (nmax) *nmax = (int)((c->freq_start - c->freq_end) / c->freq_step) + 1; /* :807 */
    return c->freq_step;
}
void pso_ctx_set_conw(pso_ctx *c, double conw_r) { /* :491-496, :721-726 */
    c->conw_r = conw_r;
    for (int j = 0; j < 3; j++) {
        c->y[c->ndata_obs + j] = 0;
        c->sig[c->ndata_obs + j] = 1 / conw_r;
    }
    c->ndata = c->ndata_obs + 3;
}
void pso_ctx_init_params(const pso_ctx *c, double freq, int m, double *cg) { /* :819-844 */
    int ncoef = c->ncoef, nph = c->nphpar;
    double prd = 1 / freq;
    for (int i = 1; i <= ncoef; i++) cg[i] = c->cg_first[i];
    cg[ncoef + 1] = beta_pole[m];
    cg[ncoef + 2] = lambda_pole[m];
    cg[ncoef + 1] = 90 - cg[ncoef + 1];
    cg[ncoef + 1] = DEG2RAD * cg[ncoef + 1];
    cg[ncoef + 2] = DEG2RAD * cg[ncoef + 2];
    cg[ncoef + 3] = 24 * 2 * PSO_PI / prd;
    for (int i = 1; i <= nph; i++) cg[ncoef + 3 + i] = c->par[i];
    cg[ncoef + 3 + nph + 2] = 1;
    cg[ncoef + 3 + nph + 1] = log(c->cl);
}

/* ------------------------------------------------------------------------------------------ */
/* Model: curv.c:10-39, blmatrix.c:9-46, phasec.c:11-24, matrix.c:11-58, bright.c:14-126,       */
/* conv.c:12-29                                                                                 */
/* ------------------------------------------------------------------------------------------ */
void pso_curv(pso_ctx *c, const double *cg) {
    int nc = c->ncoef;
    for (int i = 0; i < c->nfac; i++) {
        double g = 0;
        int n = 0;
        for (int m = 0; m <= c->mmax; m++)
            for (int l = m; l <= c->lmax; l++) {
                n++;
                double fsum = cg[n] * c->fc[i * LM1 + m];
                if (m != 0) {
                    n++;
                    fsum = fsum + cg[n] * c->fs[i * LM1 + m];
                }
                g = g + c->pleg[(i * LM1 + l) * LM1 + m] * fsum;
            }
        g = exp(g);
        c->area[i] = c->darea[i] * g;
        for (int k = 1; k <= n; k++) c->dg[i * nc + k - 1] = g * c->dsph[i * nc + k - 1];
    }
}

void pso_blmatrix(pso_ctx *c, double bet, double lam) {
    double cb, sb, cl, sl; /* sincos(), as the reference build evaluates these pairs */
    sincos(bet, &sb, &cb);
    sincos(lam, &sl, &cl);
    c->blmat[1][1] = cb * cl;
    c->blmat[1][2] = cb * sl;
    c->blmat[1][3] = -sb;
    c->blmat[2][1] = -sl;
    c->blmat[2][2] = cl;
    c->blmat[2][3] = 0;
    c->blmat[3][1] = sb * cl;
    c->blmat[3][2] = sb * sl;
    c->blmat[3][3] = cb;
    c->dblm[1][1][1] = -sb * cl;
    c->dblm[1][1][2] = -sb * sl;
    c->dblm[1][1][3] = -cb;
    c->dblm[1][2][1] = 0;
    c->dblm[1][2][2] = 0;
    c->dblm[1][2][3] = 0;
    c->dblm[1][3][1] = cb * cl;
    c->dblm[1][3][2] = cb * sl;
    c->dblm[1][3][3] = -sb;
    c->dblm[2][1][1] = -cb * sl;
    c->dblm[2][1][2] = cb * cl;
    c->dblm[2][1][3] = 0;
    c->dblm[2][2][1] = -cl;
    c->dblm[2][2][2] = -sl;
    c->dblm[2][2][3] = 0;
    c->dblm[2][3][1] = -sb * sl;
    c->dblm[2][3][2] = sb * cl;
    c->dblm[2][3][3] = 0;
}

double pso_phasec(double *dcdp, double alpha, const double *p) {
    double e = exp(-alpha / p[2]);
    double cc = 1 + p[1] * e + p[3] * alpha;
    dcdp[1] = e;
    dcdp[2] = p[1] * e * alpha / (p[2] * p[2]);
    dcdp[3] = alpha;
    return cc;
}

void pso_matrix(const pso_ctx *c, double omg, double t, double tmat[4][4], double dtm[4][4][4]) {
    double f = omg * t + c->phi0;
    f = fmod(f, 2 * PSO_PI);
    double cf, sf; /* sincos(), as the reference build evaluates this pair */
    sincos(f, &sf, &cf);
    double fmat[4][4], dfm[4][4];
    fmat[1][1] = cf;
    fmat[1][2] = sf;
    fmat[1][3] = 0;
    fmat[2][1] = -sf;
    fmat[2][2] = cf;
    fmat[2][3] = 0;
    fmat[3][1] = 0;
    fmat[3][2] = 0;
    fmat[3][3] = 1;
    dfm[1][1] = -t * sf;
    dfm[1][2] = t * cf;
    dfm[1][3] = 0;
    dfm[2][1] = -t * cf;
    dfm[2][2] = -t * sf;
    dfm[2][3] = 0;
    dfm[3][1] = 0;
    dfm[3][2] = 0;
    dfm[3][3] = 0;
    for (int i = 1; i <= 3; i++)
        for (int j = 1; j <= 3; j++) {
            tmat[i][j] = 0;
            dtm[1][i][j] = 0;
            dtm[2][i][j] = 0;
            dtm[3][i][j] = 0;
            for (int k = 1; k <= 3; k++) {
                tmat[i][j] = tmat[i][j] + fmat[i][k] * c->blmat[k][j];
                dtm[1][i][j] = dtm[1][i][j] + fmat[i][k] * c->dblm[1][k][j];
                dtm[2][i][j] = dtm[2][i][j] + fmat[i][k] * c->dblm[2][k][j];
                dtm[3][i][j] = dtm[3][i][j] + dfm[i][k] * c->blmat[k][j];
            }
        }
}

double pso_bright(pso_ctx *c, const double *ee, const double *ee0, double t, const double *cg, double *dyda,
                  int ncoef) {
    /* ee, ee0 are 1-based 3-vectors here, as in the reference */
    int nf = c->nfac, ncg = c->ncoef, nph = c->nphpar;
    int ncoef0 = ncoef - 2 - nph;
    double cl = exp(cg[ncoef - 1]);
    double cls = cg[ncoef];
    double cos_alpha = ee[1] * ee0[1] + ee[2] * ee0[2] + ee[3] * ee0[3];
    double alpha = acos(cos_alpha);
    double php[4], dphp[4];
    for (int i = 1; i <= nph; i++) php[i] = cg[ncoef0 + i];
    c->scale = pso_phasec(dphp, alpha, php);
    double tmat[4][4], dtm[4][4][4], e[4], e0[4], de[4][4], de0[4][4];
    pso_matrix(c, cg[ncoef0], t, tmat, dtm);
    double br = 0;
    for (int i = 1; i <= 3; i++) {
        e[i] = 0;
        e0[i] = 0;
        for (int j = 1; j <= 3; j++) {
            e[i] = e[i] + tmat[i][j] * ee[j];
            e0[i] = e0[i] + tmat[i][j] * ee0[j];
            de[i][j] = 0;
            de0[i][j] = 0;
            for (int k = 1; k <= 3; k++) {
                de[i][j] = de[i][j] + dtm[j][i][k] * ee[k];
                de0[i][j] = de0[i][j] + dtm[j][i][k] * ee0[k];
            }
        }
    }
    static int incl[PSO_MAX_N_FAC];
    static double mu[PSO_MAX_N_FAC], mu0[PSO_MAX_N_FAC], s[PSO_MAX_N_FAC], dbr[PSO_MAX_N_FAC],
        dsmu[PSO_MAX_N_FAC], dsmu0[PSO_MAX_N_FAC];
    for (int i = 0; i < nf; i++) {
        const double *n = &c->nor[i * 3];
        incl[i] = 0;
        mu[i] = e[1] * n[0] + e[2] * n[1] + e[3] * n[2];
        mu0[i] = e0[1] * n[0] + e0[2] * n[1] + e0[3] * n[2];
        if ((mu[i] > PSO_TINY) && (mu0[i] > PSO_TINY)) {
            incl[i] = 1;
            double dnom = mu[i] + mu0[i];
            s[i] = mu[i] * mu0[i] * (cl + cls / dnom);
            br = br + c->area[i] * s[i];
            double q0 = mu0[i] / dnom, q = mu[i] / dnom;
            dsmu[i] = cls * (q0 * q0) + cl * mu0[i];
            dsmu0[i] = cls * (q * q) + cl * mu[i];
            dbr[i] = c->darea[i] * s[i];
        }
    }
    for (int i = 1; i <= ncoef0 - 3; i++) {
        double acc = 0;
        for (int j = 0; j < nf; j++)
            if (incl[j] == 1) acc = acc + dbr[j] * c->dg[j * ncg + i - 1];
        dyda[i] = c->scale * acc;
    }
    for (int k = 1; k <= 3; k++) {
        double acc = 0;
        for (int i = 0; i < nf; i++)
            if (incl[i] == 1) {
                const double *n = &c->nor[i * 3];
                double sum = 0, sum0 = 0;
                for (int j = 1; j <= 3; j++) {
                    sum = sum + n[j - 1] * de[j][k];
                    sum0 = sum0 + n[j - 1] * de0[j][k];
                }
                acc = acc + c->area[i] * (dsmu[i] * sum + dsmu0[i] * sum0);
            }
        dyda[ncoef0 - 3 + k] = c->scale * acc;
    }
    for (int i = 1; i <= nph; i++) dyda[ncoef0 + i] = br * dphp[i];
    double d1 = 0, d2 = 0;
    for (int i = 0; i < nf; i++)
        if (incl[i] == 1) {
            d1 = d1 + mu[i] * mu0[i] * c->area[i];
            d2 = d2 + c->area[i] * mu[i] * mu0[i] / (mu[i] + mu0[i]);
        }
    dyda[ncoef - 1] = c->scale * d1 * cl;
    dyda[ncoef] = c->scale * d2;
    br *= c->scale;
    return br;
}

double pso_conv(pso_ctx *c, int nc, double *dres, int ma) {
    double res = 0;
    int ncg = c->ncoef;
    for (int j = 1; j <= ma; j++) dres[j] = 0;
    for (int i = 0; i < c->nfac; i++) {
        double n = c->nor[i * 3 + nc - 1];
        res += c->area[i] * n;
        for (int j = 1; j <= ncg; j++) dres[j] += c->darea[i] * c->dg[i * ncg + j - 1] * n;
    }
    return res;
}

void pso_ctx_set_area_dg(pso_ctx *c, const double *area, const double *dg) {
    memcpy(c->area, area, c->nfac * sizeof(double));
    memcpy(c->dg, dg, (size_t)c->nfac * c->ncoef * sizeof(double));
}
void pso_ctx_get_area_dg(const pso_ctx *c, double *area, double *dg) {
    if (area) memcpy(area, c->area, c->nfac * sizeof(double));
    if (dg) memcpy(dg, c->dg, (size_t)c->nfac * c->ncoef * sizeof(double));
}
void pso_ctx_set_blmat(pso_ctx *c, const double *b9, const double *d18) {
    for (int i = 1; i <= 3; i++)
        for (int j = 1; j <= 3; j++) {
            c->blmat[i][j] = b9[(i - 1) * 3 + j - 1];
            c->dblm[1][i][j] = d18[(i - 1) * 3 + j - 1];
            c->dblm[2][i][j] = d18[9 + (i - 1) * 3 + j - 1];
        }
}

/* ------------------------------------------------------------------------------------------ */
/* Normal equations: mrqcof.c:16-148                                                            */
/* ------------------------------------------------------------------------------------------ */
double pso_mrqcof(pso_ctx *c, const double *a, const int *ia, int ma, double *alpha, double *beta) {
    int nph = c->nphpar, mfit = 0;
    pso_curv(c, a);
    pso_blmatrix(c, a[ma - 4 - nph], a[ma - 3 - nph]);
    for (int j = 1; j <= ma; j++)
        if (ia[j]) mfit++;
#define AL(j, k) alpha[(j) * (mfit + 1) + (k)]
    for (int j = 1; j <= mfit; j++) {
        for (int k = 1; k <= j; k++) AL(j, k) = 0;
        beta[j] = 0;
    }
    double trial_chisq = 0, ave = 0;
    int np = 0, np1 = 0, np2 = 0;
    double *dyda = c->dyda, *dave = c->dave, *ytemp = c->ytemp;
    int lc_tot = c->lcurves_tot;
    for (int i = 1; i <= lc_tot; i++) {
        if (c->inrel[i] == 1) {
            ave = 0;
            for (int l = 1; l <= ma; l++) dave[l] = 0;
        }
        for (int jp = 1; jp <= c->lpoints[i]; jp++) {
            np++;
            double ymod;
            if (i < lc_tot) {
                double xx1[4], xx2[4];
                for (int ic = 1; ic <= 3; ic++) {
                    xx1[ic] = c->ee[(np - 1) * 3 + ic - 1];
                    xx2[ic] = c->ee0[(np - 1) * 3 + ic - 1];
                }
                ymod = pso_bright(c, xx1, xx2, c->tim[np - 1], a, dyda, ma);
            } else
                ymod = pso_conv(c, jp, dyda, ma);
            ytemp[jp] = ymod;
            if (c->inrel[i] == 1) ave = ave + ymod;
            for (int l = 1; l <= ma; l++) {
                c->dytemp[jp * (ma + 1) + l] = dyda[l];
                if (c->inrel[i] == 1) dave[l] = dave[l] + dyda[l];
            }
        }
        for (int jp = 1; jp <= c->lpoints[i]; jp++) {
            np1++;
            if (c->inrel[i] == 1) {
                double coef = c->sig[np1 - 1] * c->lpoints[i] / ave;
                for (int l = 1; l <= ma; l++)
                    c->dytemp[jp * (ma + 1) + l] = coef * (c->dytemp[jp * (ma + 1) + l] - ytemp[jp] * dave[l] / ave);
                ytemp[jp] = coef * ytemp[jp];
                c->dytemp[jp * (ma + 1) + 1] = 0;
            }
        }
        for (int jp = 1; jp <= c->lpoints[i]; jp++) {
            double ymod = ytemp[jp];
            for (int l = 1; l <= ma; l++) dyda[l] = c->dytemp[jp * (ma + 1) + l];
            np2++;
            double sig2i = 1 / (c->sig[np2 - 1] * c->sig[np2 - 1]);
            double wght = c->weight[np2 - 1];
            double dy = c->y[np2 - 1] - ymod;
            int j = 0;
            for (int l = 1; l <= ma; l++) {
                if (ia[l]) {
                    j++;
                    double wt = dyda[l] * sig2i * wght;
                    int k = 0;
                    for (int m = 1; m <= l; m++) {
                        if (ia[m]) {
                            k++;
                            AL(j, k) = AL(j, k) + wt * dyda[m];
                        }
                    }
                    beta[j] = beta[j] + dy * wt;
                }
            }
            trial_chisq = trial_chisq + dy * dy * sig2i * wght;
        }
    }
    for (int j = 2; j <= mfit; j++)
        for (int k = 1; k <= j - 1; k++) AL(k, j) = AL(j, k);
#undef AL
    return trial_chisq;
}

/* ------------------------------------------------------------------------------------------ */
/* Full-pivot Gauss-Jordan: gauss_errc.c:9-73 (Numerical Recipes gaussj, one right-hand side)   */
/* ------------------------------------------------------------------------------------------ */
int pso_gauss_errc(double *a, int n, double *b) {
#define A(i, j) a[(i) * (n + 1) + (j)]
    int *indxc = (int *)malloc((n + 1) * sizeof(int));
    int *indxr = (int *)malloc((n + 1) * sizeof(int));
    int *ipiv = (int *)malloc((n + 1) * sizeof(int));
    int icol = 0, irow = 0, rc = 0;
    for (int j = 1; j <= n; j++) ipiv[j] = 0;
    for (int i = 1; i <= n && rc == 0; i++) {
        double big = 0.0;
        for (int j = 1; j <= n && rc == 0; j++)
            if (ipiv[j] != 1)
                for (int k = 1; k <= n; k++) {
                    if (ipiv[k] == 0) {
                        if (fabs(A(j, k)) >= big) {
                            big = fabs(A(j, k));
                            irow = j;
                            icol = k;
                        }
                    } else if (ipiv[k] > 1) {
                        rc = 1;
                        break;
                    }
                }
        if (rc) break;
        ++(ipiv[icol]);
        if (irow != icol) {
            for (int l = 1; l <= n; l++) {
                double tmp = A(irow, l);
                A(irow, l) = A(icol, l);
                A(icol, l) = tmp;
            }
            double tmp = b[irow];
            b[irow] = b[icol];
            b[icol] = tmp;
        }
        indxr[i] = irow;
        indxc[i] = icol;
        if (A(icol, icol) == 0.0) {
            rc = 2;
            break;
        }
        double pivinv = 1.0 / A(icol, icol);
        A(icol, icol) = 1.0;
        for (int l = 1; l <= n; l++) A(icol, l) *= pivinv;
        b[icol] *= pivinv;
        for (int ll = 1; ll <= n; ll++)
            if (ll != icol) {
                double dum = A(ll, icol);
                A(ll, icol) = 0.0;
                for (int l = 1; l <= n; l++) A(ll, l) -= A(icol, l) * dum;
                b[ll] -= b[icol] * dum;
            }
    }
    if (rc == 0)
        for (int l = n; l >= 1; l--) {
            if (indxr[l] != indxc[l])
                for (int k = 1; k <= n; k++) {
                    double tmp = A(k, indxr[l]);
                    A(k, indxr[l]) = A(k, indxc[l]);
                    A(k, indxc[l]) = tmp;
                }
        }
    free(ipiv);
    free(indxc);
    free(indxr);
    return rc;
#undef A
}

/* ------------------------------------------------------------------------------------------ */
/* One LM step: mrqmin.c:11-110 (the Lastcall != 1 path used by the scan)                      */
/* ------------------------------------------------------------------------------------------ */
static int pso_mrqmin(pso_ctx *c, double *a) {
    int ma = c->ma, *ia = c->ia;
    int mf = c->mfit;
#define AL(j, k) c->alpha[(j) * (mf + 1) + (k)]
#define CV(j, k) c->covar[(j) * (mf + 1) + (k)]
    if (c->alamda < 0) {
        c->lm_mfit = mf;
        c->alamda = c->alamda_start;
        double temp = pso_mrqcof(c, a, ia, ma, c->alpha, c->beta);
        c->ochisq = temp;
        for (int j = 1; j <= ma; j++) c->atry[j] = a[j];
    }
    for (int j = 1; j <= mf; j++) {
        for (int k = 1; k <= mf; k++) CV(j, k) = AL(j, k);
        CV(j, j) = AL(j, j) * (1 + c->alamda);
        c->da[j] = c->beta[j];
    }
    int err = pso_gauss_errc(c->covar, mf, c->da);
    if (err != 0) return err;
    int j = 0;
    for (int l = 1; l <= ma; l++)
        if (ia[l]) {
            j++;
            c->atry[l] = a[l] + c->da[j];
        }
    double temp = pso_mrqcof(c, c->atry, ia, ma, c->covar, c->da);
    c->chisq = temp;
    if (temp < c->ochisq) {
        c->alamda = c->alamda / c->alamda_incr;
        for (int jj = 1; jj <= mf; jj++) {
            for (int k = 1; k <= mf; k++) AL(jj, k) = CV(jj, k);
            c->beta[jj] = c->da[jj];
        }
        for (int l = 1; l <= ma; l++) a[l] = c->atry[l];
    } else {
        c->alamda = c->alamda_incr * c->alamda;
        c->chisq = c->ochisq;
    }
    return 0;
#undef AL
#undef CV
}

/* ------------------------------------------------------------------------------------------ */
/* Driver: one (period, pole) fit, period_search_BOINC.cpp:819-911                             */
/* ------------------------------------------------------------------------------------------ */
int pso_fit(pso_ctx *c, double freq, int m, pso_fit_result *r) {
    double cg[PSO_MAX_N_PAR + 2];
    int ncoef = c->ncoef;
    pso_ctx_init_params(c, freq, m, cg);
    int n_iter_max = 0;
    double iter_diff_max = -1, rchisq = -1;
    if (c->stop_condition > 1) {
        n_iter_max = (int)c->stop_condition;
        iter_diff_max = 0;
        c->n_iter_min = 0; /* the reference overwrites its n_iter_min here (:854) */
    }
    if (c->stop_condition < 1) {
        n_iter_max = PSO_MAX_N_ITER;
        iter_diff_max = c->stop_condition;
    }
    c->alamda = -1;
    int niter = 0;
    double iter_diff = 1e40, dev_old = 1e30, dev_new = 0;
    while (((niter < n_iter_max) && (iter_diff > iter_diff_max)) || (niter < c->n_iter_min)) {
        pso_mrqmin(c, cg);
        niter++;
        if ((niter == 1) || (c->chisq < c->ochisq)) {
            c->ochisq = c->chisq;
            pso_curv(c, cg);
            for (int i = 1; i <= 3; i++) {
                c->chck[i] = 0;
                for (int j = 0; j < c->nfac; j++) c->chck[i] = c->chck[i] + c->area[j] * c->nor[j * 3 + i - 1];
            }
            rchisq = c->chisq - (c->chck[1] * c->chck[1] + c->chck[2] * c->chck[2] + c->chck[3] * c->chck[3]) *
                                    (c->conw_r * c->conw_r);
        }
        dev_new = sqrt(rchisq / (c->ndata - 3));
        if (dev_old - dev_new > 1e-10) {
            iter_diff = dev_old - dev_new;
            dev_old = dev_new;
        }
    }
    double totarea = 0;
    for (int i = 0; i < c->nfac; i++) totarea = totarea + c->area[i];
    double sum = c->chck[1] * c->chck[1] + c->chck[2] * c->chck[2] + c->chck[3] * c->chck[3];
    r->dev_new = dev_new;
    r->dark = sqrt(sum);
    r->totarea = totarea;
    r->prd = 2 * PSO_PI / cg[ncoef + 3];
    r->la = RAD2DEG * cg[ncoef + 2];
    r->be = 90 - RAD2DEG * cg[ncoef + 1];
    r->chisq = c->chisq;
    r->niter = niter;
    return 0;
}

/* Best pole of one trial period (:815-924).  Returns dark_best (not NaN-guarded). */
static void pso_period(pso_ctx *c, double freq, double *dev_best, double *per_best, double *dark_best,
                       double *la_best, double *be_best) {
    *per_best = *dark_best = *la_best = *be_best = 0;
    *dev_best = 1e40;
    for (int m = 1; m <= PSO_N_POLES; m++) {
        pso_fit_result r;
        pso_fit(c, freq, m, &r);
        if (r.dev_new < *dev_best) {
            *dev_best = r.dev_new;
            *per_best = r.prd;
            *dark_best = r.dark / r.totarea * 100;
            *la_best = r.la;
            *be_best = r.be;
        }
    }
    if (*la_best < 0) *la_best += 360;
}

/* Convexity-weight search, period_search_BOINC.cpp:487-715 */
int pso_conw_search(pso_ctx *c, double *conw_r_out, int *rounds_out) {
    int new_conw = 0, rounds = 0, nmax;
    double conw_r = c->conw_r, escl = c->escl;
    pso_ctx_freq(c, NULL, NULL, NULL, &nmax);
    int max_test_periods = 10;
    if (nmax < max_test_periods) max_test_periods = nmax;
    while ((new_conw != 1) && ((conw_r * escl * escl) < 10.0)) {
        pso_ctx_set_conw(c, conw_r);
        double sum_dark_facet = 0.0;
        for (int n = 1; n <= max_test_periods; n++) {
            double freq = c->freq_start - (n - 1) * c->freq_step;
            double dev_best, per_best, dark_best, la_best, be_best;
            pso_period(c, freq, &dev_best, &per_best, &dark_best, &la_best, &be_best);
            if (isnan(dark_best)) dark_best = 1.0;
            sum_dark_facet = sum_dark_facet + dark_best;
        }
        double ave_dark_facet = sum_dark_facet / max_test_periods;
        if (ave_dark_facet < 1.0) new_conw = 1;
        if (ave_dark_facet >= 1.0) conw_r = conw_r * 2;
        c->ndata = c->ndata_obs;
        rounds++;
    }
    *conw_r_out = conw_r;
    if (rounds_out) *rounds_out = rounds;
    return new_conw;
}

/* Main scan, period_search_BOINC.cpp:807-945.  lines[k] receives line n_from + k. */
int pso_scan(pso_ctx *c, int n_from, int n_to, double conw_r, pso_line *lines) {
    pso_ctx_set_conw(c, conw_r);
    int nmax;
    pso_ctx_freq(c, NULL, NULL, NULL, &nmax);
    if (n_to <= 0 || n_to > nmax) n_to = nmax;
    int cnt = 0;
    for (int n = n_from; n <= n_to; n++) {
        double freq = c->freq_start - (n - 1) * c->freq_step;
        double dev_best, per_best, dark_best, la_best, be_best;
        pso_period(c, freq, &dev_best, &per_best, &dark_best, &la_best, &be_best);
        pso_line *l = &lines[cnt++];
        l->per = 24 * per_best;
        l->dev = dev_best;
        l->chi2 = dev_best * dev_best * (c->ndata - 3);
        l->dark = (n == 1) ? conw_r * c->escl * c->escl : dark_best;
        l->la = la_best;
        l->be = be_best;
    }
    return cnt;
}

int pso_run(const pso_problem *p, int n_from, int n_to, pso_line *lines, double *conw_r_out) {
    pso_ctx *c = pso_ctx_create(p);
    if (!c) return -1;
    double conw_r;
    pso_conw_search(c, &conw_r, NULL);
    int cnt = pso_scan(c, n_from, n_to, conw_r, lines);
    if (conw_r_out) *conw_r_out = conw_r;
    pso_ctx_free(c);
    return cnt;
}

int pso_format_line(const pso_line *l, char *buf, int buflen) { /* :927-931 */
    return snprintf(buf, (size_t)buflen, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", l->per, l->dev, l->chi2, l->dark,
                    l->la, l->be);
}
