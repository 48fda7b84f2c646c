// ps_host.cpp -- host side of the MI355X period scan: period_search_in parsing, the Gaussian-image /
// spherical-harmonic setup tables, the convexity-weight search and period scan on top of the device
// fit batches, output formatting and checkpoint/resume.  No device code here; the fits themselves run
// in ps_device.hip.  Reference citations are relative to the reference repository root.
//
// Built with g++ and -fno-builtin-{sin,cos,sincos} (see Makefile): the setup tables are evaluated in
// the same operation order as the reference (period_search/*.c) with sincos() called exactly where
// the reference build (g++ -O3) fuses sin/cos pairs, so they match the reference bit for bit.
#define _GNU_SOURCE 1
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/periodsearch_amd.h"
#include "ps_internal.h"

namespace {
thread_local std::string g_err;
constexpr double kPi = M_PI;
constexpr int kPointsMax = 1000;  // period_search/constants.h:1
constexpr int kMaxNObs = 20000;   // constants.h:2
constexpr int kMaxLc = 200;       // constants.h:3
constexpr int kMaxNFac = 1000;    // constants.h:5
constexpr int kMaxLm = 10;        // constants.h:8
}  // namespace

int ps_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *ps_last_error(void) { return g_err.c_str(); }
extern "C" const char *ps_version(void) { return "periodsearch_amd 0.1 (gfx950)"; }

// ------------------------------------------------------------------------------------------------
// period_search_in parsing (period_search_BOINC.cpp:277-485).  Numbers are read like fscanf (any
// whitespace, including newlines, before a number); header lines then drop the rest of the line
// like the reference's fgets.
// ------------------------------------------------------------------------------------------------
namespace {
struct Reader {
    const char *s;
    size_t n, pos = 0;
    void ws() {
        while (pos < n && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n' || s[pos] == '\r' || s[pos] == '\v' ||
                           s[pos] == '\f'))
            ++pos;
    }
    bool num(double &v) {
        ws();
        if (pos >= n) return false;
        char *end = nullptr;
        double x = std::strtod(s + pos, &end);
        if (end == s + pos) return false;
        pos = size_t(end - s);
        v = x;
        return true;
    }
    bool num(int &v) {
        ws();
        if (pos >= n) return false;
        char *end = nullptr;
        long x = std::strtol(s + pos, &end, 10);
        if (end == s + pos) return false;
        pos = size_t(end - s);
        v = int(x);
        return true;
    }
    void eol() {
        while (pos < n && s[pos] != '\n') ++pos;
        if (pos < n) ++pos;
    }
};

template <class T>
T *dup_array(const std::vector<T> &v) {
    T *p = static_cast<T *>(std::malloc(sizeof(T) * (v.size() ? v.size() : 1)));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}
}  // namespace

extern "C" int ps_input_parse(const char *text, ps_input **out) {
    if (!text || !out) return ps_fail(PS_ERR_ARG, "ps_input_parse: null argument");
    *out = nullptr;
    Reader r{text, std::strlen(text)};
    ps_input in{};
    bool ok = r.num(in.per_start) && r.num(in.per_step_coef) && r.num(in.per_end) && r.num(in.ia_prd);
    r.eol();
    ok = ok && r.num(in.jd_00);
    r.eol();
    ok = ok && r.num(in.phi0);
    r.eol();
    ok = ok && r.num(in.conw);
    r.eol();
    ok = ok && r.num(in.lmax) && r.num(in.mmax);
    r.eol();
    ok = ok && r.num(in.nrows);
    r.eol();
    for (int i = 0; i < 3 && ok; ++i) {
        ok = r.num(in.par[i]) && r.num(in.ia_par[i]);
        r.eol();
    }
    ok = ok && r.num(in.cl) && r.num(in.ia_cl);
    r.eol();
    ok = ok && r.num(in.stop_condition);
    r.eol();
    ok = ok && r.num(in.n_iter_min);
    r.eol();
    ok = ok && r.num(in.alamda_incr);
    r.eol();
    ok = ok && r.num(in.alamda_start);
    r.eol();
    if (!ok) return ps_fail(PS_ERR_PARSE, "period_search_in: malformed header");
    if (in.lmax < 0 || in.lmax > kMaxLm || in.mmax < 0 || in.mmax > in.lmax || in.nrows < 1)
        return ps_fail(PS_ERR_LIMIT, "period_search_in: Lmax/Mmax/nrows out of range");
    if (8 * in.nrows * in.nrows > kMaxNFac)  // period_search_BOINC.cpp:542-543
        return ps_fail(PS_ERR_LIMIT, "number of facets is greater than MAX_N_FAC");
    if (!r.num(in.lcurves)) return ps_fail(PS_ERR_PARSE, "period_search_in: missing lightcurve count");
    if (in.lcurves < 1 || in.lcurves > kMaxLc)  // :326-327
        return ps_fail(PS_ERR_LIMIT, "number of lcs is greater than MAX_LC = %d", kMaxLc);

    std::vector<int> lpoints(in.lcurves), inrel(in.lcurves);
    std::vector<double> tim, br, sig, ee, ee0;
    double al0 = kPi, al0_abs = kPi;
    int ial0 = -1, ial0_abs = -1;
    double jd_min = 1e20, jd_max = -1e40;
    int onlyrel = 1;
    for (int i = 0; i < in.lcurves; ++i) {
        int i_temp = 0;
        if (!r.num(lpoints[i]) || !r.num(i_temp)) return ps_fail(PS_ERR_PARSE, "lightcurve %d: bad header", i + 1);
        r.eol();
        inrel[i] = 1 - i_temp;
        if (inrel[i] == 0) onlyrel = 0;
        if (lpoints[i] < 1 || lpoints[i] > kPointsMax)  // :352-353
            return ps_fail(PS_ERR_LIMIT, "number of lc points is greater than POINTS_MAX = %d", kPointsMax);
        double ave = 0;
        for (int j = 0; j < lpoints[i]; ++j) {
            if (int(tim.size()) + 1 > kMaxNObs)  // :360-361
                return ps_fail(PS_ERR_LIMIT, "number of data is greater than MAX_N_OBS = %d", kMaxNObs);
            double t, b, s[3], e[3];
            if (!r.num(t) || !r.num(b) || !r.num(s[0]) || !r.num(s[1]) || !r.num(s[2]) || !r.num(e[0]) ||
                !r.num(e[1]) || !r.num(e[2]))
                return ps_fail(PS_ERR_PARSE, "lightcurve %d point %d: expected 8 numbers", i + 1, j + 1);
            if (t < jd_min) jd_min = t;
            if (t > jd_max) jd_max = t;
            double slen = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
            double elen = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            ave += b;
            tim.push_back(t);
            br.push_back(b);
            for (int k = 0; k < 3; ++k) {
                ee.push_back(e[k] / elen);
                ee0.push_back(s[k] / slen);
            }
            if (j == 0) {  // solar phase angle of the lightcurve's first point (:384-399)
                double al = std::acos((e[0] * s[0] + e[1] * s[1] + e[2] * s[2]) / (elen * slen));
                if (al < al0) {
                    al0 = al;
                    ial0 = int(tim.size()) - 1;
                }
                if (al < al0_abs && inrel[i] == 0) {
                    al0_abs = al;
                    ial0_abs = int(tim.size()) - 1;
                }
            }
        }
        ave /= lpoints[i];
        for (int j = 0; j < lpoints[i]; ++j) sig.push_back(ave);
    }
    // (lightcurve, weight) pairs until the end of the file (:416-427); unlisted curves weigh 1.
    std::vector<double> wlc(in.lcurves, -1.0);
    for (;;) {
        int lc;
        double w;
        if (!r.num(lc) || !r.num(w)) break;
        if (lc >= 1 && lc <= in.lcurves) wlc[lc - 1] = w;
    }
    const int ndata = int(tim.size());
    double jd_0 = in.jd_00;
    if (jd_0 <= 0) jd_0 = int(jd_min);  // :431-436
    for (auto &t : tim) t = t - jd_0;
    std::vector<double> weight;
    for (int i = 0; i < in.lcurves; ++i)
        for (int j = 0; j < lpoints[i]; ++j) weight.push_back(wlc[i] == -1 ? 1.0 : wlc[i]);
    if (onlyrel == 0) {
        al0 = al0_abs;
        ial0 = ial0_abs;
    }
    // initial ellipsoid size (:465-477)
    double a = 1.05, b = 1.00, c = 0.95;
    double rfit = std::sqrt(2 * sig[ial0] / (0.5 * kPi * (1 + std::cos(al0))));
    double escl = rfit / std::sqrt((a * b + b * c + a * c) / 3);
    if (onlyrel == 0) escl *= 0.8;

    auto *res = static_cast<ps_input *>(std::calloc(1, sizeof(ps_input)));
    if (!res) return ps_fail(PS_ERR_NOMEM, "out of memory");
    *res = in;
    res->phi0 = in.phi0 * (kPi / 180);
    res->lpoints = dup_array(lpoints);
    res->inrel = dup_array(inrel);
    res->ndata = ndata;
    res->tim = dup_array(tim);
    res->brightness = dup_array(br);
    res->sig = dup_array(sig);
    res->weight = dup_array(weight);
    res->ee = dup_array(ee);
    res->ee0 = dup_array(ee0);
    res->jd_min = jd_min;
    res->jd_max = jd_max;
    res->jd_0 = jd_0;
    res->onlyrel = onlyrel;
    res->escl = escl;
    res->a_axis = a * escl;
    res->b_axis = b * escl;
    res->c_axis = c * escl;
    res->owned = 1;
    *out = res;
    return PS_OK;
}

extern "C" int ps_input_read(const char *path, ps_input **out) {
    if (!path || !out) return ps_fail(PS_ERR_ARG, "ps_input_read: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return ps_fail(PS_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    return ps_input_parse(text.c_str(), out);
}

extern "C" void ps_input_free(ps_input *in) {
    if (!in) return;
    if (in->owned) {
        std::free(in->lpoints);
        std::free(in->inrel);
        std::free(in->tim);
        std::free(in->brightness);
        std::free(in->sig);
        std::free(in->weight);
        std::free(in->ee);
        std::free(in->ee0);
        std::free(in);
    }
}

// ------------------------------------------------------------------------------------------------
// Setup tables: vertex directions (period_search_BOINC.cpp:507-539), trifac.c:10-82,
// areanorm.c:10-60, sphfunc.c:11-87, ellfit.c:11-88 with ludcmp.c/lubksb.c.
// ------------------------------------------------------------------------------------------------
namespace {
struct Mesh {
    std::vector<double> theta, phi;  // vertex directions, 0-based
    std::vector<int> tri;            // [nfac*3] 0-based vertex ids
};

Mesh build_mesh(int nrows) {
    Mesh m;
    const double dth = kPi / (2 * nrows);
    m.theta.push_back(0);
    m.phi.push_back(0);
    for (int i = 1; i <= nrows; ++i) {
        double dph = kPi / (2 * i);
        for (int j = 0; j <= 4 * i - 1; ++j) {
            m.theta.push_back(i * dth);
            m.phi.push_back(j * dph);
        }
    }
    for (int i = nrows - 1; i >= 1; --i) {
        double dph = kPi / (2 * i);
        for (int j = 0; j <= 4 * i - 1; ++j) {
            m.theta.push_back(kPi - i * dth);
            m.phi.push_back(j * dph);
        }
    }
    m.theta.push_back(kPi);
    m.phi.push_back(0);
    // Vertex numbering by latitude ring; ring r has 4*min(r, 2*nrows-r) vertices and wraps around.
    const int rings = 2 * nrows;
    std::vector<std::vector<int>> ring(rings + 1);
    int id = 0;
    ring[0] = {id++};
    for (int i = 1; i <= nrows; ++i) {
        ring[i].resize(4 * i + 1);
        for (int j = 0; j < 4 * i; ++j) ring[i][j] = id++;
        ring[i][4 * i] = ring[i][0];
    }
    for (int i = nrows - 1; i >= 1; --i) {
        int r = rings - i;
        ring[r].resize(4 * i + 1);
        for (int j = 0; j < 4 * i; ++j) ring[r][j] = id++;
        ring[r][4 * i] = ring[r][0];
    }
    ring[rings] = {id};
    auto add = [&](int a, int b, int c) {
        m.tri.push_back(a);
        m.tri.push_back(b);
        m.tri.push_back(c);
    };
    // Northern hemisphere: ring j1-1 (outer ring of the previous band) to ring j1, quadrant by quadrant.
    for (int j1 = 1; j1 <= nrows; ++j1)
        for (int q = 0; q < 4; ++q) {
            int j0 = q * j1;
            add(ring[j1 - 1][j0 - q], ring[j1][j0], ring[j1][j0 + 1]);
            for (int j2 = j0 + 1; j2 <= j0 + j1 - 1; ++j2) {
                add(ring[j1][j2], ring[j1 - 1][j2 - q], ring[j1 - 1][j2 - 1 - q]);
                add(ring[j1 - 1][j2 - q], ring[j1][j2], ring[j1][j2 + 1]);
            }
        }
    // Southern hemisphere.
    for (int j1 = nrows + 1; j1 <= rings; ++j1)
        for (int q = 0; q < 4; ++q) {
            int w = rings - j1;
            int j0 = q * w;
            add(ring[j1][j0], ring[j1 - 1][j0 + 1 + q], ring[j1 - 1][j0 + q]);
            for (int j2 = j0 + 1; j2 <= j0 + w; ++j2) {
                add(ring[j1][j2], ring[j1 - 1][j2 + q], ring[j1][j2 - 1]);
                add(ring[j1][j2], ring[j1 - 1][j2 + 1 + q], ring[j1 - 1][j2 + q]);
            }
        }
    return m;
}

// Associated Legendre polynomial coefficients in powers of cos(theta), times sin(theta)^m:
// P_l^m = sin^m(theta) * sum_n c[n][l][m] cos^n(theta)  (unnormalised, sphfunc.c:35-52 recursion).
struct LegendreCoef {
    double c[kMaxLm + 1][kMaxLm + 1][kMaxLm + 1] = {};
};

void legendre_coef(int lmax, int mmax, LegendreCoef &L) {
    auto &c = L.c;
    c[0][0][0] = 1;
    if (lmax >= 1) c[1][1][0] = 1;
    for (int l = 1; l <= lmax; ++l) c[0][l][l] = c[0][l - 1][l - 1] * (2 * l - 1);
    for (int m = 0; m <= mmax; ++m)
        for (int l = m + 1; l <= lmax; ++l) {
            int ibot = 1;
            if ((l - m) % 2 == 0) {
                c[0][l][m] = -(l + m - 1) * c[0][l - 2][m] / (1 * (l - m));
                ibot = 2;
            }
            if (l != 1)
                for (int n = ibot; n <= l - m; n += 2)
                    c[n][l][m] = ((2 * l - 1) * c[n - 1][l - 1][m] - (l + m - 1) * c[n][l - 2][m]) / (1 * (l - m));
        }
}

// Crout LU with implicit row scaling and partial pivoting, then back substitution (ludcmp.c, lubksb.c).
bool lu_solve(std::vector<double> &a, int n, std::vector<double> &b) {
    auto A = [&](int i, int j) -> double & { return a[size_t(i) * n + j]; };
    std::vector<double> v(n);
    std::vector<int> indx(n);
    for (int i = 0; i < n; ++i) {
        double big = 0.0, t;
        for (int j = 0; j < n; ++j)
            if ((t = std::fabs(A(i, j))) > big) big = t;
        if (big == 0.0) return false;
        v[i] = 1.0 / big;
    }
    int imax = -1;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < j; ++i) {
            double sum = A(i, j);
            for (int k = 0; k < i; ++k) sum -= A(i, k) * A(k, j);
            A(i, j) = sum;
        }
        double big = 0.0, dum;
        for (int i = j; i < n; ++i) {
            double sum = A(i, j);
            for (int k = 0; k < j; ++k) sum -= A(i, k) * A(k, j);
            A(i, j) = sum;
            if ((dum = v[i] * std::fabs(sum)) >= big) {
                big = dum;
                imax = i;
            }
        }
        if (j != imax) {
            for (int k = 0; k < n; ++k) std::swap(A(imax, k), A(j, k));
            v[imax] = v[j];
        }
        indx[j] = imax;
        if (A(j, j) == 0.0) A(j, j) = 1.0e-20;
        if (j != n - 1) {
            dum = 1.0 / A(j, j);
            for (int i = j + 1; i < n; ++i) A(i, j) *= dum;
        }
    }
    int ii = -1;
    for (int i = 0; i < n; ++i) {
        int ip = indx[i];
        double sum = b[ip];
        b[ip] = b[i];
        if (ii >= 0)
            for (int j = ii; j <= i - 1; ++j) sum -= A(i, j) * b[j];
        else if (sum)
            ii = i;
        b[i] = sum;
    }
    for (int i = n - 1; i >= 0; --i) {
        double sum = b[i];
        for (int j = i + 1; j < n; ++j) sum -= A(i, j) * b[j];
        b[i] = sum / A(i, i);
    }
    return true;
}
}  // namespace

extern "C" int ps_geometry_build(const ps_input *in, ps_geometry **out) {
    if (!in || !out) return ps_fail(PS_ERR_ARG, "ps_geometry_build: null argument");
    const int nrows = in->nrows, lmax = in->lmax, mmax = in->mmax;
    if (nrows < 1 || 8 * nrows * nrows > kMaxNFac || lmax < 0 || lmax > kMaxLm || mmax < 0 || mmax > lmax)
        return ps_fail(PS_ERR_LIMIT, "ps_geometry_build: nrows/lmax/mmax out of range");
    Mesh mesh = build_mesh(nrows);
    const int nfac = 8 * nrows * nrows, ndir = int(mesh.theta.size());
    int ncoef = 0, npairs = 0;
    for (int m = 0; m <= mmax; ++m)
        for (int l = m; l <= lmax; ++l) {
            ++npairs;
            ncoef += (m != 0) ? 2 : 1;
        }
    auto *g = static_cast<ps_geometry *>(std::calloc(1, sizeof(ps_geometry)));
    g->nrows = nrows;
    g->nfac = nfac;
    g->ndir = ndir;
    g->lmax = lmax;
    g->mmax = mmax;
    g->ncoef = ncoef;
    g->npairs = npairs;
    g->nor = static_cast<double *>(std::calloc(size_t(nfac) * 3, sizeof(double)));
    g->darea = static_cast<double *>(std::calloc(nfac, sizeof(double)));
    g->dsph = static_cast<double *>(std::calloc(size_t(nfac) * ncoef, sizeof(double)));
    g->pleg = static_cast<double *>(std::calloc(size_t(nfac) * npairs, sizeof(double)));
    g->fc = static_cast<double *>(std::calloc(size_t(nfac) * (mmax + 1), sizeof(double)));
    g->fs = static_cast<double *>(std::calloc(size_t(nfac) * (mmax + 1), sizeof(double)));
    g->pair_m = static_cast<int *>(std::calloc(npairs, sizeof(int)));
    g->pair_l = static_cast<int *>(std::calloc(npairs, sizeof(int)));
    g->cg_first = static_cast<double *>(std::calloc(ncoef, sizeof(double)));

    // vertex positions and facet normals/areas (areanorm.c:22-57)
    std::vector<double> x(ndir), y(ndir), z(ndir), nth(nfac), nph(nfac);
    for (int i = 0; i < ndir; ++i) {
        double st = std::sin(mesh.theta[i]);
        x[i] = st * std::cos(mesh.phi[i]);
        y[i] = st * std::sin(mesh.phi[i]);
        z[i] = std::cos(mesh.theta[i]);
    }
    for (int i = 0; i < nfac; ++i) {
        const int v1 = mesh.tri[3 * i], v2 = mesh.tri[3 * i + 1], v3 = mesh.tri[3 * i + 2];
        double ax = x[v2] - x[v1], ay = y[v2] - y[v1], az = z[v2] - z[v1];
        double bx = x[v3] - x[v1], by = y[v3] - y[v1], bz = z[v3] - z[v1];
        double cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
        double clen = std::sqrt(cx * cx + cy * cy + cz * cz);
        g->nor[3 * i] = cx / clen;
        g->nor[3 * i + 1] = cy / clen;
        g->nor[3 * i + 2] = cz / clen;
        nth[i] = std::acos(g->nor[3 * i + 2]);
        nph[i] = std::atan2(g->nor[3 * i + 1], g->nor[3 * i]);
        g->darea[i] = 0.5 * clen;
    }
    // spherical-harmonic tables at the facet normals (sphfunc.c)
    LegendreCoef L;
    legendre_coef(lmax, mmax, L);
    {
        int q = 0;
        for (int m = 0; m <= mmax; ++m)
            for (int l = m; l <= lmax; ++l, ++q) {
                g->pair_m[q] = m;
                g->pair_l[q] = l;
            }
    }
    std::vector<double> sn(lmax + 1), cs(lmax + 1);
    for (int i = 0; i < nfac; ++i) {
        sn[0] = cs[0] = 1;
        if (lmax >= 1) {
            sn[1] = std::sin(nth[i]);
            cs[1] = std::cos(nth[i]);
        }
        for (int j = 2; j <= lmax; ++j) {
            sn[j] = sn[1] * sn[j - 1];
            cs[j] = cs[1] * cs[j - 1];
        }
        double *fc = &g->fc[size_t(i) * (mmax + 1)], *fs = &g->fs[size_t(i) * (mmax + 1)];
        fs[0] = 0;
        fc[0] = 1;
        for (int j = 1; j <= mmax; ++j) {
            fs[j] = std::sin(j * nph[i]);
            fc[j] = std::cos(j * nph[i]);
        }
        int k = 0;
        for (int q = 0; q < npairs; ++q) {
            const int m = g->pair_m[q], l = g->pair_l[q];
            double p = 0;
            for (int n = (l - m) % 2; n <= l - m; n += 2) p = p + L.c[n][l][m] * cs[n] * sn[m];
            g->pleg[size_t(i) * npairs + q] = p;
            g->dsph[size_t(i) * ncoef + k++] = fc[m] * p;
            if (m != 0) g->dsph[size_t(i) * ncoef + k++] = fs[m] * p;
        }
    }
    // least-squares fit of the log-curvature of the initial ellipsoid (ellfit.c)
    const double a = in->a_axis, b = in->b_axis, c = in->c_axis;
    std::vector<double> er(nfac), fmat(size_t(nfac) * ncoef), fitmat(size_t(ncoef) * ncoef), fitvec(ncoef);
    for (int i = 0; i < nfac; ++i) {
        double st, ct, sf, cf;
        sincos(nth[i], &st, &ct);  // the reference build evaluates these pairs with sincos()
        sincos(nph[i], &sf, &cf);
        double u = a * st * cf, v = b * st * sf, w = c * ct;
        er[i] = 2 * (std::log(a * b * c) - std::log(u * u + v * v + w * w));
        int k = 0;
        for (int q = 0; q < npairs; ++q) {
            const int m = g->pair_m[q];
            const double p = g->pleg[size_t(i) * npairs + q];
            if (m != 0) {
                fmat[size_t(i) * ncoef + k++] = p * std::cos(m * nph[i]);
                fmat[size_t(i) * ncoef + k++] = p * std::sin(m * nph[i]);
            } else {
                fmat[size_t(i) * ncoef + k++] = p;
            }
        }
    }
    for (int i = 0; i < ncoef; ++i) {
        for (int j = 0; j < ncoef; ++j) {
            double s = 0;
            for (int k = 0; k < nfac; ++k) s = s + fmat[size_t(k) * ncoef + i] * fmat[size_t(k) * ncoef + j];
            fitmat[size_t(i) * ncoef + j] = s;
        }
        double s = 0;
        for (int j = 0; j < nfac; ++j) s = s + fmat[size_t(j) * ncoef + i] * er[j];
        fitvec[i] = s;
    }
    if (!lu_solve(fitmat, ncoef, fitvec)) {
        ps_geometry_free(g);
        return ps_fail(PS_ERR_ARG, "singular matrix in the initial ellipsoid fit");
    }
    for (int i = 0; i < ncoef; ++i) g->cg_first[i] = fitvec[i];
    *out = g;
    return PS_OK;
}

extern "C" void ps_geometry_free(ps_geometry *g) {
    if (!g) return;
    std::free(g->nor);
    std::free(g->darea);
    std::free(g->dsph);
    std::free(g->pleg);
    std::free(g->fc);
    std::free(g->fs);
    std::free(g->pair_m);
    std::free(g->pair_l);
    std::free(g->cg_first);
    std::free(g);
}

// ------------------------------------------------------------------------------------------------
// Scan driver on top of ps_fit_batch: best pole per period (period_search_BOINC.cpp:815-924),
// convexity-weight search (:487-715), main scan (:807-945).
// ------------------------------------------------------------------------------------------------
namespace {
struct PeriodBest {
    double dev, per, dark, la, be;
};

// Fits periods n_from..n_to (1-based) x 10 poles in device batches and reduces each period to its
// best pole exactly as the reference does (first pole wins ties; lambda wrapped into [0, 360)).
int best_poles(ps_handle *h, int n_from, int n_to, double conw_r, std::vector<PeriodBest> &best) {
    ps_info info;
    ps_get_info(h, &info);
    const int nper = n_to - n_from + 1;
    best.assign(nper, PeriodBest{});
    const int chunk = std::max(1, (1 << 16) / PS_N_POLES);
    std::vector<double> freq;
    std::vector<int> pole;
    std::vector<ps_fit_result> res;
    for (int c0 = 0; c0 < nper; c0 += chunk) {
        const int cn = std::min(chunk, nper - c0);
        freq.resize(size_t(cn) * PS_N_POLES);
        pole.resize(freq.size());
        res.resize(freq.size());
        for (int i = 0; i < cn; ++i) {
            const int n = n_from + c0 + i;
            const double f = info.freq_start - (n - 1) * info.freq_step;  // :812
            for (int m = 0; m < PS_N_POLES; ++m) {
                freq[size_t(i) * PS_N_POLES + m] = f;
                pole[size_t(i) * PS_N_POLES + m] = m + 1;
            }
        }
        int rc = ps_fit_batch(h, int(freq.size()), freq.data(), pole.data(), conw_r, res.data());
        if (rc != PS_OK) return rc;
        for (int i = 0; i < cn; ++i) {
            PeriodBest b{1e40, 0, 0, 0, 0};
            for (int m = 0; m < PS_N_POLES; ++m) {
                const ps_fit_result &r = res[size_t(i) * PS_N_POLES + m];
                if (r.dev < b.dev) {
                    b.dev = r.dev;
                    b.per = r.prd;
                    b.dark = r.dark / r.totarea * 100;
                    b.la = r.la;
                    b.be = r.be;
                }
            }
            if (b.la < 0) b.la += 360;
            best[c0 + i] = b;
        }
    }
    return PS_OK;
}
}  // namespace

extern "C" int ps_conw_round(ps_handle *h, double conw_r, double *ave_dark) {
    if (!h || !ave_dark) return ps_fail(PS_ERR_ARG, "ps_conw_round: null argument");
    ps_info info;
    ps_get_info(h, &info);
    const int ntest = std::min(10, info.nmax);  // :577-582
    std::vector<PeriodBest> best;
    int rc = best_poles(h, 1, ntest, conw_r, best);
    if (rc != PS_OK) return rc;
    double sum = 0.0;
    for (auto &b : best) sum = sum + (std::isnan(b.dark) ? 1.0 : b.dark);  // :701-704
    *ave_dark = sum / ntest;
    return PS_OK;
}

extern "C" int ps_conw_search(ps_handle *h, double *conw_r_io, int *new_conw_out) {
    if (!h || !conw_r_io) return ps_fail(PS_ERR_ARG, "ps_conw_search: null argument");
    ps_info info;
    ps_get_info(h, &info);
    double conw_r = info.conw_r0;
    int new_conw = 0;
    while (new_conw != 1 && conw_r * info.escl * info.escl < 10.0) {  // :489
        double ave;
        int rc = ps_conw_round(h, conw_r, &ave);
        if (rc != PS_OK) return rc;
        if (ave < 1.0) new_conw = 1;
        if (ave >= 1.0) conw_r = conw_r * 2;
    }
    *conw_r_io = conw_r;
    if (new_conw_out) *new_conw_out = new_conw;
    return PS_OK;
}

extern "C" int ps_scan(ps_handle *h, int n_from, int n_to, double conw_r, ps_line *lines) {
    if (!h || !lines) return ps_fail(PS_ERR_ARG, "ps_scan: null argument");
    ps_info info;
    ps_get_info(h, &info);
    if (n_to <= 0 || n_to > info.nmax) n_to = info.nmax;
    if (n_from < 1 || n_from > n_to) return ps_fail(PS_ERR_ARG, "ps_scan: bad period range %d..%d", n_from, n_to);
    std::vector<PeriodBest> best;
    int rc = best_poles(h, n_from, n_to, conw_r, best);
    if (rc != PS_OK) return rc;
    for (int i = 0; i < int(best.size()); ++i) {
        const int n = n_from + i;
        ps_line &l = lines[i];
        l.per = 24 * best[i].per;
        l.dev = best[i].dev;
        l.chi2 = best[i].dev * best[i].dev * info.ndata;  // dev^2 * (ndata - 3), ndata incl. conv rows
        l.dark = (n == 1) ? conw_r * info.escl * info.escl : best[i].dark;
        l.la = best[i].la;
        l.be = best[i].be;
    }
    return PS_OK;
}

extern "C" int ps_format_line(const ps_line *l, char *buf, int buflen) {
    if (!l || !buf || buflen <= 0) return ps_fail(PS_ERR_ARG, "ps_format_line: bad argument");
    return std::snprintf(buf, size_t(buflen), "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", l->per, l->dev, l->chi2,
                         l->dark, l->la, l->be);
}

// ------------------------------------------------------------------------------------------------
// Process contract with checkpoint/resume (period_search_BOINC.cpp:85-101, :234-261, :926-937):
// the state file holds the number of output lines written; on restart the output is appended from
// line nlines+1.  The convexity-weight search reruns on restart, as in the plain reference.
// ------------------------------------------------------------------------------------------------
extern "C" int ps_run_files(const char *in_path, const char *out_path, const char *state_path, int device,
                            int precision) {
    if (!in_path || !out_path) return ps_fail(PS_ERR_ARG, "ps_run_files: null path");
    int n_start_from = 1;
    if (state_path) {
        if (FILE *st = std::fopen(state_path, "r")) {
            int nlines;
            if (std::fscanf(st, "%d", &nlines) == 1 && nlines >= 0) n_start_from = nlines + 1;
            std::fclose(st);
        }
    }
    ps_input *in = nullptr;
    ps_geometry *geo = nullptr;
    ps_handle *h = nullptr;
    int rc = ps_input_read(in_path, &in);
    if (rc == PS_OK) rc = ps_geometry_build(in, &geo);
    if (rc == PS_OK) rc = ps_prepare(in, geo, device, &h);
    if (rc == PS_OK) rc = ps_set_precision(h, precision);
    double conw_r = 0;
    if (rc == PS_OK) rc = ps_conw_search(h, &conw_r, nullptr);
    FILE *out = nullptr;
    if (rc == PS_OK) {
        out = std::fopen(out_path, n_start_from > 1 ? "a" : "w");
        if (!out) rc = ps_fail(PS_ERR_IO, "cannot open %s", out_path);
    }
    if (rc == PS_OK) {
        ps_info info;
        ps_get_info(h, &info);
        const int block = 512;  // periods per device batch between checkpoints
        std::vector<ps_line> lines(block);
        for (int n0 = n_start_from; n0 <= info.nmax && rc == PS_OK; n0 += block) {
            const int n1 = std::min(info.nmax, n0 + block - 1);
            rc = ps_scan(h, n0, n1, conw_r, lines.data());
            if (rc != PS_OK) break;
            char buf[256];
            for (int i = 0; i <= n1 - n0; ++i) {
                ps_format_line(&lines[i], buf, sizeof buf);
                std::fputs(buf, out);
            }
            std::fflush(out);
            if (state_path) {  // write-then-rename, like do_checkpoint
                std::string tmp = std::string(state_path) + ".tmp";
                if (FILE *st = std::fopen(tmp.c_str(), "w")) {
                    std::fprintf(st, "%d", n1);
                    std::fclose(st);
                    std::rename(tmp.c_str(), state_path);
                }
            }
        }
    }
    if (out) std::fclose(out);
    ps_release(h);
    ps_geometry_free(geo);
    ps_input_free(in);
    return rc;
}

