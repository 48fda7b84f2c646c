"""Seeded synthetic period_search_in inputs (the benchmark and parity workloads).

A convex asteroid (log-facet-area spherical-harmonic coefficients, c00 fixed, the rest N(0, sigma))
with a random pole and a fixed sidereal period is observed from Earth (circular orbit, 1 AU) while it
moves on a circular, slightly inclined orbit at 2.5 AU, which gives solar phase angles of 2-24 deg.
Brightness follows the reference's scattering model -- Lommel-Seeliger + cl * Lambert, exponential-
linear phase function (period_search/bright.c:60-75, phasec.c:11-24) -- on the triangulated Gaussian
image (the setup tables come from libpsamd's ps_geometry_build).  Dense lightcurves are written as
relative (normalised to mean 1), sparse points as one or more calibrated lightcurves (<= 1000 points
each, the plain reference's POINTS_MAX).  Gaussian noise, unit weights and an explicit "1 1" weights
line close the file, as the reference parser expects (period_search_BOINC.cpp:416-427).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

JD0 = 2455000.0
OBLIQ_INC_DEG = 8.0  # inclination of the asteroid orbit to the ecliptic


@dataclass
class SynthSpec:
    seed: int = 1000
    truth_lmax: int = 3
    truth_sigma: float = 0.3
    truth_nrows: int = 8
    period_h: float = 7.35
    n_dense: int = 10
    dense_points: int = 60
    dense_hours: float = 6.0
    dense_noise: float = 0.01
    n_sparse: int = 0
    sparse_noise: float = 0.02
    span_years: float = 2.0
    # fit header
    per_start: float = 6.9
    per_end: float = 7.8
    per_step_coef: float = 1.0
    fit_lmax: int = 3
    fit_nrows: int = 6
    stop_condition: float = 50
    n_iter_min: int = 10
    conw: float = 0.1
    phase: tuple = (0.5, 0.1, -0.5)
    cl: float = 0.1
    cl_free: bool = False  # fit the Lambert weight too (the input's cl flag)
    extra: dict = field(default_factory=dict)


def _geometry(nrows, lmax):
    """Facet normals, areas and spherical-harmonic basis from the product's setup code."""
    inp = L.PsInput()
    inp.nrows, inp.lmax, inp.mmax = nrows, lmax, lmax
    inp.a_axis = inp.b_axis = inp.c_axis = 1.0
    g = C.POINTER(L.PsGeometry)()
    L.check(L.lib().ps_geometry_build(C.byref(inp), C.byref(g)), "ps_geometry_build")
    gc = g.contents
    nf, nc = gc.nfac, gc.ncoef
    out = (np.ctypeslib.as_array(gc.nor, (nf * 3,)).reshape(nf, 3).copy(),
           np.ctypeslib.as_array(gc.darea, (nf,)).copy(),
           np.ctypeslib.as_array(gc.dsph, (nf * nc,)).reshape(nf, nc).copy())
    L.lib().ps_geometry_free(g)
    return out


def _positions(jd):
    """Asteroid-centric Sun and Earth vectors [AU] in ecliptic coordinates."""
    t = jd - JD0
    th_e = 2 * np.pi * t / 365.25 + 0.3
    th_a = 2 * np.pi * t / (365.25 * 2.5 ** 1.5) + 1.1
    inc = np.radians(OBLIQ_INC_DEG)
    earth = np.stack([np.cos(th_e), np.sin(th_e), np.zeros_like(t)], -1)
    ast = 2.5 * np.stack([np.cos(th_a), np.sin(th_a) * np.cos(inc), np.sin(th_a) * np.sin(inc)], -1)
    return -ast, earth - ast


def _observable(jd):
    sun, earth = _positions(jd)
    cos_a = np.sum(sun * earth, -1) / (np.linalg.norm(sun, axis=-1) * np.linalg.norm(earth, axis=-1))
    alpha = np.degrees(np.arccos(np.clip(cos_a, -1, 1)))
    # solar elongation seen from Earth: angle between (Sun - Earth) and (asteroid - Earth)
    e2s = sun - earth
    e2a = -earth
    cos_el = np.sum(e2s * e2a, -1) / (np.linalg.norm(e2s, axis=-1) * np.linalg.norm(e2a, axis=-1))
    elong = np.degrees(np.arccos(np.clip(cos_el, -1, 1)))
    return (alpha >= 2.0) & (elong >= 100.0)


def model_brightness(nor, darea, dsph, coef, lam_deg, beta_deg, period_h, jd, sun, earth, phase, cl, cls=1.0,
                     t0=JD0):
    """Disk-integrated brightness of the convex model (bright.c:14-126 without derivatives)."""
    area = darea * np.exp(dsph @ coef)
    e = earth / np.linalg.norm(earth, axis=-1, keepdims=True)
    e0 = sun / np.linalg.norm(sun, axis=-1, keepdims=True)
    b, lmb = np.radians(90.0 - beta_deg), np.radians(lam_deg)
    cb, sb, cl_, sl = np.cos(b), np.sin(b), np.cos(lmb), np.sin(lmb)
    bl = np.array([[cb * cl_, cb * sl, -sb], [-sl, cl_, 0.0], [sb * cl_, sb * sl, cb]])
    omega = 24 * 2 * np.pi / period_h
    f = np.mod(omega * (jd - t0), 2 * np.pi)
    cf, sf = np.cos(f), np.sin(f)
    fm = np.zeros((len(jd), 3, 3))
    fm[:, 0, 0], fm[:, 0, 1], fm[:, 1, 0], fm[:, 1, 1], fm[:, 2, 2] = cf, sf, -sf, cf, 1.0
    tm = fm @ bl
    eb = np.einsum("nij,nj->ni", tm, e)
    e0b = np.einsum("nij,nj->ni", tm, e0)
    mu = eb @ nor.T
    mu0 = e0b @ nor.T
    vis = (mu > 1e-8) & (mu0 > 1e-8)
    s = np.where(vis, mu * mu0 * (cl + cls / np.where(vis, mu + mu0, 1.0)), 0.0)
    br = s @ area
    alpha = np.arccos(np.clip(np.sum(e * e0, -1), -1, 1))
    a, d, k = phase
    return br * (1 + a * np.exp(-alpha / d) + k * alpha)


def generate(spec: SynthSpec):
    """Returns (period_search_in text, truth dict)."""
    rng = np.random.default_rng(spec.seed)
    nor, darea, dsph = _geometry(spec.truth_nrows, spec.truth_lmax)
    coef = np.zeros(dsph.shape[1])
    coef[1:] = rng.normal(0.0, spec.truth_sigma, dsph.shape[1] - 1)
    lam = rng.uniform(0, 360)
    beta = np.degrees(np.arcsin(rng.uniform(-1, 1)))
    span = spec.span_years * 365.25
    # candidate nights: observable dates over the span
    grid = JD0 + np.arange(0, span, 0.5)
    ok = grid[_observable(grid)]
    lcs = []
    if spec.n_dense:
        nights = np.sort(rng.choice(ok, size=spec.n_dense, replace=False))
        for jd_n in nights:
            jd = jd_n + np.linspace(0, spec.dense_hours / 24, spec.dense_points)
            sun, earth = _positions(jd)
            y = model_brightness(nor, darea, dsph, coef, lam, beta, spec.period_h, jd, sun, earth, spec.phase,
                                 spec.cl)
            y = y * (1 + spec.dense_noise * rng.standard_normal(len(y)))
            lcs.append((0, jd, y / y.mean(), sun, earth))
    if spec.n_sparse:
        # one point per observable night while there are enough nights, else several per night (sparse
        # surveys revisit a field 2-3 times a night)
        jd = np.sort(rng.choice(ok, size=spec.n_sparse, replace=spec.n_sparse > len(ok)) +
                     rng.uniform(0, 0.5, spec.n_sparse))
        sun, earth = _positions(jd)
        y = model_brightness(nor, darea, dsph, coef, lam, beta, spec.period_h, jd, sun, earth, spec.phase, spec.cl)
        y = y * (1 + spec.sparse_noise * rng.standard_normal(len(y)))
        for c0 in range(0, spec.n_sparse, 1000):
            sl = slice(c0, min(spec.n_sparse, c0 + 1000))
            lcs.append((1, jd[sl], y[sl], sun[sl], earth[sl]))
    out = []
    out.append(f"{spec.per_start} {spec.per_step_coef} {spec.per_end} 1\tperiod_start period_step period_end fixed/free")
    out.append("0\t\tepoch of zero time t0")
    out.append("0\t\tinitial fixed rotation angle fi0")
    out.append(f"{spec.conw}\t\tthe weight factor for conv. reg.")
    out.append(f"{spec.fit_lmax} {spec.fit_lmax}\t\tdegree and order of the Laplace series")
    out.append(f"{spec.fit_nrows}\t\tnr. of triangulation rows per octant")
    out.append(f"{spec.phase[0]} 1\t\tinitial guesses for phase funct. params.")
    out.append(f"{spec.phase[1]} 1")
    out.append(f"{spec.phase[2]} 1")
    out.append(f"{spec.cl} {int(spec.cl_free)}\t\tinitial Lambert coeff. (L-S=1)")
    out.append(f"{spec.stop_condition:g}\t\tstop condition")
    out.append(f"{spec.n_iter_min}\t\tminimum number of iterations")
    out.append("5\t\tAlamda multiplicative factor")
    out.append("0.1\t\tAlamda initial value")
    out.append(f"{len(lcs)}")
    for calib, jd, y, sun, earth in lcs:
        out.append(f"{len(jd)} {calib}")
        for i in range(len(jd)):
            s, e = sun[i], earth[i]
            out.append(f"{jd[i]:.6f} {y[i]:.6e} {s[0]:.6e} {s[1]:.6e} {s[2]:.6e} {e[0]:.6e} {e[1]:.6e} {e[2]:.6e}")
    out.append("1 1")
    truth = {"coef": coef, "lambda": lam, "beta": beta, "period_h": spec.period_h,
             "ndata": sum(len(l[1]) for l in lcs)}
    return "\n".join(out) + "\n", truth


# Workload shapes (dense + sparse photometry, fit settings); seeds 1000 + k.
def config0(**kw):
    """Plain reference method, FP64, CPU-runnable: 10 dense lightcurves x 60 points over 2 years,
    Lmax = 3 fit on 288 facets, scan 6.9-7.8 h."""
    base = dict(seed=1000)
    base.update(kw)
    return SynthSpec(**base)


def config1(**kw):
    """Production shape: 40 dense lightcurves x 50 points + 400 sparse calibrated points over 15 years,
    Lmax = Mmax = 6 fit (49 shape coefficients) on 288 facets, scan 2-20 h."""
    base = dict(seed=1001, n_dense=40, dense_points=50, n_sparse=400, span_years=15.0, per_start=2.0, per_end=20.0,
                fit_lmax=6, fit_nrows=6)
    base.update(kw)
    return SynthSpec(**base)


def config3(**kw):
    """Sparse survey: 5000 calibrated points over 10 years (3% noise), Lmax = 6 fit on 512 facets,
    scan 2-100 h."""
    base = dict(seed=1003, n_dense=0, n_sparse=5000, sparse_noise=0.03, span_years=10.0, per_start=2.0,
                per_end=100.0, fit_lmax=6, fit_nrows=8)
    base.update(kw)
    return SynthSpec(**base)


# The batched uneven workload: many asteroids of different sizes and model shapes.
CONFIG4_N = 512
CONFIG4_MAX_DENSE = 150  # dense lightcurves per asteroid (the reference allows fewer than 200 in all)


def config4(i, **kw):
    """Asteroid i (0 <= i < 512) of the batched workload: ndata log-uniform in 100-20000 (5/6 of it in
    dense relative lightcurves of 50 points, at most 150 of them, the rest sparse calibrated points),
    Lmax = Mmax in {3, 4, 6}, nrows in {6, 8}, a random true period in 2.5-18 h, and a scan window of
    1000-20000 trial periods (log-uniform) around it.  The convergence stop (rms improvement below
    0.001, 10-100 iterations) gives the fits uneven iteration counts.  Seeds 1004 (the shape of the
    batch) and 40000 + i (the asteroid)."""
    rng = np.random.default_rng([1004, i])
    ndata = int(round(np.exp(rng.uniform(np.log(100), np.log(19997)))))
    n_dense = min(CONFIG4_MAX_DENSE, int(round(ndata * 5 / 6 / 50)))
    n_sparse = ndata - 50 * n_dense
    lmax = int(rng.choice([3, 4, 6]))
    nrows = int(rng.choice([6, 8]))
    period = float(rng.uniform(2.5, 18.0))
    nper = int(round(np.exp(rng.uniform(np.log(1000), np.log(20000)))))
    span_years = 15.0
    # trial frequencies f0 - (n - 1) df with df = 0.5 / (span_days * 24) (period_search_BOINC.cpp:784-786)
    df = 0.5 / (span_years * 365.25 * 24)
    f_true = 1 / period
    per_start = 1 / (f_true + 0.5 * nper * df)
    per_end = 1 / (f_true - 0.5 * nper * df)
    base = dict(seed=40000 + i, n_dense=n_dense, dense_points=50, n_sparse=n_sparse, span_years=span_years,
                period_h=period, per_start=round(per_start, 6), per_end=round(per_end, 6), fit_lmax=lmax,
                fit_nrows=nrows, stop_condition=0.001, n_iter_min=10)
    base.update(kw)
    return SynthSpec(**base)


def config2_vectors(ps, nvec=4096, seed=1002):
    """Parameter vectors of the differentiation workload at a prepared problem's shape (0-based, ma
    entries): shape coefficients N(0, 0.3) in units of each basis function's largest facet value (the
    unnormalised Legendre functions reach ~1e4 at l = m = 6, so raw N(0, 0.3) coefficients overflow
    exp()), c00 from the initial ellipsoid, poles uniform on the sphere (beta as colatitude), periods
    U(2, 20) h, the phase-function and Lambert parameters at the input's initial values, cls = 1."""
    rng = np.random.default_rng(seed)
    geo = ps.geometry()
    inp = ps.input
    nc = len(geo["cg_first"])
    span = np.abs(geo["dsph"]).max(0)
    a = np.zeros((nvec, nc + 8))
    a[:, 0] = geo["cg_first"][0]
    a[:, 1:nc] = rng.normal(0, 0.3, (nvec, nc - 1)) / span[1:]
    a[:, nc] = np.arccos(rng.uniform(-1, 1, nvec))
    a[:, nc + 1] = rng.uniform(0, 2 * np.pi, nvec)
    a[:, nc + 2] = 24 * 2 * np.pi / rng.uniform(2, 20, nvec)  # omega [rad/d]
    a[:, nc + 3:nc + 6] = [inp.par[0], inp.par[1], inp.par[2]]
    a[:, nc + 6] = np.log(inp.cl)
    a[:, nc + 7] = 1.0
    return a
