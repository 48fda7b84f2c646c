"""Scan-level parity at production shape against committed reference-core fixtures
(tests/golden/scan_<workload>.json, written by tests/golden/make_scan_fixtures.py from the reference's
own plain-C core and, as a second valid FP64 rounding, the FMA-contracted oracle).

Per workload every (period, pole) fit of the fixture's window runs through ps_fit_batch and every
period_search_out line through ps_scan (period_search_BOINC.cpp:807-945):
  config0  the whole 462-period grid of the CPU-runnable case after the convexity-weight search
  config1  64 periods around the true period of the production shape (Lmax 6, 288 facets, ndata 2400)
  config3  32 periods around the true period of the sparse survey (Lmax 6, 512 facets, ndata 5000)

Bar (FP64): the best-period index is identical; the iteration count of every fit is identical; rms,
chi2 and period of every fit and line agree with the reference core to 1e-9 relative or -- for the
rounding-sensitive fits, where the reference's two FP64 roundings themselves disagree by more -- lie
within one spread-width of their midpoint (the GPU is a third FP64 rounding); dark, lambda and beta
likewise, to print precision.  A printed line equals the reference's, the FMA build's, or lies within
that spread.  NaN (the plain reference has no NaN guard on the dark-facet figure of the main scan,
period_search_BOINC.cpp:900-902) must be NaN on the GPU too.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from periodsearch_amd import PeriodSearch, format_line, synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ["n", "pole", "dev", "prd", "dark", "totarea", "la", "be", "chisq", "niter"]


def load(name):
    with open(os.path.join(GOLD, f"scan_{name}.json")) as f:
        return json.load(f)


def agrees(g, a, b, rel=1e-9, abs_slack=0.0):
    """g matches the reference value a to rel (or abs_slack), or lies within one spread-width of the
    midpoint of the reference's two roundings a and b; NaN only matches NaN."""
    g, a, b = (np.asarray(x, float) for x in (g, a, b))
    nan = np.isnan(a) | np.isnan(b)
    tol = np.maximum(rel * np.abs(a), abs_slack)
    close = np.abs(g - a) <= tol
    d = np.abs(a - b)
    spread = (d > tol) & (np.abs(g - (a + b) / 2) <= d + tol)
    return np.where(nan, np.isnan(g) | (np.isnan(a) != np.isnan(b)), close | spread)


def last_digit_close(t, tr):
    """Two printed lines whose columns differ by at most one unit of their last printed digit."""
    a, b = t.split(), tr.split()
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        unit = 10.0 ** -(len(y.split(".")[1]) if "." in y else 0)
        if x != y and not (x.lower() != "nan" and abs(float(x) - float(y)) <= 1.000001 * unit):
            return False
    return True


@pytest.mark.parametrize("name", ["config0", "config1", "config3"])
def test_scan_matches_reference_core_fixture(name):
    fx = load(name)
    text, _ = synth.generate(getattr(synth, name)())
    assert hashlib.sha256(text.encode()).hexdigest() == fx["sha256"], "the seeded generator changed"
    ref = np.array(fx["fits_ref"], float)
    fma = np.array(fx["fits_fma"], float)
    col = {k: i for i, k in enumerate(FIELDS)}
    ps = PeriodSearch.from_text(text).prepare(device=0)
    try:
        conw_r = fx["conw_r"]
        if fx["conw_search"]:
            cw, new_conw = ps.conw_search()
            assert (cw, new_conw) == (conw_r, fx["new_conw"])
        n = ref[:, col["n"]].astype(int)
        pole = ref[:, col["pole"]].astype(int)
        got = ps.fit_batch(ps.freq(n), pole, conw_r)
        assert np.array_equal(got["niter"], ref[:, col["niter"]].astype(int))
        bad = {}
        for key in ("dev", "chisq", "prd"):
            ok = agrees(got[key], ref[:, col[key]], fma[:, col[key]])
            bad[key] = np.flatnonzero(~ok)
        ok = agrees(got["totarea"], ref[:, col["totarea"]], fma[:, col["totarea"]])
        bad["totarea"] = np.flatnonzero(~ok)
        for key in ("la", "be"):
            ok = agrees(got[key], ref[:, col[key]], fma[:, col[key]], rel=0, abs_slack=1e-6)
            bad[key] = np.flatnonzero(~ok)
        assert all(len(v) == 0 for v in bad.values()), {k: [(int(n[i]), int(pole[i])) for i in v[:5]]
                                                       for k, v in bad.items() if len(v)}

        lines = ps.scan(fx["n_from"], fx["n_to"], conw_r)
        got_l = np.stack([lines[k] for k in ("per", "dev", "chi2", "dark", "la", "be")], 1)
        lr, lf = np.array(fx["lines_ref"], float), np.array(fx["lines_fma"], float)
        assert int(np.argmin(got_l[:, 1])) == int(np.argmin(lr[:, 1])), "best-period index"
        for c, (rel, slack) in enumerate([(1e-9, 0), (1e-9, 0), (1e-9, 0), (1e-9, 1e-9), (0, 1e-6), (0, 1e-6)]):
            ok = agrees(got_l[:, c], lr[:, c], lf[:, c], rel=rel, abs_slack=slack)
            assert ok.all(), (c, np.flatnonzero(~ok)[:5] + fx["n_from"])
        # line for line: wherever the reference's two roundings print the same line, the printed text
        # equals it, up to one unit in the last printed digit of a column (a value that agrees to
        # 1e-12 can still sit on the other side of a rounding boundary of %.6f)
        same_text = 0
        for i in range(len(got_l)):
            t, tr, tf = (format_line(dict(zip(("per", "dev", "chi2", "dark", "la", "be"), x)))
                         for x in (got_l[i], lr[i], lf[i]))
            same_text += t == tr
            if tr == tf and t != tr:
                assert last_digit_close(t, tr), (i + fx["n_from"], t, tr)
        print(f"{name}: {same_text} of {len(got_l)} lines print exactly as the reference core's")
    finally:
        ps.close()
