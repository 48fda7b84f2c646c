"""The batched uneven workload (512 seeded asteroids of periodsearch_amd.synth.config4: ndata 100-20000,
Lmax 3/4/6, nrows 6/8, convergence stop so that iteration counts differ) through ps_batch_run -- one
graph launch of persistent workgroups draining a device work queue -- against per-problem ps_fit_batch
(bit for bit) and, on the smallest asteroid of every (Lmax, nrows) class, against the oracle; in FP64
(the J-free queue kernel) and in mixed precision (the mixed queue kernel, against per-problem mixed
fits bit for bit and against the FP64 batch within the mixed path's accuracy).  The per-job body is
period_search_BOINC.cpp:817-921."""
import numpy as np
import pytest

from periodsearch_amd import PS_FP64, PS_MIXED, Batch, PeriodSearch, synth
from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu

FIELDS = ("dev", "prd", "dark", "totarea", "la", "be", "chisq")


@pytest.fixture(scope="module")
def config4():
    """All 512 asteroids prepared, each with a window of its true period's grid point x 10 poles."""
    texts, searches, freq, pole, prob, conw = [], [], [], [], [], []
    for i in range(synth.CONFIG4_N):
        text, truth = synth.generate(synth.config4(i))
        ps = PeriodSearch.from_text(text).prepare(device=0)
        info = ps.info()
        n_true = int(round((info["freq_start"] - 1 / truth["period_h"]) / info["freq_step"])) + 1
        n_true = min(max(n_true, 1), info["nmax"])
        texts.append(text)
        searches.append(ps)
        conw.append(info["conw_r0"])
        freq.extend(ps.freq([n_true] * 10))
        pole.extend(range(1, 11))
        prob.extend([i] * 10)
    yield texts, searches, np.array(conw), np.array(prob), np.array(freq), np.array(pole)
    for ps in searches:
        ps.close()


def test_config4_batch_matches_per_problem_fits(config4):
    texts, searches, conw, prob, freq, pole = config4
    with Batch(searches) as b:
        got = b.run(conw, prob, freq, pole)
        again = b.run(conw, prob, freq, pole)  # the kept graph, replayed
    assert np.array_equal(got, again)
    niter = got["niter"]
    assert niter.min() >= 10 and niter.max() <= 100 and len(np.unique(niter)) > 5, "uneven iteration counts"
    # per-problem launches give the same bits (every 8th asteroid: all shapes, sizes up to the largest)
    for i in range(0, len(searches), 8):
        sel = prob == i
        ref = searches[i].fit_batch(freq[sel], pole[sel], conw[i])
        for k in FIELDS + ("niter", "nsolve_fail"):
            assert np.array_equal(got[k][sel], ref[k], equal_nan=True), (i, k)


def test_config4_batch_matches_oracle_on_each_shape(config4):
    texts, searches, conw, prob, freq, pole = config4
    with Batch(searches) as b:
        got = b.run(conw, prob, freq, pole)
    ndata = np.array([ps.info()["ndata"] for ps in searches])
    shape = [(ps.info()["ncoef"], ps.info()["nfac"]) for ps in searches]
    checked = 0
    for key in sorted(set(shape)):
        i = min((j for j in range(len(searches)) if shape[j] == key), key=lambda j: ndata[j])
        orc = ob.OracleCtx.from_text(texts[i])
        orc.set_conw(conw[i])
        for m in (1, 6):
            j = np.flatnonzero((prob == i) & (pole == m))[0]
            r = orc.fit(freq[j], m)
            assert got["niter"][j] == r["niter"], (i, m)
            for k, rk in (("dev", "dev_new"), ("chisq", "chisq"), ("prd", "prd")):
                assert abs(got[k][j] - r[rk]) <= 1e-9 * abs(r[rk]), (i, m, k, got[k][j], r[rk])
            checked += 1
    assert checked == 2 * len(set(shape)) == 12


def test_config4_mixed_batch_matches_per_problem_mixed_fits(config4):
    texts, searches, conw, prob, freq, pole = config4
    with Batch(searches) as b:
        fp64 = b.run(conw, prob, freq, pole)
    for ps in searches:
        ps.set_precision(PS_MIXED)
    try:
        with Batch(searches) as b:
            got = b.run(conw, prob, freq, pole)
            again = b.run(conw, prob, freq, pole)
        for k in FIELDS + ("niter", "nsolve_fail"):  # the kept graph, replayed (a NaN matches a NaN)
            bad = np.flatnonzero(~((got[k] == again[k]) | (np.isnan(got[k]) & np.isnan(again[k]))
                                   if got[k].dtype.kind == "f" else got[k] == again[k]))
            assert len(bad) == 0, (k, bad[:10], got[bad[:3]], again[bad[:3]])
        for i in range(0, len(searches), 8):
            sel = prob == i
            ref = searches[i].fit_batch(freq[sel], pole[sel], conw[i])
            for k in FIELDS + ("niter", "nsolve_fail"):
                assert np.array_equal(got[k][sel], ref[k], equal_nan=True), (i, k)
    finally:
        for ps in searches:
            ps.set_precision(PS_FP64)
    # the mixed path's accuracy against FP64 (FP32 facet sums): chi2 per fit, median and bulk (measured:
    # median 3e-6, 90th percentile 1.2e-3; convexity rows weighted with a stale conw_r gave 0.1 and 1)
    ok = np.isfinite(fp64["chisq"]) & np.isfinite(got["chisq"])
    err = np.abs(got["chisq"][ok] - fp64["chisq"][ok]) / np.abs(fp64["chisq"][ok])
    print(f"mixed vs FP64 batch: {ok.mean():.4f} finite; chi2 rel err median {np.median(err):.3g}, "
          f"p90 {np.quantile(err, 0.9):.3g}, max {err.max():.3g}; same niter {np.mean(got['niter'] == fp64['niter']):.3f}")
    assert ok.mean() > 0.99
    assert np.median(err) < 1e-4 and np.quantile(err, 0.9) < 1e-2, (np.median(err), np.quantile(err, 0.9))
