"""The committed scan fixtures (tests/golden/scan_<workload>.json) are what they claim to be: the
reference core's and the FMA oracle's fits of the seeded workloads, and lines formed from them with the
reference driver's arithmetic (period_search_BOINC.cpp:913-931).  CPU only; a few fits per fixture
are recomputed on the oracle, which is bit-exact with the reference core."""
import hashlib
import json
import os

import numpy as np
import pytest

from periodsearch_amd import synth
from tests import oracle_binding as ob

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name,sample", [("config0", 3), ("config1", 1), ("config3", 1)])
def test_scan_fixture_is_the_reference_core(name, sample, tmp_path):
    with open(os.path.join(GOLD, f"scan_{name}.json")) as f:
        fx = json.load(f)
    text, truth = synth.generate(getattr(synth, name)())
    assert hashlib.sha256(text.encode()).hexdigest() == fx["sha256"]
    ref, fma = np.array(fx["fits_ref"], float), np.array(fx["fits_fma"], float)
    nper = fx["n_to"] - fx["n_from"] + 1
    assert ref.shape == fma.shape == (10 * nper, 10)
    assert np.array_equal(ref[:, 0], np.repeat(np.arange(fx["n_from"], fx["n_to"] + 1), 10))
    # the window holds the true period's grid point
    path = tmp_path / "period_search_in"
    path.write_text(text)
    orc = ob.OracleCtx.from_file(str(path))
    orc_fma = ob.OracleCtx.from_file(str(path), fma=True)
    f0, _, df, nmax = orc.freq_grid()
    n_true = int(round((f0 - 1 / truth["period_h"]) / df)) + 1
    assert fx["n_from"] <= n_true <= fx["n_to"]
    if fx["conw_search"]:
        assert (fx["n_from"], fx["n_to"]) == (1, nmax)
    orc.set_conw(fx["conw_r"])
    orc_fma.set_conw(fx["conw_r"])
    rng = np.random.default_rng(7)
    for i in rng.choice(len(ref), sample, replace=False):
        n, m = int(ref[i, 0]), int(ref[i, 1])
        for o, rows in ((orc, ref), (orc_fma, fma)):
            r = o.fit(o.freq(n), m)
            got = [r["dev_new"], r["prd"], r["dark"], r["totarea"], r["la"], r["be"], r["chisq"], r["niter"]]
            assert np.array_equal(np.array(got, float), rows[i, 2:], equal_nan=True), (name, n, m)
    # lines from the fits: best pole by rms, first pole wins ties, lambda wrapped, chi2 = rms^2 * ndata
    lines = np.array(fx["lines_ref"], float)
    dev = ref[:, 2].reshape(nper, 10)
    k = np.argmin(dev, axis=1)
    best = ref.reshape(nper, 10, 10)[np.arange(nper), k]
    assert np.array_equal(lines[:, 0], 24 * best[:, 3])
    assert np.array_equal(lines[:, 1], best[:, 2])
    assert np.array_equal(lines[:, 2], best[:, 2] * best[:, 2] * fx["ndata"])
    la = np.where(best[:, 6] < 0, best[:, 6] + 360, best[:, 6])
    assert np.array_equal(lines[:, 4], la) and np.array_equal(lines[:, 5], best[:, 7])
