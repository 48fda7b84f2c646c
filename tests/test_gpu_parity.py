"""GPU parity of the HIP fit kernels against the oracle (FP64 path), through the C ABI.

Tolerances: chi2 / rms / fitted parameters within 1e-9 relative (the FP64 parity bar); normal
equations within 1e-11 of the matrix scale; fitted-period index and best-pole selection exact.
"""
import numpy as np
import pytest

from periodsearch_amd import PeriodSearch, synth
from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu

SMALL = synth.config0(n_dense=4, dense_points=30, stop_condition=10)


@pytest.fixture(scope="module")
def small():
    text, _ = synth.generate(SMALL)
    ps = PeriodSearch.from_text(text).prepare(device=0)
    orc = ob.OracleCtx.from_text(text)
    yield ps, orc
    ps.close()


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_eval_matches_oracle_mrqcof(small):
    ps, orc = small
    conw_r = ps.info()["conw_r0"]
    orc.set_conw(conw_r)
    rng = np.random.default_rng(5)
    vecs = []
    for pole in (1, 4, 7):
        a = orc.init_params(orc.freq(1), pole)
        a[: orc.ncoef] += rng.normal(0, 0.05, orc.ncoef)
        vecs.append(a)
    vecs = np.array(vecs)
    got = ps.eval(vecs, conw_r, raw=True)
    for v, a in enumerate(vecs):
        al, be, chi = orc.mrqcof(a)
        scale = np.abs(al).max()
        assert np.abs(got["alpha"][v] - al).max() <= 1e-11 * scale
        assert np.abs(got["beta"][v] - be).max() <= 1e-11 * np.abs(be).max()
        assert abs(got["chisq"][v] - chi) <= 1e-12 * chi
        # raw bright() outputs of the first observed points
        arr = ps.input_arrays()
        for p in (0, 7, 33):
            br, dy = orc.bright(a, arr["ee"][p], arr["ee0"][p], arr["tim"][p])
            assert abs(got["ymod"][v, p] - br) <= 1e-13 * abs(br)
            assert np.abs(got["dyda"][v, p] - dy).max() <= 1e-12 * np.abs(dy).max()


def test_fit_batch_matches_oracle(small):
    ps, orc = small
    conw_r = ps.info()["conw_r0"]
    orc.set_conw(conw_r)
    n = np.array([1, 1, 1, 2, 2, 3])
    poles = np.array([1, 5, 10, 2, 8, 6])
    got = ps.fit_batch(ps.freq(n), poles, conw_r)
    for i in range(len(n)):
        ref = orc.fit(orc.freq(int(n[i])), int(poles[i]))
        assert got["niter"][i] == ref["niter"]
        assert _rel(got["dev"][i], ref["dev_new"]) < 1e-9
        assert _rel(got["chisq"][i], ref["chisq"]) < 1e-9
        assert _rel(got["prd"][i], ref["prd"]) < 1e-9
        assert _rel(got["totarea"][i], ref["totarea"]) < 1e-9
        assert abs(got["la"][i] - ref["la"]) < 1e-6 and abs(got["be"][i] - ref["be"]) < 1e-6


def test_scan_and_conw_match_oracle(small):
    ps, orc = small
    cw_gpu, nc_gpu = ps.conw_search()
    cw_orc, nc_orc = orc.conw_search()
    assert cw_gpu == cw_orc and nc_gpu == nc_orc
    lines = ps.scan(1, 4, cw_gpu)
    ref = orc.scan(1, 4, cw_orc)
    got = np.stack([lines[k] for k in ("per", "dev", "chi2", "dark", "la", "be")], 1)
    assert _rel(got[:, :3], ref[:, :3]) < 1e-9
    assert np.argmin(got[:, 1]) == np.argmin(ref[:, 1])
    assert np.abs(got[:, 3:] - ref[:, 3:]).max() < 1e-4
