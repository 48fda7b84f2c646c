"""The differentiation workload at the production shape (2400 points, Lmax 6, 288 facets): 4096 seeded
parameter vectors, one model evaluation each through ps_eval -- the forward dyda of every parameter
and the fused normal equations (J^T W J, J^T W r, chi2) -- against the reference core's values
committed in tests/golden/diff_config2.npz (mrqcof.c:16-148, bright.c:14-126), for both kernel forms
(the J-free one accumulates the normal equations without storing J), and the analytic dyda against
central finite differences of the GPU's own model."""
import os

import numpy as np
import pytest

from periodsearch_amd import PS_KERNEL_JFREE, PS_KERNEL_LEGACY, PeriodSearch, synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def prob():
    text, _ = synth.generate(synth.config1())
    ps = PeriodSearch.from_text(text).prepare(device=0)
    fx = np.load(os.path.join(GOLD, "diff_config2.npz"))
    yield ps, fx, synth.config2_vectors(ps)
    ps.close()


@pytest.mark.parametrize("kernel", [PS_KERNEL_JFREE, PS_KERNEL_LEGACY], ids=["jfree", "legacy"])
def test_all_vectors_match_reference_core(prob, kernel):
    ps, fx, vecs = prob
    ps.set_kernel(kernel)
    got = ps.eval(vecs, float(fx["conw_r"]))
    chi_err = np.abs(got["chisq"] - fx["chisq"]) / fx["chisq"]
    be_err = np.abs(got["beta"] - fx["beta"]).max(1) / np.abs(fx["beta"]).max(1)
    au = got["alpha"] @ fx["u"]
    au_err = np.abs(au - fx["alpha_u"]).max(1) / np.abs(fx["alpha_u"]).max(1)
    assert chi_err.max() <= 1e-12, chi_err.max()
    assert be_err.max() <= 1e-11, be_err.max()
    assert au_err.max() <= 1e-11, au_err.max()
    full = got["alpha"][fx["full_idx"]]
    assert np.abs(full - fx["alpha_full"]).max() <= 1e-11 * np.abs(fx["alpha_full"]).max()
    assert np.array_equal(full, np.transpose(full, (0, 2, 1)))  # symmetric as mrqcof.c:143-145 makes it
    # the forward rows of 16 vectors against the reference bright() at 64 points
    raw = ps.eval(vecs[fx["bright_vec"]], float(fx["conw_r"]), raw=True)
    y = raw["ymod"][:, fx["bright_pts"]]
    d = raw["dyda"][:, fx["bright_pts"], :]
    assert np.abs(y - fx["ymod"]).max() <= 1e-13 * np.abs(fx["ymod"]).max()
    d_err = np.abs(d - fx["dyda"]).max((0, 1)) / np.abs(fx["dyda"]).max((0, 1))
    assert d_err.max() <= 1e-11, (d_err.max(), int(d_err.argmax()))


def test_dyda_matches_central_differences(prob):
    ps, fx, vecs = prob
    ps.set_kernel(PS_KERNEL_JFREE)
    conw = float(fx["conw_r"])
    info = ps.info()
    ma, nc = info["ma"], info["ncoef"]
    span = np.abs(ps.geometry()["dsph"]).max(0)
    pts = np.array([0, 321, 1999, 2000, 2399])
    for v in (5, 3000):
        a = vecs[v]
        an = ps.eval(a[None], conw, raw=True)["dyda"][0][pts]
        # relative step 1e-6 (shape coefficients: of their basis scale); omega 1e-9 rad/d absolute, since
        # omega*t reaches ~1e5 rad on these epochs
        h = 1e-6 * np.maximum(np.abs(a), 1.0)
        h[:nc] = 1e-6 * np.maximum(np.abs(a[:nc]), 1.0 / span)
        h[nc + 2] = 1e-9
        plus, minus = np.tile(a, (ma, 1)), np.tile(a, (ma, 1))
        plus[np.arange(ma), np.arange(ma)] += h
        minus[np.arange(ma), np.arange(ma)] -= h
        yp = ps.eval(plus, conw, raw=True)["ymod"][:, pts]
        ym = ps.eval(minus, conw, raw=True)["ymod"][:, pts]
        fd = ((yp - ym) / (2 * h[:, None])).T
        err = np.abs(fd - an).max(0) / np.maximum(np.abs(an).max(0), 1e-300)
        assert err.max() <= 1e-5, (v, err.max(), int(err.argmax()))
