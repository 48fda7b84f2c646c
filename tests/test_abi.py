"""The C ABI of libpsamd.so without a GPU: exports, host-side parsing/setup parity with the oracle,
error paths and the output format.  No kernel is launched here."""
import os
import re

import numpy as np
import pytest

from periodsearch_amd import PeriodSearch, PsError, format_line, lib, synth
from periodsearch_amd import _lib as L
from tests import oracle_binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
HEADER = os.path.join(ROOT, "include", "periodsearch_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 18
    L_ = lib()
    for s in syms:
        assert hasattr(L_, s), s
    assert set(syms) == set(L.SIGNATURES), "ctypes signatures out of sync with the header"
    assert b"gfx950" in L_.ps_version()


def test_gfx950_code_object_present():
    data = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in data


INPUTS = [os.path.join(GOLD, "fixture118_period_search_in"), os.path.join(GOLD, "sample_period_search_in")]


@pytest.mark.parametrize("path", INPUTS)
def test_parse_and_setup_match_oracle_bitwise(path):
    ps = PeriodSearch.from_file(path)
    c = ob.OracleCtx.from_file(path)
    g = ps.geometry()
    nor, da, ds = np.zeros(c.nfac * 3), np.zeros(c.nfac), np.zeros(c.nfac * c.ncoef)
    c.o.pso_ctx_geometry(c.c, ob._p(nor), ob._p(da), ob._p(ds))
    assert np.array_equal(g["nor"], nor.reshape(-1, 3))
    assert np.array_equal(g["darea"], da)
    assert np.array_equal(g["dsph"], ds.reshape(c.nfac, c.ncoef))
    cg = np.zeros(c.ma + 1)
    c.o.pso_ctx_cg_first(c.c, ob._p(cg))
    assert np.array_equal(g["cg_first"], cg[1:c.ncoef + 1])
    assert ps.input.escl == c.escl
    assert ps.input.ndata == c.ndata


def test_synthetic_inputs_parse_identically():
    for spec in (synth.config0(), synth.config1()):
        text, truth = synth.generate(spec)
        ps = PeriodSearch.from_text(text)
        c = ob.OracleCtx.from_text(text)
        assert ps.input.ndata == c.ndata == truth["ndata"]
        assert ps.input.escl == c.escl
        arr = ps.input_arrays()
        assert np.all(np.abs(np.linalg.norm(arr["ee"], axis=1) - 1) < 1e-15)


def test_synthetic_generator_is_seeded():
    t1, _ = synth.generate(synth.config0())
    t2, _ = synth.generate(synth.config0())
    t3, _ = synth.generate(synth.config0(seed=5))
    assert t1 == t2 and t1 != t3


def test_parse_errors():
    with pytest.raises(PsError) as e:
        PeriodSearch.from_file("/nonexistent/period_search_in")
    assert e.value.code == -7
    with pytest.raises(PsError) as e:
        PeriodSearch.from_text("garbage")
    assert e.value.code == -2
    text, _ = synth.generate(synth.config0())
    lines = text.splitlines()
    lines[15] = "1001 0"  # POINTS_MAX = 1000 (period_search_BOINC.cpp:352)
    with pytest.raises(PsError) as e:
        PeriodSearch.from_text("\n".join(lines))
    assert e.value.code == -3


def test_format_line_matches_reference_printf():
    line = {"per": 89.99905543, "dev": 0.108304, "chi2": 5.864857, "dark": 0.1, "la": 176.2, "be": 5.6}
    assert format_line(line) == "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n" % (89.99905543, 0.108304, 5.864857, 0.1,
                                                                         176.2, 5.6)
    ref = open(os.path.join(GOLD, "sample_period_search_out")).read().replace("\r", "").splitlines()[0]
    vals = [float(x) for x in ref.split()]
    assert format_line(dict(zip(["per", "dev", "chi2", "dark", "la", "be"], vals))).rstrip("\n") == ref


def test_device_calls_need_prepare():
    text, _ = synth.generate(synth.config0())
    ps = PeriodSearch.from_text(text)
    with pytest.raises(RuntimeError):
        ps.fit_batch([0.14], [1], 0.1)


def test_prepare_rejects_shapes_beyond_the_limits():
    """Shape limits are checked before a device is touched.  The reference's own limit of 200
    lightcurves (MAX_LC, period_search_BOINC.cpp:324-327; PS_ERR_LIMIT) binds before the kernels' LDS
    normalisation table, which holds (lightcurves + 1) * (ma + 4) <= 288 * 49 doubles (230 lightcurves
    at Lmax 6, ma = 57); 200 lightcurves pass both (and, without a GPU here, fail on the device).  Lmax 7 (ma = 72 > 63) is
    beyond the kernels (PS_ERR_UNSUPPORTED)."""
    codes = {k: v for v, k in L.ERRORS.items()}

    def code(**kw):
        text, _ = synth.generate(synth.config0(dense_points=3, **kw))
        with pytest.raises(PsError) as e:
            PeriodSearch.from_text(text).prepare(device=0)
        return e.value.code
    assert code(fit_lmax=6, n_dense=201) == codes["PS_ERR_LIMIT"]
    assert code(fit_lmax=6, n_dense=200) not in (codes["PS_ERR_LIMIT"], codes["PS_ERR_UNSUPPORTED"])
    assert code(fit_lmax=7, n_dense=4) == codes["PS_ERR_UNSUPPORTED"]
    # Lmax 8 and 10 (ma = 89, 129: the reference's MAX_LM = 10 is the parser's limit) are rejected the same
    # way, before any per-parameter table is filled
    assert code(fit_lmax=8, n_dense=4) == codes["PS_ERR_UNSUPPORTED"]
    assert code(fit_lmax=10, n_dense=4) == codes["PS_ERR_UNSUPPORTED"]


def test_prepare_rejects_an_inconsistent_lightcurve_table():
    """A caller that fills ps_input itself must describe exactly ndata points with the lightcurve
    lengths (the kernels index points through that table): a mismatch is PS_ERR_ARG, found before a
    device is touched."""
    codes = {k: v for v, k in L.ERRORS.items()}
    text, _ = synth.generate(synth.config0(n_dense=3, dense_points=10))
    ps = PeriodSearch.from_text(text)
    ps.input.lpoints[0] += 1
    with pytest.raises(PsError) as e:
        ps.prepare(device=0)
    assert e.value.code == codes["PS_ERR_ARG"]
    ps.input.lpoints[0] -= 1  # consistent again, but one lightcurve empty
    ps.input.lpoints[0] += ps.input.lpoints[1]
    ps.input.lpoints[1] = 0
    with pytest.raises(PsError) as e:
        ps.prepare(device=0)
    assert e.value.code == codes["PS_ERR_ARG"]
    ps.close()
