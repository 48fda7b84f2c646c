"""The on-disk formats end to end on the reference's own real inputs (all 100 sparse-survey
period_search_in files the reference ships in period_search/input_period_search/, copied to
tests/golden/real_inputs/): parsing and the setup tables against the oracle, bit for bit.  GPU scans
of three of them are in test_gpu_parity.py, the executable in test_executable.py and
test_gpu_executable.py."""
import glob
import os

import numpy as np
import pytest

from periodsearch_amd import PeriodSearch
from tests import oracle_binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REAL = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "real_inputs", "period_search_in_*")))


def test_real_inputs_present():
    assert len(REAL) == 100


@pytest.mark.parametrize("path", REAL, ids=[os.path.basename(p) for p in REAL])
def test_real_input_parse_and_setup_match_oracle(path):
    ps = PeriodSearch.from_file(path)
    c = ob.OracleCtx.from_file(path)
    inp = ps.input
    assert inp.ndata == c.ndata and inp.escl == c.escl
    assert (inp.lmax, inp.nrows) == (3, 6)
    g = ps.geometry()
    nor, da, ds = np.zeros(c.nfac * 3), np.zeros(c.nfac), np.zeros(c.nfac * c.ncoef)
    c.o.pso_ctx_geometry(c.c, ob._p(nor), ob._p(da), ob._p(ds))
    assert np.array_equal(g["nor"], nor.reshape(-1, 3))
    assert np.array_equal(g["darea"], da)
    assert np.array_equal(g["dsph"], ds.reshape(c.nfac, c.ncoef))
    cg = np.zeros(c.ma + 1)
    c.o.pso_ctx_cg_first(c.c, ob._p(cg))
    assert np.array_equal(g["cg_first"], cg[1:c.ncoef + 1])
    arr = ps.input_arrays()
    # lightcurve bookkeeping: point counts add up, unit direction vectors, mean-brightness sigmas
    assert int(np.sum(arr["lpoints"])) == inp.ndata
    assert np.all(np.abs(np.linalg.norm(arr["ee"], axis=1) - 1) < 1e-15)
    assert np.all(np.abs(np.linalg.norm(arr["ee0"], axis=1) - 1) < 1e-15)
    assert np.all(arr["sig"] > 0)
