"""The differentiation-workload fixture (tests/golden/diff_config2.npz) is what it claims to be: the
reference core's mrqcof (mrqcof.c:16-148) and bright() (bright.c:14-126) at the seeded parameter
vectors of synth.config2_vectors on the production-shape input.  CPU only: a sample is recomputed on
the oracle, which is bit-exact with the reference core."""
import hashlib
import os

import numpy as np

from periodsearch_amd import PeriodSearch, synth
from tests import oracle_binding as ob

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_diff_fixture_is_the_reference_core(tmp_path):
    fx = np.load(os.path.join(GOLD, "diff_config2.npz"))
    text, _ = synth.generate(synth.config1())
    assert hashlib.sha256(text.encode()).hexdigest() == str(fx["text_sha256"])
    ps = PeriodSearch.from_text(text)
    vecs = synth.config2_vectors(ps)
    assert hashlib.sha256(vecs.tobytes()).hexdigest() == str(fx["vec_sha256"])
    path = tmp_path / "period_search_in"
    path.write_text(text)
    orc = ob.OracleCtx.from_file(str(path))
    orc.set_conw(float(fx["conw_r"]))
    for v in (0, 1717, 4095):
        al, be, chi = orc.mrqcof(vecs[v])
        assert chi == fx["chisq"][v]
        assert np.array_equal(be, fx["beta"][v])
        assert np.array_equal(al @ fx["u"], fx["alpha_u"][v])
    i = int(np.flatnonzero(fx["full_idx"] == 512)[0])
    assert np.array_equal(orc.mrqcof(vecs[512])[0], fx["alpha_full"][i])
    arr = ps.input_arrays()
    for i, j in ((0, 0), (7, 63), (15, 40)):
        v, p = int(fx["bright_vec"][i]), int(fx["bright_pts"][j])
        br, dy = orc.bright(vecs[v], arr["ee"][p], arr["ee0"][p], arr["tim"][p])
        assert br == fx["ymod"][i, j] and np.array_equal(dy, fx["dyda"][i, j])
