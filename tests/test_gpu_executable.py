"""The period_search_amd executable on the GPU, the reference's process contract end to end
(period_search/period_search_BOINC.cpp:147-954): in a working directory holding period_search_in it
writes period_search_out and period_search_state and exits 0.  The lines are the oracle's (the same
text, or -- on the rounding-sensitive lines -- within the validator's tolerances, the restated
compare_results of validator/period_search_validator4.cpp:71-109); a run resumed from the state file
after k lines (the reference's state: the line count, :85-101, :237-251) appends exactly the rest."""
import os
import subprocess

import numpy as np
import pytest

from periodsearch_amd import EXE_PATH, format_line, synth
from tests import oracle_binding as ob
from tests.test_oracle import validator_v4

pytestmark = pytest.mark.gpu


def test_executable_run_and_resume(tmp_path):
    text, _ = synth.generate(synth.config0(n_dense=4, dense_points=30, stop_condition=6, per_end=7.25))
    full = tmp_path / "full"
    full.mkdir()
    (full / "period_search_in").write_text(text)
    r = subprocess.run([EXE_PATH], cwd=full, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = (full / "period_search_out").read_text().splitlines(keepends=True)
    assert (full / "period_search_state").read_text().strip() == str(len(out))
    # the oracle's run of the same input (convexity-weight search, then every period)
    orc = ob.OracleCtx.from_text(text)
    cw, _ = orc.conw_search()
    _, _, _, nmax = orc.freq_grid()
    lines = orc.scan(1, nmax, cw)
    keys = ("per", "dev", "chi2", "dark", "la", "be")
    ref_text = [format_line(dict(zip(keys, x))) for x in lines]
    assert len(out) == len(ref_text) == nmax > 10
    same = sum(a == b for a, b in zip(out, ref_text))
    assert same >= len(out) - 3, (same, len(out))
    parse = lambda ls: np.array([[float(v) for v in l.split()] for l in ls])  # noqa: E731
    assert validator_v4(parse(out), parse(ref_text))
    # resume: the state file says k lines were written and the output holds them
    part = tmp_path / "part"
    part.mkdir()
    (part / "period_search_in").write_text(text)
    k = len(out) // 2
    (part / "period_search_out").write_text("".join(out[:k]))
    (part / "period_search_state").write_text(str(k))
    r = subprocess.run([EXE_PATH], cwd=part, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert (part / "period_search_out").read_text() == "".join(out)
    assert (part / "period_search_state").read_text().strip() == str(len(out))
