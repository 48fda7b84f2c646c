"""The period_search_amd executable's exit statuses and messages where no device is needed (the
reference's contract, period_search/period_search_BOINC.cpp: exit(1) when a file cannot be opened,
:252-261; exit(2) on an input limit, :327, :353, :361).  A run on the GPU is in test_gpu_executable.py."""
import os
import subprocess

import pytest

from periodsearch_amd import EXE_PATH

pytestmark = pytest.mark.skipif(not os.path.exists(EXE_PATH), reason="period_search_amd not built")


def run(args, cwd):
    return subprocess.run([EXE_PATH] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_missing_input_exits_1(tmp_path):
    r = run([], tmp_path)  # no period_search_in in the working directory
    assert r.returncode == 1
    assert "period_search_in" in r.stderr


def test_malformed_input_exits_2(tmp_path):
    (tmp_path / "period_search_in").write_text("6.9 1.0 7.8 1\nnot a number\n")
    r = run([], tmp_path)
    assert r.returncode == 2
    assert "period_search_amd:" in r.stderr and not (tmp_path / "period_search_out").exists()


def test_reference_limit_exits_2(tmp_path):
    # a lightcurve of 1001 points: above the plain reference's POINTS_MAX (constants.h:1, check :352)
    head = ["6.9 1.0 7.8 1", "0", "0", "0.1", "3 3", "6", "0.5 1", "0.1 1", "-0.5 1", "0.1 0", "50", "10", "5",
            "0.1", "1", "1001 0"]
    pts = [f"{2455000 + 0.001 * i:.6f} 1.0 1 0 0 0 1 0" for i in range(1001)]
    (tmp_path / "period_search_in").write_text("\n".join(head + pts + ["1 1"]) + "\n")
    r = run([], tmp_path)
    assert r.returncode == 2 and "period_search_amd:" in r.stderr


def test_usage_exits_2(tmp_path):
    r = run(["--bogus"], tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr
