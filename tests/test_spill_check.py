"""The build-time guard against VGPR spills placed before a divergent join's exec restore
(periodsearch_amd/csrc/check_spill_placement.py): it flags the faulty shape, passes the correct one, and
the built device assembly passes it."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "periodsearch_amd", "csrc")
sys.path.insert(0, CSRC)
import check_spill_placement as chk  # noqa: E402

# exit block of a divergent loop: the loop leaves with exec == 0, the spill runs before the restore
BAD = """_Zkernel:
.LBB0_3:
	v_add_u32_e32 v1, 0x200, v1
	s_andn2_b64 exec, exec, s[2:3]
	s_cbranch_execnz .LBB0_3
.LBB0_4:
	s_mov_b32 s18, 0xfefa39ef
	scratch_store_dwordx2 off, v[150:151], off offset:92 ; 8-byte Folded Spill
	s_or_b64 exec, exec, s[0:1]
	s_barrier
	s_endpgm
"""
# the same spill after the restore
GOOD = """_Zkernel:
.LBB0_3:
	v_add_u32_e32 v1, 0x200, v1
	s_andn2_b64 exec, exec, s[2:3]
	s_cbranch_execnz .LBB0_3
.LBB0_4:
	s_or_b64 exec, exec, s[0:1]
	scratch_store_dwordx2 off, v[150:151], off offset:92 ; 8-byte Folded Spill
	s_endpgm
"""


def test_flags_spill_before_exec_restore():
    hits = chk.find(BAD.split("\n"))
    assert len(hits) == 1 and hits[0][0] == "_Zkernel" and hits[0][1] == 8


def test_passes_spill_after_restore():
    assert chk.find(GOOD.split("\n")) == []


def test_built_device_assembly_passes():
    asm = os.path.join(ROOT, "periodsearch_amd", "_build", "ps_device.s")
    if not os.path.exists(asm):
        pytest.skip("device assembly not built (make -C periodsearch_amd/csrc)")
    r = subprocess.run([sys.executable, os.path.join(CSRC, "check_spill_placement.py"), asm],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
