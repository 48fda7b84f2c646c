"""The autotuner (ps_autotune): it times both kernel forms on a sample of the problem's own fits, keeps
the faster, and remembers the choice per problem shape."""
import pytest

from periodsearch_amd import PS_KERNEL_JFREE, PS_KERNEL_LEGACY, PS_MIXED, PeriodSearch, synth

pytestmark = pytest.mark.gpu


def test_autotune_picks_the_faster_form_and_remembers_it():
    text, _ = synth.generate(synth.config0())
    ps = PeriodSearch.from_text(text).prepare(device=0)
    try:
        conw = ps.info()["conw_r0"]
        kern, ms_legacy, ms_jfree = ps.autotune(conw)
        assert ms_legacy > 0 and ms_jfree > 0
        assert kern == (PS_KERNEL_JFREE if ms_jfree <= ms_legacy else PS_KERNEL_LEGACY)
        assert ps.kernel() == kern
        # a second problem of the same shape gets the stored times without timing again
        ps2 = PeriodSearch.from_text(text).prepare(device=0)
        assert ps2.autotune(conw) == (kern, ms_legacy, ms_jfree)
        ps2.close()
        # mixed precision has one kernel form
        ps.set_precision(PS_MIXED)
        assert ps.autotune(conw)[0] == PS_KERNEL_LEGACY
    finally:
        ps.close()
