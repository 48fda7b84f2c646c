"""The damped system goes singular: with the phase-function amplitude a = 0 (and a free), the derivative
of the model with respect to the phase parameter d, a*exp(-alpha/d)*alpha/d^2 (phasec.c:16-21), is zero
at every point, so alpha has a zero row and column and its damped form (diagonal x (1 + lambda),
mrqmin.c:53-59) stays singular.  gauss_errc fails on every step (a zero pivot, gauss_errc.c:45-50):
the step is skipped, the iteration still counts (mrqmin.c:61-63, period_search_BOINC.cpp:870-871) and
the parameters stay at their starting values.  Both kernel forms, against the oracle: the iteration
count, the fitted period and pole (unchanged), the convexity figures; every step counted as a solver
failure.  The reported chi2 and rms are the starting point's (the oracle's mrqcof at the initial
parameters): on a failure in the first step the reference would report its previous fit's global chi2
instead."""
import numpy as np
import pytest

from periodsearch_amd import PS_KERNEL_JFREE, PS_KERNEL_LEGACY, PeriodSearch, synth
from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel", [PS_KERNEL_JFREE, PS_KERNEL_LEGACY], ids=["jfree", "legacy"])
def test_singular_damped_system_skips_every_step(kernel):
    text, _ = synth.generate(synth.config0(n_dense=4, dense_points=30, stop_condition=8, phase=(0.0, 0.1, -0.5)))
    ps = PeriodSearch.from_text(text).prepare(device=0)
    ps.set_kernel(kernel)
    orc = ob.OracleCtx.from_text(text)
    try:
        conw = ps.info()["conw_r0"]
        orc.set_conw(conw)
        freq = ps.freq([3, 3, 3])
        pole = np.array([1, 6, 9])
        got = ps.fit_batch(freq, pole, conw)
        ndata = ps.info()["ndata"]
        for i in range(3):
            ref = orc.fit(freq[i], int(pole[i]))
            a0 = orc.init_params(freq[i], int(pole[i]))
            nc = orc.ncoef
            assert got["niter"][i] == ref["niter"] == 8
            assert got["nsolve_fail"][i] == 8
            # the starting period and pole (degrees converted back with an ulp of difference at most)
            assert got["prd"][i] == ref["prd"] == 2 * np.pi / a0[nc + 2]
            assert abs(got["la"][i] - ref["la"]) <= 1e-12 and abs(got["be"][i] - ref["be"]) <= 1e-12
            # |sum Area*n| of the starting ellipsoid is rounding noise (a closed surface): bounded by the area
            assert abs(got["totarea"][i] - ref["totarea"]) <= 1e-12 * ref["totarea"]
            assert abs(got["dark"][i] - ref["dark"]) <= 1e-12 * ref["totarea"]
            chi0 = orc.mrqcof(a0)[2]
            assert abs(got["chisq"][i] - chi0) <= 1e-12 * chi0
            rchisq = chi0 - got["dark"][i] ** 2 * conw * conw
            assert abs(got["dev"][i] - np.sqrt(rchisq / ndata)) <= 1e-12 * got["dev"][i]
    finally:
        ps.close()
