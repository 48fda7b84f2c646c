"""ps_fit_batch_noalloc, the entry point that allocates nothing (device-resident jobs and results, the
per-workgroup buffers in a static device array), against ps_fit_batch through the C ABI: two calls in a
row, the second with a smaller input, bit for bit.  A job with an invalid pole comes back with
niter = -1 (the kernel checks the jobs it reads from device memory)."""
import ctypes as C

import numpy as np
import pytest

from periodsearch_amd import FIT_DTYPE, PeriodSearch, lib, synth

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory from the HIP runtime libpsamd.so itself uses (torch bundles a second HIP runtime,
    which cannot share the process with it)."""

    _hip = None

    def __init__(self, host):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.nbytes, 1))) == 0
        assert self._hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def read(self, dtype):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        assert self._hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        self._hip.hipFree(self.ptr)


def noalloc(ps, freq, pole, conw_r):
    d_freq = DeviceBuffer(np.ascontiguousarray(freq, np.float64))
    d_pole = DeviceBuffer(np.ascontiguousarray(pole, np.int32))
    d_out = DeviceBuffer(np.zeros(len(freq), dtype=FIT_DTYPE))
    try:
        rc = lib().ps_fit_batch_noalloc(ps._handle(), len(freq), d_freq.ptr, d_pole.ptr, float(conw_r), d_out.ptr)
        assert rc == 0, lib().ps_last_error().decode()
        return d_out.read(FIT_DTYPE)
    finally:
        for b in (d_freq, d_pole, d_out):
            b.free()


def test_noalloc_matches_fit_batch_twice():
    text, truth = synth.generate(synth.config1())
    ps = PeriodSearch.from_text(text).prepare(device=0)
    try:
        info = ps.info()
        conw = info["conw_r0"]
        n_true = int(round((info["freq_start"] - 1 / truth["period_h"]) / info["freq_step"])) + 1
        for nper in (60, 7):  # the second call smaller than the first
            n = n_true - nper // 2 + np.arange(nper)
            freq = ps.freq(np.repeat(n, 10))
            pole = np.tile(np.arange(1, 11), nper).astype(np.int32)
            ref = ps.fit_batch(freq, pole, conw)
            got = noalloc(ps, freq, pole, conw)
            for k in FIT_DTYPE.names:
                assert np.array_equal(got[k], ref[k], equal_nan=True), (nper, k)
        bad = noalloc(ps, freq[:3], np.array([1, 0, 3], np.int32), conw)
        assert list(bad["niter"]) == [ref["niter"][0], -1, ref["niter"][2]]
        assert np.isnan(bad["dev"][1]) and bad["dev"][0] == ref["dev"][0]
    finally:
        ps.close()
