"""MI355X-native convex-inversion period scan (the hot path of Asteroids@home period_search).

The product is libpsamd.so (HIP kernels for gfx950 + host C++ behind the C ABI in
include/periodsearch_amd.h) and the period_search_amd executable; this package holds the ctypes
front end (api), the synthetic-input generator (synth) and the build helper.
"""
from ._lib import (EXE_PATH, LIB_PATH, N_POLES, PS_FP64, PS_KERNEL_AUTO, PS_KERNEL_JFREE, PS_KERNEL_LEGACY,  # noqa: F401
                   PS_MIXED, PsError, lib)
from .api import FIT_DTYPE, LINE_DTYPE, Batch, PeriodSearch, fit_multi, format_line, run_files  # noqa: F401

__all__ = ["PeriodSearch", "Batch", "fit_multi", "format_line", "run_files", "PS_FP64", "PS_MIXED", "N_POLES", "PsError", "lib",
           "LIB_PATH", "EXE_PATH", "LINE_DTYPE", "FIT_DTYPE"]
