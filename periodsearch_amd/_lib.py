"""ctypes bindings of the C ABI in include/periodsearch_amd.h (libpsamd.so, built in-tree).

The library is the product: every entry point here forwards to it.  Loading fails loudly when the
shared library has not been built; there is no Python fallback for any computation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsamd.so")
EXE_PATH = os.path.join(_HERE, "bin", "period_search_amd")

PS_OK = 0
PS_FP64 = 0
PS_MIXED = 1
PS_KERNEL_AUTO = 0
PS_KERNEL_LEGACY = 1
PS_KERNEL_JFREE = 2
N_POLES = 10

ERRORS = {
    -1: "PS_ERR_ARG",
    -2: "PS_ERR_PARSE",
    -3: "PS_ERR_LIMIT",
    -4: "PS_ERR_DEVICE",
    -5: "PS_ERR_UNSUPPORTED",
    -6: "PS_ERR_NOMEM",
    -7: "PS_ERR_IO",
}

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class PsInput(C.Structure):
    _fields_ = [
        ("per_start", C.c_double), ("per_step_coef", C.c_double), ("per_end", C.c_double),
        ("ia_prd", C.c_int), ("jd_00", C.c_double), ("phi0", C.c_double), ("conw", C.c_double),
        ("lmax", C.c_int), ("mmax", C.c_int), ("nrows", C.c_int),
        ("par", C.c_double * 3), ("ia_par", C.c_int * 3), ("cl", C.c_double), ("ia_cl", C.c_int),
        ("stop_condition", C.c_double), ("n_iter_min", C.c_int),
        ("alamda_incr", C.c_double), ("alamda_start", C.c_double),
        ("lcurves", C.c_int), ("lpoints", IP), ("inrel", IP), ("ndata", C.c_int),
        ("tim", DP), ("brightness", DP), ("sig", DP), ("weight", DP), ("ee", DP), ("ee0", DP),
        ("jd_min", C.c_double), ("jd_max", C.c_double), ("jd_0", C.c_double),
        ("onlyrel", C.c_int), ("escl", C.c_double),
        ("a_axis", C.c_double), ("b_axis", C.c_double), ("c_axis", C.c_double), ("owned", C.c_int),
    ]


class PsGeometry(C.Structure):
    _fields_ = [
        ("nrows", C.c_int), ("nfac", C.c_int), ("ndir", C.c_int), ("lmax", C.c_int), ("mmax", C.c_int),
        ("ncoef", C.c_int), ("npairs", C.c_int),
        ("nor", DP), ("darea", DP), ("dsph", DP), ("pleg", DP), ("fc", DP), ("fs", DP),
        ("pair_m", IP), ("pair_l", IP), ("cg_first", DP),
    ]


class PsLine(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("per", "dev", "chi2", "dark", "la", "be")]


class PsFitResult(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("dev", "prd", "dark", "totarea", "la", "be", "chisq")] + [
        ("niter", C.c_int), ("nsolve_fail", C.c_int)]


class PsInfo(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("ndata", "npts", "nlc", "nfac", "ncoef", "ma", "mfit", "nmax")] + [
        (k, C.c_double) for k in ("freq_start", "freq_end", "freq_step", "escl", "conw_r0")] + [
        ("slots", C.c_int), ("precision", C.c_int)]


# Every symbol declared in include/periodsearch_amd.h, with its ctypes signature.
SIGNATURES = {
    "ps_last_error": (C.c_char_p, []),
    "ps_version": (C.c_char_p, []),
    "ps_input_read": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(PsInput))]),
    "ps_input_parse": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(PsInput))]),
    "ps_input_free": (None, [C.POINTER(PsInput)]),
    "ps_geometry_build": (C.c_int, [C.POINTER(PsInput), C.POINTER(C.POINTER(PsGeometry))]),
    "ps_geometry_free": (None, [C.POINTER(PsGeometry)]),
    "ps_prepare": (C.c_int, [C.POINTER(PsInput), C.POINTER(PsGeometry), C.c_int, C.POINTER(C.c_void_p)]),
    "ps_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "ps_set_kernel": (C.c_int, [C.c_void_p, C.c_int]),
    "ps_get_kernel": (C.c_int, [C.c_void_p]),
    "ps_autotune": (C.c_int, [C.c_void_p, C.c_double, IP, DP]),
    "ps_get_info": (C.c_int, [C.c_void_p, C.POINTER(PsInfo)]),
    "ps_fit_batch": (C.c_int, [C.c_void_p, C.c_int, DP, IP, C.c_double, C.POINTER(PsFitResult)]),
    "ps_fit_batch_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "ps_fit_multi": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, DP, C.c_int, IP, DP, IP, C.POINTER(PsFitResult)]),
    "ps_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "ps_batch_run": (C.c_int, [C.c_void_p, DP, C.c_int, IP, DP, IP, C.POINTER(PsFitResult)]),
    "ps_batch_kernel_time": (C.c_int, [C.c_void_p, DP, IP, C.c_int]),
    "ps_batch_release": (None, [C.c_void_p]),
    "ps_conw_round": (C.c_int, [C.c_void_p, C.c_double, DP]),
    "ps_conw_search": (C.c_int, [C.c_void_p, DP, IP]),
    "ps_scan": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.POINTER(PsLine)]),
    "ps_eval": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_double, DP, DP, DP, DP, DP]),
    "ps_kernel_time": (C.c_int, [C.c_void_p, DP, IP, C.c_int]),
    "ps_release": (None, [C.c_void_p]),
    "ps_format_line": (C.c_int, [C.POINTER(PsLine), C.c_char_p, C.c_int]),
    "ps_run_files": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int]),
}

_lib = None


def lib():
    """Loads libpsamd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
                "or make -C periodsearch_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class PsError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().ps_last_error().decode(errors="replace")
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(rc, where):
    if rc < 0:
        raise PsError(rc, where)
    return rc
