"""ctypes binding of the oracle (oracle/libps_oracle.so) and, where it was built, of the reference
core (oracle/_ref/libps_ref.so).  Test infrastructure: the checker, never the thing measured."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libps_oracle.so")
ORACLE_FMA_SO = os.path.join(ROOT, "oracle", "libps_oracle_fma.so")  # the same restatement, FMA-contracted
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libps_ref.so")
ORACLE_CLI = os.path.join(ROOT, "oracle", "ps_oracle_cli")

VP = C.c_void_p
DP = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(DP)


class _Line(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("per", "dev", "chi2", "dark", "la", "be")]


class _Fit(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("dev_new", "prd", "dark", "totarea", "la", "be", "chisq")] + [
        ("niter", C.c_int)]


_libs = {}


def oracle(fma=False):
    """The oracle library: the plain FP64 build, or (fma=True) the FMA-contracted build -- a second
    valid FP64 rounding of the reference arithmetic."""
    path = ORACLE_FMA_SO if fma else ORACLE_SO
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{os.path.relpath(path, ROOT)} missing: run make -C oracle")
        o = C.CDLL(path)
        o.pso_ctx_from_file.restype = VP
        o.pso_ctx_from_file.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        o.pso_ctx_from_text.restype = VP
        o.pso_ctx_from_text.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        for name in ("pso_ctx_free", "pso_ctx_ma", "pso_ctx_ncoef", "pso_ctx_nfac", "pso_ctx_mfit", "pso_ctx_ndata",
                     "pso_ctx_stop_iters"):
            getattr(o, name).argtypes = [VP]
        o.pso_ctx_escl.restype = C.c_double
        o.pso_ctx_escl.argtypes = [VP]
        o.pso_ctx_conw_r.restype = C.c_double
        o.pso_ctx_conw_r.argtypes = [VP]
        o.pso_ctx_freq.restype = C.c_double
        o.pso_ctx_freq.argtypes = [VP, DP, DP, DP, C.POINTER(C.c_int)]
        o.pso_ctx_set_conw.argtypes = [VP, C.c_double]
        o.pso_ctx_init_params.argtypes = [VP, C.c_double, C.c_int, DP]
        o.pso_ctx_ia.argtypes = [VP, C.POINTER(C.c_int)]
        o.pso_ctx_cg_first.argtypes = [VP, DP]
        o.pso_ctx_geometry.argtypes = [VP, DP, DP, DP]
        o.pso_fit.argtypes = [VP, C.c_double, C.c_int, C.POINTER(_Fit)]
        o.pso_conw_search.argtypes = [VP, DP, C.POINTER(C.c_int)]
        o.pso_scan.argtypes = [VP, C.c_int, C.c_int, C.c_double, C.POINTER(_Line)]
        o.pso_mrqcof.restype = C.c_double
        o.pso_mrqcof.argtypes = [VP, DP, C.POINTER(C.c_int), C.c_int, DP, DP]
        o.pso_bright.restype = C.c_double
        o.pso_bright.argtypes = [VP, DP, DP, C.c_double, DP, DP, C.c_int]
        o.pso_conv.restype = C.c_double
        o.pso_conv.argtypes = [VP, C.c_int, DP, C.c_int]
        o.pso_curv.argtypes = [VP, DP]
        o.pso_blmatrix.argtypes = [VP, C.c_double, C.c_double]
        o.pso_ctx_set_area_dg.argtypes = [VP, DP, DP]
        o.pso_ctx_get_area_dg.argtypes = [VP, DP, DP]
        o.pso_ctx_set_blmat.argtypes = [VP, DP, DP]
        o.pso_gauss_errc.argtypes = [DP, C.c_int, DP]
        o.pso_phasec.restype = C.c_double
        o.pso_phasec.argtypes = [DP, C.c_double, DP]
        o.pso_trifac.argtypes = [C.c_int, C.POINTER(C.c_int)]
        o.pso_vertex_dirs.argtypes = [C.c_int, DP, DP]
        o.pso_areanorm.argtypes = [DP, DP, C.c_int, C.c_int, C.POINTER(C.c_int), DP, DP, DP, DP]
        o.pso_sphfunc.argtypes = [C.c_int, DP, DP, C.c_int, C.c_int, DP, DP, DP, DP]
        o.pso_ellfit.argtypes = [DP, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, DP, DP, DP, C.c_int,
                                 C.c_int]
        o.pso_ncoef.argtypes = [C.c_int, C.c_int]
        _libs[path] = o
    return _libs[path]


def ref_available():
    return os.path.exists(REF_SO)


class OracleCtx:
    """One parsed problem in the oracle (period_search_BOINC.cpp state, without globals)."""

    def __init__(self, ptr, fma=False):
        self.o = oracle(fma)
        self.c = VP(ptr)
        self.ma = self.o.pso_ctx_ma(self.c)
        self.ncoef = self.o.pso_ctx_ncoef(self.c)
        self.nfac = self.o.pso_ctx_nfac(self.c)
        self.mfit = self.o.pso_ctx_mfit(self.c)
        self.ndata = self.o.pso_ctx_ndata(self.c)

    @classmethod
    def from_file(cls, path, fma=False):
        err = C.c_int()
        p = oracle(fma).pso_ctx_from_file(str(path).encode(), C.byref(err))
        if not p:
            raise RuntimeError(f"oracle cannot load {path}: {err.value}")
        return cls(p, fma)

    @classmethod
    def from_text(cls, text, fma=False):
        err = C.c_int()
        p = oracle(fma).pso_ctx_from_text(text.encode(), C.byref(err))
        if not p:
            raise RuntimeError(f"oracle cannot parse input: {err.value}")
        return cls(p, fma)

    def __del__(self):
        try:
            self.o.pso_ctx_free(self.c)
        except Exception:
            pass

    def freq_grid(self):
        f0, f1, df = C.c_double(), C.c_double(), C.c_double()
        nmax = C.c_int()
        self.o.pso_ctx_freq(self.c, C.byref(f0), C.byref(f1), C.byref(df), C.byref(nmax))
        return f0.value, f1.value, df.value, nmax.value

    def freq(self, n):
        f0, _, df, _ = self.freq_grid()
        return f0 - (n - 1) * df

    @property
    def escl(self):
        return self.o.pso_ctx_escl(self.c)

    @property
    def conw_r0(self):
        return self.o.pso_ctx_conw_r(self.c)

    def set_conw(self, conw_r):
        self.o.pso_ctx_set_conw(self.c, conw_r)

    def ia(self):
        a = (C.c_int * (self.ma + 1))()
        self.o.pso_ctx_ia(self.c, a)
        return np.array(a[1:])

    def init_params(self, freq, pole):
        cg = np.zeros(self.ma + 1)
        self.o.pso_ctx_init_params(self.c, freq, pole, _p(cg))
        return cg[1:].copy()

    def fit(self, freq, pole):
        r = _Fit()
        self.o.pso_fit(self.c, freq, pole, C.byref(r))
        return {k: getattr(r, k) for k, _ in _Fit._fields_}

    def conw_search(self):
        cw = C.c_double()
        rounds = C.c_int()
        nc = self.o.pso_conw_search(self.c, C.byref(cw), C.byref(rounds))
        return cw.value, nc

    def scan(self, n_from, n_to, conw_r):
        lines = (_Line * (n_to - n_from + 1))()
        cnt = self.o.pso_scan(self.c, n_from, n_to, conw_r, lines)
        return np.array([[getattr(l, k) for k, _ in _Line._fields_] for l in lines[:cnt]])

    def mrqcof(self, a):
        """alpha (mfit x mfit), beta (mfit), chi2 for 0-based parameters a (conv rows included)."""
        cg = np.concatenate([[0.0], a])
        ia = (C.c_int * (self.ma + 1))()
        self.o.pso_ctx_ia(self.c, ia)
        al = np.zeros((self.mfit + 1) ** 2)
        be = np.zeros(self.ma + 1)
        chi = self.o.pso_mrqcof(self.c, _p(cg), ia, self.ma, _p(al), _p(be))
        return al.reshape(self.mfit + 1, self.mfit + 1)[1:, 1:].copy(), be[1:self.mfit + 1].copy(), chi

    def conv(self, a, comp):
        """conv() (value, dres[ma]) for convexity component comp in 1..3 after curv of a."""
        cg = np.concatenate([[0.0], a])
        self.o.pso_curv(self.c, _p(cg))
        dres = np.zeros(self.ma + 1)
        res = self.o.pso_conv(self.c, comp, _p(dres), self.ma)
        return res, dres[1:].copy()

    def bright(self, a, ee, ee0, t):
        """bright() after curv/blmatrix of a: (brightness, dyda[ma])."""
        cg = np.concatenate([[0.0], a])
        self.o.pso_curv(self.c, _p(cg))
        self.o.pso_blmatrix(self.c, cg[self.ncoef + 1], cg[self.ncoef + 2])
        e = np.concatenate([[0.0], ee])
        e0 = np.concatenate([[0.0], ee0])
        dy = np.zeros(self.ma + 1)
        br = self.o.pso_bright(self.c, _p(e), _p(e0), t, _p(cg), _p(dy), self.ma)
        return br, dy[1:].copy()


_r = None


def reference():
    """The reference core (g++ -O3 build of the reference's own sources) with the harness."""
    global _r
    if _r is None:
        r = C.CDLL(REF_SO)
        r.psr_load_file.argtypes = [C.c_char_p]
        r.psr_run.argtypes = [C.c_int, C.c_int, DP, DP]
        r.psr_fit.argtypes = [C.c_double, C.c_int, DP]
        r.psr_set_conw.argtypes = [C.c_double]
        r.psr_mrqcof.restype = C.c_double
        r.psr_mrqcof.argtypes = [DP, DP, DP]
        r.psr_bright.restype = C.c_double
        r.psr_bright.argtypes = [DP, DP, C.c_double, DP, DP, C.c_int, DP]
        r.psr_curv.argtypes = [DP, DP, DP]
        r.psr_blmatrix.argtypes = [C.c_double, C.c_double]
        r.psr_geometry.argtypes = [DP, DP, DP, DP, DP]
        r.psr_cg_first.argtypes = [DP]
        r.psr_gauss_errc.argtypes = [DP, C.c_int, DP]
        _r = r
    return _r
