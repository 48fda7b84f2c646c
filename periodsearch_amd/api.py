"""Python front end of the C ABI: the period scan as the reference's host driver sees it.

    ps = PeriodSearch.from_file("period_search_in")     # period_search_BOINC.cpp:277-485 (parse)
    ps.prepare(device=0)                                 # CUDAPrepare
    conw_r, new_conw = ps.conw_search()                  # CUDAPrecalc (convexity-weight search)
    lines = ps.scan(1, 100, conw_r)                      # CUDAStart: period_search_out lines
    text = "".join(format_line(l) for l in lines)

All computation happens in libpsamd.so (HIP kernels for gfx950 plus host C++); this module only
marshals numpy arrays.
"""
import ctypes as C

import numpy as np

from . import _lib as L

LINE_DTYPE = np.dtype([(k, "f8") for k in ("per", "dev", "chi2", "dark", "la", "be")])
FIT_DTYPE = np.dtype([(k, "f8") for k in ("dev", "prd", "dark", "totarea", "la", "be", "chisq")] +
                     [("niter", "i4"), ("nsolve_fail", "i4")])


def _dp(a):
    return a.ctypes.data_as(L.DP) if a is not None else None


def format_line(line):
    """One period_search_out line, format of period_search_BOINC.cpp:927-931."""
    ln = L.PsLine(*[float(line[k]) for k in LINE_DTYPE.names])
    buf = C.create_string_buffer(256)
    L.check(L.lib().ps_format_line(C.byref(ln), buf, 256), "ps_format_line")
    return buf.value.decode()


class PeriodSearch:
    """Parsed input + setup tables + (after prepare) a device handle."""

    def __init__(self, inp_ptr):
        self._lib = L.lib()
        self._inp = inp_ptr
        self._geo = C.POINTER(L.PsGeometry)()
        L.check(self._lib.ps_geometry_build(self._inp, C.byref(self._geo)), "ps_geometry_build")
        self._h = None

    @classmethod
    def from_file(cls, path):
        p = C.POINTER(L.PsInput)()
        L.check(L.lib().ps_input_read(str(path).encode(), C.byref(p)), "ps_input_read")
        return cls(p)

    @classmethod
    def from_text(cls, text):
        p = C.POINTER(L.PsInput)()
        L.check(L.lib().ps_input_parse(text.encode(), C.byref(p)), "ps_input_parse")
        return cls(p)

    # ---- parsed input and setup tables (host) ----
    @property
    def input(self):
        return self._inp.contents

    def input_arrays(self):
        i = self._inp.contents
        n = i.ndata
        return {
            "tim": np.ctypeslib.as_array(i.tim, (n,)).copy(),
            "brightness": np.ctypeslib.as_array(i.brightness, (n,)).copy(),
            "sig": np.ctypeslib.as_array(i.sig, (n,)).copy(),
            "weight": np.ctypeslib.as_array(i.weight, (n,)).copy(),
            "ee": np.ctypeslib.as_array(i.ee, (n * 3,)).reshape(n, 3).copy(),
            "ee0": np.ctypeslib.as_array(i.ee0, (n * 3,)).reshape(n, 3).copy(),
            "lpoints": np.ctypeslib.as_array(i.lpoints, (i.lcurves,)).copy(),
            "inrel": np.ctypeslib.as_array(i.inrel, (i.lcurves,)).copy(),
        }

    def geometry(self):
        g = self._geo.contents
        nf, nc = g.nfac, g.ncoef
        return {
            "nor": np.ctypeslib.as_array(g.nor, (nf * 3,)).reshape(nf, 3).copy(),
            "darea": np.ctypeslib.as_array(g.darea, (nf,)).copy(),
            "dsph": np.ctypeslib.as_array(g.dsph, (nf * nc,)).reshape(nf, nc).copy(),
            "cg_first": np.ctypeslib.as_array(g.cg_first, (nc,)).copy(),
        }

    # ---- device ----
    def prepare(self, device=0, precision=L.PS_FP64):
        h = C.c_void_p()
        L.check(self._lib.ps_prepare(self._inp, self._geo, device, C.byref(h)), "ps_prepare")
        self._h = h
        self.set_precision(precision)
        return self

    def _need(self):
        if self._h is None:
            raise RuntimeError("call prepare() first")
        return self._h

    def set_precision(self, precision):
        L.check(self._lib.ps_set_precision(self._need(), precision), "ps_set_precision")

    def set_kernel(self, kernel):
        """Kernel form of single-problem FP64 fits: PS_KERNEL_AUTO, _LEGACY or _JFREE."""
        L.check(self._lib.ps_set_kernel(self._need(), kernel), "ps_set_kernel")

    def kernel(self):
        return L.check(self._lib.ps_get_kernel(self._need()), "ps_get_kernel")

    def autotune(self, conw_r):
        """Times both kernel forms on a sample of this problem's fits and keeps the faster: returns
        (kernel form, ms per fit of the legacy form, ms per fit of the J-free form)."""
        k = C.c_int()
        ms = (C.c_double * 2)()
        L.check(self._lib.ps_autotune(self._need(), float(conw_r), C.byref(k), ms), "ps_autotune")
        return k.value, ms[0], ms[1]

    def info(self):
        inf = L.PsInfo()
        L.check(self._lib.ps_get_info(self._need(), C.byref(inf)), "ps_get_info")
        return {k: getattr(inf, k) for k, _ in L.PsInfo._fields_}

    def freq(self, n):
        """Trial frequency [1/h] of 1-based period index n (period_search_BOINC.cpp:812)."""
        inf = self.info()
        return inf["freq_start"] - (np.asarray(n) - 1) * inf["freq_step"]

    def fit_batch(self, freq, pole, conw_r):
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        pole = np.ascontiguousarray(pole, dtype=np.int32)
        out = np.zeros(len(freq), dtype=FIT_DTYPE)
        L.check(self._lib.ps_fit_batch(self._need(), len(freq), _dp(freq), pole.ctypes.data_as(L.IP),
                                       float(conw_r), out.ctypes.data_as(C.POINTER(L.PsFitResult))),
                "ps_fit_batch")
        return out

    def _handle(self):
        return self._need()

    def conw_round(self, conw_r):
        ave = C.c_double()
        L.check(self._lib.ps_conw_round(self._need(), float(conw_r), C.byref(ave)), "ps_conw_round")
        return ave.value

    def conw_search(self):
        cw = C.c_double()
        nc = C.c_int()
        L.check(self._lib.ps_conw_search(self._need(), C.byref(cw), C.byref(nc)), "ps_conw_search")
        return cw.value, nc.value

    def scan(self, n_from, n_to, conw_r):
        nmax = self.info()["nmax"]
        if n_to <= 0 or n_to > nmax:
            n_to = nmax
        out = np.zeros(n_to - n_from + 1, dtype=LINE_DTYPE)
        L.check(self._lib.ps_scan(self._need(), n_from, n_to, float(conw_r),
                                  out.ctypes.data_as(C.POINTER(L.PsLine))), "ps_scan")
        return out

    def eval(self, a, conw_r, raw=False):
        """Normal equations (alpha, beta, chi2) for parameter vectors a[nvec, ma] (0-based cg)."""
        a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64)
        inf = self.info()
        nv, mf, npts, ma = a.shape[0], inf["mfit"], inf["npts"], inf["ma"]
        if a.shape[1] != ma:
            raise ValueError(f"parameter vectors must have {ma} entries")
        alpha = np.zeros((nv, mf, mf))
        beta = np.zeros((nv, mf))
        chisq = np.zeros(nv)
        ymod = np.zeros((nv, npts)) if raw else None
        dyda = np.zeros((nv, npts, ma)) if raw else None
        L.check(self._lib.ps_eval(self._need(), nv, _dp(a), float(conw_r), _dp(alpha), _dp(beta), _dp(chisq),
                                  _dp(ymod), _dp(dyda)), "ps_eval")
        res = {"alpha": alpha, "beta": beta, "chisq": chisq}
        if raw:
            res["ymod"] = ymod
            res["dyda"] = dyda
        return res

    def kernel_time(self, reset=False):
        ms = C.c_double()
        n = C.c_int()
        L.check(self._lib.ps_kernel_time(self._need(), C.byref(ms), C.byref(n), int(reset)), "ps_kernel_time")
        return ms.value, n.value

    def close(self):
        if self._h is not None:
            self._lib.ps_release(self._h)
            self._h = None
        if self._geo:
            self._lib.ps_geometry_free(self._geo)
            self._geo = C.POINTER(L.PsGeometry)()
        if self._inp:
            self._lib.ps_input_free(self._inp)
            self._inp = C.POINTER(L.PsInput)()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_files(in_path, out_path, state_path=None, device=0, precision=L.PS_FP64):
    """The reference process contract: period_search_in -> period_search_out (+ state file)."""
    L.check(L.lib().ps_run_files(str(in_path).encode(), str(out_path).encode(),
                                 str(state_path).encode() if state_path else None, device, precision),
            "ps_run_files")


def fit_multi(searches, conw_r, prob, freq, pole):
    """Fits of several prepared problems as one batch (ps_fit_multi): job j fits freq[j] with pole
    pole[j] of searches[prob[j]] at convexity weight conw_r[prob[j]].  Returns FIT_DTYPE records in job
    order."""
    searches = list(searches)
    if not searches:
        raise ValueError("fit_multi: no problems")
    lib = searches[0]._lib
    ptrs = [s._handle() for s in searches]
    hs = (C.c_void_p * len(searches))(*[p.value if isinstance(p, C.c_void_p) else p for p in ptrs])
    cw = np.ascontiguousarray(conw_r, dtype=np.float64)
    prob = np.ascontiguousarray(prob, dtype=np.int32)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    pole = np.ascontiguousarray(pole, dtype=np.int32)
    if not (len(cw) == len(searches) and len(prob) == len(freq) == len(pole)):
        raise ValueError("fit_multi: inconsistent sizes")
    out = np.zeros(len(freq), dtype=FIT_DTYPE)
    L.check(lib.ps_fit_multi(hs, len(searches), _dp(cw), len(freq), prob.ctypes.data_as(L.IP), _dp(freq),
                             pole.ctypes.data_as(L.IP), out.ctypes.data_as(C.POINTER(L.PsFitResult))), "ps_fit_multi")
    return out


class Batch:
    """Many prepared problems (asteroids) fitted as one persistent launch over a device work queue
    (ps_batch_create / ps_batch_run): job j fits freq[j] with pole pole[j] of searches[prob[j]] at
    convexity weight conw_r[prob[j]].  The searches must stay open while the batch is used."""

    def __init__(self, searches):
        self._searches = list(searches)
        if not self._searches:
            raise ValueError("Batch: no problems")
        self._lib = self._searches[0]._lib
        ptrs = [s._handle() for s in self._searches]
        hs = (C.c_void_p * len(ptrs))(*[p.value if isinstance(p, C.c_void_p) else p for p in ptrs])
        b = C.c_void_p()
        L.check(self._lib.ps_batch_create(hs, len(ptrs), C.byref(b)), "ps_batch_create")
        self._b = b

    def run(self, conw_r, prob, freq, pole):
        cw = np.ascontiguousarray(conw_r, dtype=np.float64)
        prob = np.ascontiguousarray(prob, dtype=np.int32)
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        pole = np.ascontiguousarray(pole, dtype=np.int32)
        if not (len(cw) == len(self._searches) and len(prob) == len(freq) == len(pole)):
            raise ValueError("Batch.run: inconsistent sizes")
        out = np.zeros(len(freq), dtype=FIT_DTYPE)
        L.check(self._lib.ps_batch_run(self._b, _dp(cw), len(freq), prob.ctypes.data_as(L.IP), _dp(freq),
                                       pole.ctypes.data_as(L.IP), out.ctypes.data_as(C.POINTER(L.PsFitResult))),
                "ps_batch_run")
        return out

    def kernel_time(self, reset=False):
        ms = C.c_double()
        n = C.c_int()
        L.check(self._lib.ps_batch_kernel_time(self._b, C.byref(ms), C.byref(n), int(reset)), "ps_batch_kernel_time")
        return ms.value, n.value

    def close(self):
        if self._b is not None:
            self._lib.ps_batch_release(self._b)
            self._b = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
