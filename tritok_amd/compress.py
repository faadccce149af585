"""Compressed matrices of the C-ABI: cross approximation (lib_ACA.f),
stretched-Haar wavelets (lib_WT.f) and block matrices of compressed parts
(lib_ctree.f).  Matrices are numpy arrays (column-major copies are made);
index arrays are the reference's, 1-based."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P, PD, PI, check, i32

i64 = C.c_int64
_aca_c = _lib._sig("tt_aca_compress", i32, i32, i32, PD, i32, C.c_double, PD, C.POINTER(P), C.POINTER(i32),
                   C.POINTER(C.c_double))
_aca_f = _lib._sig("tt_aca_factors", i32, P, PD, PD)
_aca_mv = _lib._sig("tt_aca_mv", i32, P, PD, PD)
_aca_abxv = _lib._sig("tt_aca_abxv", i32, i32, i32, i32, PD, PD, PD, PD)
_aca_free = _lib._sig("tt_aca_free", None, P)
_wt_c = _lib._sig("tt_wt_compress", i32, i32, i32, PD, C.c_double, C.POINTER(P), C.POINTER(i32))
_wt_csr = _lib._sig("tt_wt_csr", i32, P, PI, PI, PD)
_wt_mv = _lib._sig("tt_wt_mv", i32, P, PD, PD)
_wt_wtawx = _lib._sig("tt_wt_wtawx", i32, i32, i32, i32, PI, PI, PD, PD, PD)
_wt_free = _lib._sig("tt_wt_free", None, P)
_hm_c = _lib._sig("tt_hmat_create", i32, i32, i32, PD, i64, PD, PD, i32, C.c_double, i32, i32, C.c_double,
                  C.c_double, C.POINTER(P))
_hm_cp = _lib._sig("tt_hmat_create_parts", i32, i32, i32, PD, i64, i32, PI, PI, PI, PI, PI, i32, i32, C.c_double,
                   C.c_double, C.POINTER(P))
_hm_part = _lib._sig("tt_hmat_partition", i32, i32, PD, i32, PD, i32, C.c_double, C.POINTER(i32), C.POINTER(i64),
                     C.POINTER(i64), PI, PI, PI, PI, PI)
_hm_info = _lib._sig("tt_hmat_info", i32, P, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64),
                     C.POINTER(i32))
_hm_parts = _lib._sig("tt_hmat_parts", i32, P, PI, PI, PI, PI, PI, PI)
_hm_data = _lib._sig("tt_hmat_part_data", i32, P, i32, PD, PI)
_hm_mv = _lib._sig("tt_hmat_mv", i32, P, i32, PD, PD)
_hm_free = _lib._sig("tt_hmat_free", None, P)


def _f64(a, order="C"):
    return np.require(a, np.float64, [order[0] + "_CONTIGUOUS", "ALIGNED"])


class _Handle:
    _free = None

    def close(self):
        if getattr(self, "_h", None):
            type(self)._free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Aca(_Handle):
    """AcaDo: cross approximation of `a` (type 1, 2 or 3; vec required for 3)."""
    _free = _aca_free

    def __init__(self, a, type_=2, eps=1e-6, vec=None):
        a = _f64(a, "F")
        self.m, self.n = a.shape
        h, r, nrm = P(), i32(), C.c_double()
        v = None if vec is None else _f64(vec)
        check(_aca_c(self.m, self.n, a.ctypes.data_as(PD), int(type_), float(eps),
                     None if v is None else v.ctypes.data_as(PD), C.byref(h), C.byref(r), C.byref(nrm)),
              "tt_aca_compress")
        self._h, self.rank, self.norm = h, r.value, nrm.value

    def factors(self):
        a = np.zeros((self.m, self.rank), order="F")
        b = np.zeros((self.rank, self.n), order="F")
        check(_aca_f(self._h, a.ctypes.data_as(PD), b.ctypes.data_as(PD)), "tt_aca_factors")
        return a, b

    def mv(self, x):
        x = _f64(x)
        y = np.zeros(self.m)
        check(_aca_mv(self._h, x.ctypes.data_as(PD), y.ctypes.data_as(PD)), "tt_aca_mv")
        return y


def aca_abxv(a, b, x):
    a, b, x = _f64(a, "F"), _f64(b, "F"), _f64(x)
    y = np.zeros(a.shape[0])
    check(_aca_abxv(a.shape[0], b.shape[1], a.shape[1], a.ctypes.data_as(PD), b.ctypes.data_as(PD),
                    x.ctypes.data_as(PD), y.ctypes.data_as(PD)), "tt_aca_abxv")
    return y


class Wavelet(_Handle):
    """Che_WcomMatrix: W A W^T with entries below kappa * mean|.| dropped."""
    _free = _wt_free

    def __init__(self, a, kappa=1e-5):
        a = _f64(a, "F")
        self.m, self.n = a.shape
        h, nnz = P(), i32()
        check(_wt_c(self.m, self.n, a.ctypes.data_as(PD), float(kappa), C.byref(h), C.byref(nnz)), "tt_wt_compress")
        self._h, self.nnz = h, nnz.value

    def csr(self):
        zvr = np.zeros(self.m + 1, np.int32)
        c = np.zeros(max(self.nnz, 1), np.int32)
        v = np.zeros(max(self.nnz, 1))
        ref = check(_wt_csr(self._h, zvr.ctypes.data_as(PI), c.ctypes.data_as(PI), v.ctypes.data_as(PD)),
                    "tt_wt_csr")
        return dict(zvr=zvr, c=c[:self.nnz], v=v[:self.nnz], nnz_ref=ref)

    def mv(self, x):
        x = _f64(x)
        y = np.zeros(self.m)
        check(_wt_mv(self._h, x.ctypes.data_as(PD), y.ctypes.data_as(PD)), "tt_wt_mv")
        return y


def wt_wtawx(zvr, c, v, x, nrow):
    zvr, c = np.ascontiguousarray(zvr, np.int32), np.ascontiguousarray(c, np.int32)
    v, x = _f64(v), _f64(x)
    y = np.zeros(nrow)
    check(_wt_wtawx(nrow, len(x), len(c), zvr.ctypes.data_as(PI), c.ctypes.data_as(PI), v.ctypes.data_as(PD),
                    x.ctypes.data_as(PD), y.ctypes.data_as(PD)), "tt_wt_wtawx")
    return y


def partition(xr, xc, leaf=64, eta=1.0):
    """The library's block partition (host only): (rptr, rnodes, cptr, cnodes, admiss)."""
    xr, xc = _f64(xr, "F"), _f64(xc, "F")
    args = (len(xr), xr.ctypes.data_as(PD), len(xc), xc.ctypes.data_as(PD), int(leaf), float(eta))
    n, rl, cl = i32(), i64(), i64()
    check(_hm_part(*args, C.byref(n), C.byref(rl), C.byref(cl), None, None, None, None, None), "tt_hmat_partition")
    rptr, cptr = np.zeros(n.value + 1, np.int32), np.zeros(n.value + 1, np.int32)
    rn, cn, adm = np.zeros(rl.value, np.int32), np.zeros(cl.value, np.int32), np.zeros(n.value, np.int32)
    check(_hm_part(*args, None, None, None, *[v.ctypes.data_as(PI) for v in (rptr, rn, cptr, cn, adm)]),
          "tt_hmat_partition")
    return rptr, rn, cptr, cn, adm


class HMatrix(_Handle):
    """Block matrix of dense / wavelet (ctype 2) / cross-approximated (ctype 3)
    parts.  Either the library partitions by geometry (points xr, xc, leaf,
    eta) or `parts` = (rptr, rnodes, cptr, cnodes, admiss) is given."""
    _free = _hm_free

    def __init__(self, a, xr=None, xc=None, leaf=64, eta=1.0, ctype=3, aca_type=2, aca_eps=1e-6, wt_kappa=1e-5,
                 parts=None):
        a = _f64(a, "F")
        self.nrow, self.ncol = a.shape
        h = P()
        if parts is None:
            xr, xc = _f64(xr, "F"), _f64(xc, "F")
            check(_hm_c(self.nrow, self.ncol, a.ctypes.data_as(PD), self.nrow, xr.ctypes.data_as(PD),
                        xc.ctypes.data_as(PD), int(leaf), float(eta), int(ctype), int(aca_type), float(aca_eps),
                        float(wt_kappa), C.byref(h)), "tt_hmat_create")
        else:
            rptr, rn, cptr, cn, adm = [np.ascontiguousarray(v, np.int32) for v in parts]
            check(_hm_cp(self.nrow, self.ncol, a.ctypes.data_as(PD), self.nrow, len(adm), rptr.ctypes.data_as(PI),
                         rn.ctypes.data_as(PI), cptr.ctypes.data_as(PI), cn.ctypes.data_as(PI),
                         adm.ctypes.data_as(PI), int(ctype), int(aca_type), float(aca_eps), float(wt_kappa),
                         C.byref(h)), "tt_hmat_create_parts")
        self._adopt(h)

    def _adopt(self, h):
        self._h = h
        np_, rl, cl, st, mr = i32(), i64(), i64(), i64(), i32()
        check(_hm_info(h, C.byref(np_), C.byref(rl), C.byref(cl), C.byref(st), C.byref(mr)), "tt_hmat_info")
        self.nparts, self.stored, self.max_rank = np_.value, st.value, mr.value
        self._rl, self._cl = rl.value, cl.value

    @classmethod
    def from_handle(cls, h, nrow, ncol):
        """Wraps a tt_hmat made by another entry point (tt_sd_hmat)."""
        obj = cls.__new__(cls)
        obj.nrow, obj.ncol = int(nrow), int(ncol)
        obj._adopt(h)
        return obj

    def parts(self):
        n = self.nparts
        rptr, cptr = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
        rn, cn = np.zeros(self._rl, np.int32), np.zeros(self._cl, np.int32)
        ct, rk = np.zeros(n, np.int32), np.zeros(n, np.int32)
        check(_hm_parts(self._h, *[v.ctypes.data_as(PI) for v in (rptr, rn, cptr, cn, ct, rk)]), "tt_hmat_parts")
        return dict(rptr=rptr, rnodes=rn, cptr=cptr, cnodes=cn, ctype=ct, rank=rk)

    def part_data(self, p):
        """(doubles, ints) of part p in the layout of tt_hmat_part_data."""
        nd = check(_hm_data(self._h, int(p), None, None), "tt_hmat_part_data")
        info = self.parts()
        nr = info["rptr"][p + 1] - info["rptr"][p]
        d = np.zeros(max(nd, 1))
        iv = np.zeros(nr + 1 + max(nd, 0), np.int32)
        check(_hm_data(self._h, int(p), d.ctypes.data_as(PD), iv.ctypes.data_as(PI)), "tt_hmat_part_data")
        return d[:nd], iv

    def mv(self, x):
        x = np.asarray(x, np.float64)
        one = x.ndim == 1
        X = _f64(x.reshape(self.ncol, -1), "F")
        Y = np.zeros((self.nrow, X.shape[1]), order="F")
        check(_hm_mv(self._h, X.shape[1], X.ctypes.data_as(PD), Y.ctypes.data_as(PD)), "tt_hmat_mv")
        return Y[:, 0] if one else Y
