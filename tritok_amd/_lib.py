"""ctypes binding of libtritok_amd.so (the C-ABI declared in include/tritok_bem.h).

The library is loaded from this package directory only; there is no fallback.
If it is missing the import fails with a message that says how to build it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtritok_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libtritok_amd.so not built: run `python tritok_amd/build.py` "
                      "(or __graft_entry__.build()) first")

lib = C.CDLL(LIB_PATH)

i32 = C.c_int32
i64 = C.c_int64
dbl = C.c_double
flt = C.c_float
P = C.c_void_p
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)


class MeshView(C.Structure):
    _fields_ = [("nnodes", i32), ("nbnodes", i32), ("nbelem", i32), ("nicell", i32), ("nq", i32),
                ("x", PD), ("ibc", PI), ("idc", PI), ("ibf", PI), ("xq", PD), ("flpm", PI),
                ("gbn", PI), ("lbn", PI), ("nx", PD), ("ny", PD), ("nz", PD)]


class Quad(C.Structure):
    _fields_ = [("kmBr", i32), ("kmBs", i32), ("kmDr", i32), ("kmDs", i32), ("iiDiv", i32),
                ("gi", PD), ("ome", PD), ("ng1", PI), ("ng2", PI)]


class CaseView(C.Structure):
    _fields_ = [("mesh", MeshView), ("quad", Quad), ("copy", i32),
                ("kode", PI), ("kobc", PI), ("wkode", PI), ("wqkode", PI), ("Tkode", PI), ("Tqkode", PI),
                ("nunk", i32 * 3), ("nbk", i32 * 3), ("Tnunk", i32), ("Tnb", i32),
                ("kmu", PI), ("nsmcol", i32), ("nrhscol", i32), ("iDT", i32),
                ("ReN", dbl), ("RaN", dbl), ("PrN", dbl), ("gx", dbl), ("gy", dbl), ("gz", dbl), ("tstep", dbl),
                ("urBw", dbl * 3), ("urDv", dbl * 3), ("urDw", dbl * 3), ("urDT", dbl), ("eps", dbl),
                ("nnlit", i32), ("lsqs_maxit", i32), ("iDLSE", i32),
                ("lsqs_eps", flt), ("dlse_r", flt), ("dlse_epsth", flt), ("iBw", i32)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


tt_last_error = _sig("tt_last_error", C.c_char_p)


class TTError(RuntimeError):
    pass


def check(rc, what=""):
    if rc is not None and rc < 0:
        raise TTError("%s failed (%d): %s" % (what, rc, tt_last_error().decode(errors="replace")))
    return rc


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


def iptr(a):
    assert a.dtype == np.int32
    return a.ctypes.data_as(PI)


def fortran_view(arr):
    """(rows, cols) column-major array as a flat Fortran-ordered copy."""
    return np.asfortranarray(arr)
