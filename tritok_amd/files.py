"""The reference's files through the C-ABI (host only): the integral caches
tri.macro.int / tri.sdkm.int.N, the restart file tri.rst (version 4) and the
iteration history tri.sta / tri.ite.  Arrays are column-major, as the
reference holds them."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import PD, PI, check, i32

_PPD = C.POINTER(PD)
_wmacro = _lib._sig("tt_write_macro_int", i32, C.c_char_p, i32, i32, i32, *([PD] * 13))
_rmacro = _lib._sig("tt_read_macro_int", i32, C.c_char_p, i32, PI, PI, *([PD] * 13))
_wsdkm = _lib._sig("tt_write_sdkm_int", i32, C.c_char_p, i32, i32, i32, i32, *([PD] * 8))
_rsdkm = _lib._sig("tt_read_sdkm_int", i32, C.c_char_p, i32, PI, *([PD] * 8))
_wrst = _lib._sig("tt_write_restart", i32, C.c_char_p, i32, i32, i32, i32, C.c_double, _PPD)
_rrst = _lib._sig("tt_read_restart", i32, C.c_char_p, i32, i32, PI, PI, PI, PD, _PPD)
_wsta = _lib._sig("tt_write_sta_ite", i32, C.c_char_p, C.c_char_p, C.c_char_p, i32, PI, PI, PD, PI)

MACRO = ("H", "G", "B", "Abdx", "Abdy", "Abdz", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz")
SDKM = ("H", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz")
RESTART = ("velocity", "pnlsvelocity", "vorticity", "ptsvorticity", "pptsvorticity", "pnlsvorticity", "qvorticity",
           "temp", "ptstemp", "pptstemp", "pnlstemp", "qtemp", "dc", "ptsdc", "pptsdc", "pnlsdc", "qdc")
_RST_SHAPE = {"velocity": ("n", 3), "pnlsvelocity": ("n", 3), "vorticity": ("n", 3), "ptsvorticity": ("n", 3),
              "pptsvorticity": ("n", 3), "pnlsvorticity": ("n", 3), "qvorticity": ("q", 3)}


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def _b(path):
    return str(path).encode()


def write_macro_int(path, m, nnodes, nbnodes):
    """WriteMacroIntDisk: m = dict of the twelve matrices (keys MACRO) and 'ont' (11)."""
    arrs = [_f(m[k]) for k in MACRO] + [np.ascontiguousarray(m["ont"], np.float64)]
    nicell = arrs[0].shape[0] // 51
    check(_wmacro(_b(path), nicell, int(nnodes), int(nbnodes), *[a.ctypes.data_as(PD) for a in arrs]),
          "tt_write_macro_int")


def read_macro_int(path, nicell):
    m1, m2 = nicell * 51, nicell * 27
    out = {k: np.zeros((m2 if k.startswith("Ht") else m1, 24 if k == "G" else 27), order="F") for k in MACRO}
    out["ont"] = np.zeros(11)
    nn, nb = i32(), i32()
    check(_rmacro(_b(path), int(nicell), C.byref(nn), C.byref(nb), *[out[k].ctypes.data_as(PD) for k in MACRO],
                  out["ont"].ctypes.data_as(PD)), "tt_read_macro_int")
    out["nnodes"], out["nbnodes"] = nn.value, nb.value
    return out


def write_sdkm_int(path, sd, nnodes, nbnodes, nproc=1):
    """WriteSDIntDisk: sd = dict of the seven matrices (rows x nbnodes / nnodes) and 'ont' (10)."""
    arrs = [_f(sd[k]) for k in SDKM] + [np.ascontiguousarray(sd["ont"], np.float64)]
    rows = arrs[0].shape[0]
    check(_wsdkm(_b(path), int(nnodes), int(nbnodes), int(nproc), rows, *[a.ctypes.data_as(PD) for a in arrs]),
          "tt_write_sdkm_int")


def read_sdkm_int(path, rows, nnodes, nbnodes):
    out = {k: np.zeros((rows, nbnodes if k.startswith("H") else nnodes), order="F") for k in SDKM}
    out["ont"] = np.zeros(10)
    hdr = np.zeros(3, np.int32)
    check(_rsdkm(_b(path), int(rows), hdr.ctypes.data_as(PI), *[out[k].ctypes.data_as(PD) for k in SDKM],
                 out["ont"].ctypes.data_as(PD)), "tt_read_sdkm_int")
    out["header"] = hdr
    return out


def _rst_arrays(fields, nn, nq, alloc=False):
    arrs = []
    for k in RESTART:
        rows = nq if (k.startswith("q") or k == "qdc") else nn
        cols = 3 if k in _RST_SHAPE else 1
        a = np.zeros((rows, cols), order="F") if alloc else _f(np.asarray(fields[k]).reshape((rows, cols), order="F"))
        arrs.append(a)
    return arrs


def write_restart(path, fields, nnodes, nq, tstep=1, glnnlit=0, rtime=0.0):
    """WrstF (version 4): fields = dict keyed by RESTART."""
    arrs = _rst_arrays(fields, nnodes, nq)
    ptrs = (PD * 17)(*[a.ctypes.data_as(PD) for a in arrs])
    check(_wrst(_b(path), int(nnodes), int(nq), int(tstep), int(glnnlit), float(rtime), ptrs), "tt_write_restart")


def read_restart(path, nnodes, nq):
    """RrstF: returns (fields, version, tstep, glnnlit, rtime)."""
    arrs = _rst_arrays(None, nnodes, nq, alloc=True)
    ptrs = (PD * 17)(*[a.ctypes.data_as(PD) for a in arrs])
    ver, ts, gl, rt = i32(), i32(), i32(), C.c_double()
    check(_rrst(_b(path), int(nnodes), int(nq), C.byref(ver), C.byref(ts), C.byref(gl), C.byref(rt), ptrs),
          "tt_read_restart")
    return dict(zip(RESTART, arrs)), ver.value, ts.value, gl.value, rt.value


def write_sta_ite(sta_path, ite_path, err, nit, tstep=None, tsnnlit=None, id_line=None):
    """tri.sta / tri.ite from per-iteration changes err (n x 9) and solver
    counts nit (n x 12); tstep defaults to 1, tsnnlit to 1..n."""
    err = np.ascontiguousarray(err, np.float64).reshape(-1, 9)
    n = err.shape[0]
    nit = np.ascontiguousarray(nit, np.int32).reshape(n, 12)
    ts = np.ascontiguousarray(np.ones(n) if tstep is None else tstep, np.int32)
    it = np.ascontiguousarray(np.arange(1, n + 1) if tsnnlit is None else tsnnlit, np.int32)
    check(_wsta(None if sta_path is None else _b(sta_path), None if ite_path is None else _b(ite_path),
                None if id_line is None else id_line.encode(), n, ts.ctypes.data_as(PI), it.ctypes.data_as(PI),
                err.ctypes.data_as(PD), nit.ctypes.data_as(PI)), "tt_write_sta_ite")
