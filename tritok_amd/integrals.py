"""Integral assembly entry points of the C-ABI (FMATkmdcSteady, FMATsdkm)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import MeshView, Quad, check, i32

_macro = _lib._sig("tt_macro_integrals", i32, C.POINTER(MeshView), C.POINTER(Quad), i32, *([_lib.PD] * 12),
                   _lib.PD)
_sd = _lib._sig("tt_sd_integrals", i32, C.POINTER(MeshView), C.POINTER(Quad), i32, i32, *([_lib.PD] * 7), _lib.PD)

MACRO = ["H", "G", "B", "Abdx", "Abdy", "Abdz", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]
SDKM = ["H", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]
_keep = []


def quad(kmBr=4, kmBs=4, kmDr=4, kmDs=4, iiDiv=2, table=None):
    """tt_quad with optional GaussType table dict(gi, ome, ng1, ng2)."""
    q = Quad(kmBr, kmBs, kmDr, kmDs, iiDiv)
    if table is not None:
        t = [np.ascontiguousarray(table["gi"], np.float64), np.ascontiguousarray(table["ome"], np.float64),
             np.ascontiguousarray(table["ng1"], np.int32), np.ascontiguousarray(table["ng2"], np.int32)]
        _keep.append(t)
        q.gi, q.ome = t[0].ctypes.data_as(_lib.PD), t[1].ctypes.data_as(_lib.PD)
        q.ng1, q.ng2 = t[2].ctypes.data_as(_lib.PI), t[3].ctypes.data_as(_lib.PI)
    return q


def macro_integrals(view, q, copy=0):
    """view: CaseView (or its .mesh) of a Host.  Returns dict of column-major
    (rows, cols) arrays and ont[11]."""
    m = view.mesh if hasattr(view, "mesh") else view
    m1, m2 = m.nicell * 51, m.nicell * 27
    shp = {"G": (m1, 24), "Htx": (m2, 27), "Hty": (m2, 27), "Htz": (m2, 27)}
    out = {k: np.zeros(shp.get(k, (m1, 27)), order="F") for k in MACRO}
    ont = np.zeros(11)
    check(_macro(C.byref(m), C.byref(q), int(copy), *[out[k].ctypes.data_as(_lib.PD) for k in MACRO],
                 ont.ctypes.data_as(_lib.PD)), "tt_macro_integrals")
    out["ont"] = ont
    return out


def sd_integrals(view, q, row0=0, row1=None):
    m = view.mesh if hasattr(view, "mesh") else view
    row1 = m.nbnodes if row1 is None else row1
    r = row1 - row0
    out = {k: np.zeros((r, m.nbnodes if k.startswith("H") else m.nnodes), order="F") for k in SDKM}
    ont = np.zeros(10)
    check(_sd(C.byref(m), C.byref(q), int(row0), int(row1), *[out[k].ctypes.data_as(_lib.PD) for k in SDKM],
              ont.ctypes.data_as(_lib.PD)), "tt_sd_integrals")
    out["ont"] = ont
    return out
