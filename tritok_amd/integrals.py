"""Integral assembly entry points of the C-ABI (FMATkmdcSteady, FMATmodHelm,
FMATsdkm)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import MeshView, Quad, check, i32

_macro = _lib._sig("tt_macro_integrals", i32, C.POINTER(MeshView), C.POINTER(Quad), i32, *([_lib.PD] * 12),
                   _lib.PD)
_macro_mh = _lib._sig("tt_macro_integrals_mh", i32, C.POINTER(MeshView), C.POINTER(Quad), i32, _lib.PD, _lib.PD,
                      C.c_double, *([_lib.PD] * 12), _lib.PD, _lib.PD)
_stats = _lib._sig("tt_last_integral_stats", i32, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int64))
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


class ArrayMesh:
    """A tt_mesh_view over numpy arrays (1-based indices, the layouts of
    Host.mesh_arrays()); keeps the arrays alive with the view."""

    def __init__(self, a):
        self.arrays = {
            "x": np.asfortranarray(a["x"], np.float64), "ibc": np.asfortranarray(a["ibc"], np.int32),
            "idc": np.asfortranarray(a["idc"], np.int32), "ibf": np.asfortranarray(a["ibf"], np.int32),
            "xq": np.asfortranarray(a["xq"], np.float64), "flpm": np.asfortranarray(a["flpm"], np.int32),
            "gbn": np.ascontiguousarray(a["gbn"], np.int32), "lbn": np.ascontiguousarray(a["lbn"], np.int32)}
        r = self.arrays
        m = MeshView()
        m.nnodes, m.nbnodes, m.nbelem = len(r["x"]), len(r["gbn"]), len(r["ibc"])
        m.nicell, m.nq = len(r["idc"]), len(r["xq"])
        for k in ("x", "xq"):
            setattr(m, k, r[k].ctypes.data_as(_lib.PD))
        for k in ("ibc", "idc", "ibf", "flpm", "gbn", "lbn"):
            setattr(m, k, r[k].ctypes.data_as(_lib.PI))
        self.mesh = m


def macro_integrals_mh(view, q, diff, qdiff, beta, copy=0):
    """FMATmodHelm: diff (nnodes) and qdiff (nq) diffusivities, mu = sqrt(beta/diff)
    per source.  Returns the dict of macro_integrals plus 'mumax'."""
    m = view.mesh if hasattr(view, "mesh") else view
    diff = np.ascontiguousarray(diff, np.float64)
    qdiff = np.ascontiguousarray(qdiff, np.float64)
    if diff.shape != (m.nnodes,) or qdiff.shape != (m.nq,):
        raise ValueError("diff must have nnodes and qdiff nq entries")
    m1, m2 = m.nicell * 51, m.nicell * 27
    shp = {"G": (m1, 24), "Htx": (m2, 27), "Hty": (m2, 27), "Htz": (m2, 27)}
    out = {k: np.zeros(shp.get(k, (m1, 27)), order="F") for k in MACRO}
    ont = np.zeros(11)
    mumax = C.c_double()
    check(_macro_mh(C.byref(m), C.byref(q), int(copy), diff.ctypes.data_as(_lib.PD), qdiff.ctypes.data_as(_lib.PD),
                    float(beta), *[out[k].ctypes.data_as(_lib.PD) for k in MACRO], ont.ctypes.data_as(_lib.PD),
                    C.byref(mumax)), "tt_macro_integrals_mh")
    out["ont"] = ont
    out["mumax"] = mumax.value
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


def last_integral_stats():
    """Device time (ms) and kernel evaluations (domain, boundary) of this
    thread's last macro assembly."""
    ms, d, b = C.c_double(), C.c_int64(), C.c_int64()
    _stats(C.byref(ms), C.byref(d), C.byref(b))
    return dict(ms=ms.value, domain_points=d.value, boundary_points=b.value)
