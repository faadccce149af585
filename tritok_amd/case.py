"""Whole-case driver of the C-ABI (tt_case_*): the reference's setup
(main.f90:237-1084) and steady Picard loop (main.f90:1200-1589) on the device.

    h = Host.cube(20).read_bic(bic).read_inp(inp)
    with Case(h, prec="fp32") as c:
        sta, ite = c.run()
        fields = c.fields()
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CaseView, P, check, i32, i64

_create = _lib._sig("tt_case_create", i32, C.POINTER(CaseView), _lib.PD, _lib.PD, _lib.PD, _lib.PD, _lib.PD, i32,
                    C.c_double, C.POINTER(P))
_step = _lib._sig("tt_case_picard_step", i32, P, _lib.PD, _lib.PI)
_step_batch = _lib._sig("tt_case_picard_step_batch", i32, C.POINTER(P), i32, _lib.PD, _lib.PI, _lib.PI)
_run = _lib._sig("tt_case_run", i32, P, i32, _lib.PD, _lib.PI)
_fields = _lib._sig("tt_case_fields", i32, P, _lib.PD, _lib.PD, _lib.PD, _lib.PD, _lib.PD)
_timings = _lib._sig("tt_case_timings", i32, P, _lib.PD)
_bytes = _lib._sig("tt_case_device_bytes", i64, P)
_free = _lib._sig("tt_case_free", None, P)
_reset = _lib._sig("tt_case_reset", i32, P, _lib.PD, _lib.PD, _lib.PD, _lib.PD, _lib.PD)
_setprec = _lib._sig("tt_case_set_precision", i32, P, i32)
_setanchor = _lib._sig("tt_case_set_anchor_period", i32, P, i32)
_probe = _lib._sig("tt_case_probe", i32, P, i32)
_probe_stats = _lib._sig("tt_case_probe_stats", i32, P, i32, C.POINTER(C.c_double), C.POINTER(i64), C.POINTER(i64))
_sysinfo = _lib._sig("tt_case_system_info", i32, P, i32, C.POINTER(i64))
_setgrid = _lib._sig("tt_case_set_grid", i32, P, i32)
_tuning = _lib._sig("tt_case_system_tuning", i32, P, i32, _lib.PI, _lib.PI)
_setthreads = _lib._sig("tt_case_set_unknown_threads", i32, P, i32)
_setgraphs = _lib._sig("tt_case_set_graphs", i32, P, i32)
_wrst = _lib._sig("tt_case_write_restart", i32, P, C.c_char_p, i32, i32, C.c_double)

PREC = {"fp64": 0, "fp32": 1, "bf16": 2, "fp16": 3}
TIMING_KEYS = ["boundary_vorticity", "km", "temperature", "vte", "convergence", "reserved", "step_total",
               "setup_macro", "setup_sdkm", "setup_lu", "setup_systems"]


def _fd(a):
    return np.asfortranarray(a, dtype=np.float64)


class Case:
    """Device-resident case.  `host` is a tritok_amd.Host with mesh, boundary
    conditions and tri.inp parameters; `quad` optionally replaces the
    quadrature description (a tt_quad, e.g. carrying the host's Gauss table).
    `prec`: "fp64" everywhere, or "fp32" / "fp16" / "bf16": FP32
    least-squares blocks anchored on FP64 residuals, and the dense kinematics
    products in that format anchored by an FP64 step every few iterations."""

    def __init__(self, host, prec="fp64", quad=None, fields=None, refine_tol=1e-12):
        self.host = host
        self.view = host.view()
        if quad is not None:
            self.view.quad = quad
            self._quad = quad
        f = fields if fields is not None else host.fields()
        self._init = [_fd(f[k]) for k in ("velocity", "vorticity", "qvorticity", "temp", "qtemp")]
        m = self.view.mesh
        self.nn, self.nq = m.nnodes, m.nq
        h = P()
        check(_create(C.byref(self.view), *[a.ctypes.data_as(_lib.PD) for a in self._init], PREC[prec],
                      float(refine_tol), C.byref(h)), "tt_case_create")
        self._h = h

    def step(self):
        err = np.zeros(9)
        nit = np.zeros(12, np.int32)
        rc = check(_step(self._h, err.ctypes.data_as(_lib.PD), nit.ctypes.data_as(_lib.PI)), "tt_case_picard_step")
        return rc, err, nit

    def run(self, max_iter=100000):
        cap = max(1, min(int(max_iter), int(self.view.nnlit)))
        sta = np.zeros((cap, 9))
        ite = np.zeros((cap, 12), np.int32)
        n = check(_run(self._h, cap, sta.ctypes.data_as(_lib.PD), ite.ctypes.data_as(_lib.PI)), "tt_case_run")
        return sta[:n], ite[:n]

    def fields(self):
        nn, nq = self.nn, self.nq
        out = dict(velocity=np.zeros((nn, 3), order="F"), vorticity=np.zeros((nn, 3), order="F"),
                   qvorticity=np.zeros((nq, 3), order="F"), temp=np.zeros(nn), qtemp=np.zeros(nq))
        check(_fields(self._h, *[out[k].ctypes.data_as(_lib.PD)
                                 for k in ("velocity", "vorticity", "qvorticity", "temp", "qtemp")]), "tt_case_fields")
        return out

    def timings(self):
        t = np.zeros(11)
        check(_timings(self._h, t.ctypes.data_as(_lib.PD)), "tt_case_timings")
        return dict(zip(TIMING_KEYS, t))

    def reset(self, fields=None):
        """Restarts from the initial fields (or the given ones) on the resident matrices."""
        if fields is not None:
            self._init = [_fd(fields[k]) for k in ("velocity", "vorticity", "qvorticity", "temp", "qtemp")]
        check(_reset(self._h, *[a.ctypes.data_as(_lib.PD) for a in self._init]), "tt_case_reset")

    def set_precision(self, prec):
        check(_setprec(self._h, PREC[prec]), "tt_case_set_precision")

    def set_anchor_period(self, period):
        check(_setanchor(self._h, int(period)), "tt_case_set_anchor_period")

    def probe(self, stride):
        check(_probe(self._h, int(stride)), "tt_case_probe")

    def probe_stats(self, rows=51):
        avg, sampled, seen = C.c_double(), i64(), i64()
        check(_probe_stats(self._h, int(rows), C.byref(avg), C.byref(sampled), C.byref(seen)), "tt_case_probe_stats")
        return dict(avg_ms=avg.value, sampled=sampled.value, launches=seen.value)

    def set_grid(self, grid):
        """Cell-kernel grid for every least-squares system (0: autotuned)."""
        check(_setgrid(self._h, int(grid)), "tt_case_set_grid")

    def set_graphs(self, on):
        """Least-squares iterations replayed from captured HIP graphs (True)
        or launched kernel by kernel on the stream (False, the default)."""
        check(_setgraphs(self._h, 1 if on else 0), "tt_case_set_graphs")

    def system_info(self, sys):
        """sys: 0..2 vorticity x, y, z; 3 energy; 4 kinematics."""
        info = (i64 * 9)()
        check(_sysinfo(self._h, int(sys), info), "tt_case_system_info")
        return dict(zip(["m", "n", "nnz", "cells", "rows_per_cell", "cols_per_cell", "grid_fp64", "grid_fp32",
                         "grid_fp16"], list(info)))

    def write_restart(self, path, tstep=1, glnnlit=0, rtime=1e20):
        """tri.rst (WrstF, version 4) of the current state."""
        check(_wrst(self._h, str(path).encode(), int(tstep), int(glnnlit), float(rtime)), "tt_case_write_restart")

    def set_unknown_threads(self, threads):
        """Threads per workgroup of the unknown kernel (256/512/1024; 0: tuned)."""
        check(_setthreads(self._h, int(threads)), "tt_case_set_unknown_threads")

    def system_tuning(self, sys):
        """Autotuned launch shapes of a system: {'grid': [fp64, fp32, fp16], 'threads': [...]}."""
        g, t = np.zeros(3, np.int32), np.zeros(3, np.int32)
        check(_tuning(self._h, int(sys), g.ctypes.data_as(_lib.PI), t.ctypes.data_as(_lib.PI)),
              "tt_case_system_tuning")
        return {"grid": g.tolist(), "threads": t.tolist()}

    def device_bytes(self):
        return int(_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_batch(cases):
    """One Picard iteration of each Case in `cases` (1..16, same block
    precision), their LSQR iterations sharing launches
    (tt_case_picard_step_batch).  Returns (status, err, nit): status[i] 1 =
    continue, 0 = finished, -1 = failed; err (k, 9); nit (k, 12)."""
    k = len(cases)
    hs = (P * k)(*[c._h for c in cases])
    err = np.zeros((k, 9))
    nit = np.zeros((k, 12), np.int32)
    st = np.zeros(k, np.int32)
    check(_step_batch(hs, k, err.ctypes.data_as(_lib.PD), nit.ctypes.data_as(_lib.PI), st.ctypes.data_as(_lib.PI)),
          "tt_case_picard_step_batch")
    return st, err, nit
