"""Request loop over whole flow cases (tt_server_*): cases are queued as
host-side inputs (Host objects with boundary conditions and tri.inp read),
admitted by device memory into a fixed number of concurrent slots, iterated to
convergence on the device and returned with their timings."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P, PD, check, i32

i64 = C.c_int64


class ServerInfo(C.Structure):
    _fields_ = [("status", i32), ("picard_iterations", i32), ("lsqr_iterations", i64), ("final_change", C.c_double),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_setup", C.c_double), ("t_end", C.c_double),
                ("estimated_bytes", i64), ("device_bytes", i64), ("error", C.c_char * 256), ("batch_peak", i32)]


_create = _lib._sig("tt_server_create_batched", i32, i32, i32, i64, i32, i32, C.POINTER(P))
_submit = _lib._sig("tt_server_submit", i32, P, P, i32, C.POINTER(i64))
_wait = _lib._sig("tt_server_wait", i32, P, i64, C.POINTER(ServerInfo))
_fields = _lib._sig("tt_server_fields", i32, P, i64, PD, PD, PD, PD, PD)
_release = _lib._sig("tt_server_release", i32, P, i64)
_load = _lib._sig("tt_server_load", i32, P, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64))
_free = _lib._sig("tt_server_free", None, P)

PREC = {"fp64": 0, "fp32": 1, "bf16": 2, "fp16": 3}


class Server:
    def __init__(self, slots=4, mem_budget=0, anchor_period=16, max_iter=0, max_batch=1):
        h = P()
        check(_create(int(slots), int(max_batch), int(mem_budget), int(anchor_period), int(max_iter), C.byref(h)),
              "tt_server_create_batched")
        self._h = h
        self._hosts = {}

    def submit(self, host, prec="fp64"):
        rid = i64()
        check(_submit(self._h, host._h, PREC[prec] if isinstance(prec, str) else int(prec), C.byref(rid)),
              "tt_server_submit")
        self._hosts[rid.value] = host  # keep the inputs alive while queued / running
        return rid.value

    def wait(self, rid):
        info = ServerInfo()
        rc = _wait(self._h, int(rid), C.byref(info))
        out = {k: getattr(info, k) for k, _ in ServerInfo._fields_ if k != "error"}
        out["error"] = info.error.decode(errors="replace")
        if rc < 0 and not out["error"]:
            out["error"] = _lib.tt_last_error().decode(errors="replace")
        return out

    def fields(self, rid):
        m = self._hosts[rid].view().mesh
        nn, nq = m.nnodes, m.nq
        f = dict(velocity=np.zeros((nn, 3), order="F"), vorticity=np.zeros((nn, 3), order="F"),
                 qvorticity=np.zeros((nq, 3), order="F"), temp=np.zeros(nn), qtemp=np.zeros(nq))
        check(_fields(self._h, int(rid), *[f[k].ctypes.data_as(PD) for k in
                                           ("velocity", "vorticity", "qvorticity", "temp", "qtemp")]),
              "tt_server_fields")
        return f

    def release(self, rid):
        check(_release(self._h, int(rid)), "tt_server_release")
        self._hosts.pop(rid, None)

    def load(self):
        q, b, bud = i32(), i64(), i64()
        check(_load(self._h, C.byref(q), C.byref(b), C.byref(bud)), "tt_server_load")
        return dict(queued=q.value, bytes_in_use=b.value, budget=bud.value)

    def close(self):
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None
            self._hosts.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
