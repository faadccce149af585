"""Host-side preprocessing (the reference's Fortran host side) exposed through
the C-ABI: mesh reading / generation, flux-node macro mesh, boundary
conditions, unknown numbering and tri.inp parameters."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CaseView, check, dptr, i32, iptr, lib, P

_from_arrays = _lib._sig("tt_host_from_arrays", i32, i32, i32, i32, i32, _lib.PD, _lib.PI, _lib.PI, _lib.PI,
                         C.POINTER(P))
_read_geo = _lib._sig("tt_host_read_geo", i32, C.c_char_p, C.POINTER(P))
_cube = _lib._sig("tt_host_cube", i32, i32, C.POINTER(P))
_read_bic = _lib._sig("tt_host_read_bic", i32, P, C.c_char_p)
_read_inp = _lib._sig("tt_host_read_inp", i32, P, C.c_char_p)
_case_view = _lib._sig("tt_host_case_view", i32, P, C.POINTER(CaseView))
_fields = _lib._sig("tt_host_fields", i32, P, C.POINTER(_lib.PD), C.POINTER(_lib.PD), C.POINTER(_lib.PD),
                    C.POINTER(_lib.PD), C.POINTER(_lib.PD))
_write_geo = _lib._sig("tt_host_write_geo", i32, P, C.c_char_p)
_free = _lib._sig("tt_host_free", None, P)


def _arr(ptr, n, dtype, shape=None):
    if n == 0:
        return np.zeros(0, dtype)
    ctype = C.c_double if dtype == np.float64 else C.c_int32
    a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()
    if shape is not None:
        a = a.reshape(shape[::-1]).T  # column-major (rows, cols)
    return a


class Host:
    """One case's host data.  Arrays are returned as numpy copies with the
    reference's shapes (column-major order preserved via transposes)."""

    def __init__(self, handle):
        self._h = handle if isinstance(handle, P) else P(handle)

    @classmethod
    def from_geo(cls, path):
        h = P()
        check(_read_geo(path.encode(), C.byref(h)), "tt_host_read_geo")
        return cls(h)

    @classmethod
    def from_arrays(cls, x, ibc, idc, bewn, nbnodes=0):
        x = np.asfortranarray(x, dtype=np.float64)
        ibc = np.asfortranarray(ibc, dtype=np.int32)
        idc = np.asfortranarray(idc, dtype=np.int32)
        bewn = np.ascontiguousarray(bewn, dtype=np.int32)
        h = P()
        check(_from_arrays(len(x), int(nbnodes), len(ibc), len(idc), dptr(x), iptr(ibc), iptr(idc), iptr(bewn),
                           C.byref(h)), "tt_host_from_arrays")
        return cls(h)

    @classmethod
    def cube(cls, n):
        h = P()
        check(_cube(int(n), C.byref(h)), "tt_host_cube")
        return cls(h)

    def read_bic(self, text):
        check(_read_bic(self._h, text.encode()), "tt_host_read_bic")
        return self

    def read_inp(self, text):
        check(_read_inp(self._h, text.encode()), "tt_host_read_inp")
        return self

    def write_geo(self, path):
        check(_write_geo(self._h, path.encode()), "tt_host_write_geo")

    def view(self):
        v = CaseView()
        check(_case_view(self._h, C.byref(v)), "tt_host_case_view")
        return v

    def mesh_arrays(self):
        """Mesh and macro-mesh arrays as numpy (1-based indices)."""
        v = self.view()
        m = v.mesh
        nn, nb, ne, nc, nq = m.nnodes, m.nbnodes, m.nbelem, m.nicell, m.nq
        return dict(
            nn=nn, nb=nb, ne=ne, nc=nc, nq=nq,
            x=_arr(m.x, nn * 3, np.float64, (nn, 3)), ibc=_arr(m.ibc, ne * 9, np.int32, (ne, 9)),
            idc=_arr(m.idc, nc * 27, np.int32, (nc, 27)), ibf=_arr(m.ibf, nc * 24, np.int32).reshape(4, 6, nc).T,
            xq=_arr(m.xq, nq * 3, np.float64, (nq, 3)), flpm=_arr(m.flpm, nc * 6, np.int32, (nc, 6)),
            gbn=_arr(m.gbn, nb, np.int32), lbn=_arr(m.lbn, nn, np.int32),
            nx=_arr(m.nx, nb, np.float64), ny=_arr(m.ny, nb, np.float64), nz=_arr(m.nz, nb, np.float64))

    def fields(self):
        ps = [_lib.PD() for _ in range(5)]
        check(_fields(self._h, *[C.byref(p) for p in ps]), "tt_host_fields")
        m = self.view().mesh
        nn, nq = m.nnodes, m.nq
        return dict(velocity=_arr(ps[0], nn * 3, np.float64, (nn, 3)),
                    vorticity=_arr(ps[1], nn * 3, np.float64, (nn, 3)),
                    qvorticity=_arr(ps[2], nq * 3, np.float64, (nq, 3)),
                    temp=_arr(ps[3], nn, np.float64), qtemp=_arr(ps[4], nq, np.float64))

    def close(self):
        if self._h:
            _free(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
