"""Sparse and dense solver entry points of the C-ABI: CRSxV, prelsqr2 +
slvlsqr2 (LSQR on a CSR system), dgefa / dgesl, and the device-resident
single-domain kinematics (sdKm_corInt + sdKm_sysm + FormPREfm, sdkm_solve).
Index arrays are the reference's: 1-based, int32."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import P, check, i32

PD, PI = _lib.PD, _lib.PI
_csr_mv = _lib._sig("tt_csr_mv", i32, i32, i32, i32, PI, PI, PD, PD, PD)
_lsqr_setup = _lib._sig("tt_lsqr_setup", i32, i32, i32, i32, PI, PI, PD, PD, i32, C.POINTER(P))
_lsqr_solve = _lib._sig("tt_lsqr_solve", i32, P, PD, PD, i32, C.c_float, PI, PI)
_lsqr_solve_noalloc = _lib._sig("tt_lsqr_solve_noalloc", i32, P, P, P, i32, C.c_float, PI, PI)
_lsqr_free = _lib._sig("tt_lsqr_free", None, P)
_lu = _lib._sig("tt_dense_lu", i32, i32, PD, PD)
_lu_solve = _lib._sig("tt_dense_lu_solve", i32, i32, PD, PD, PD)
_sd_create = _lib._sig("tt_sd_create", i32, C.POINTER(_lib.MeshView), C.POINTER(_lib.Quad), PI, C.POINTER(P))
_sd_solve = _lib._sig("tt_sd_solve", i32, P, PD, PD, PD)
_sd_refine = _lib._sig("tt_sd_refine", i32, P, i32, i32, i32, PD, PD, C.c_double, i32, PI, PD, PD)
SD_PREC = {"fp64": 0, "fp32": 1, "bf16": 2, "fp16": 3}
_sd_free = _lib._sig("tt_sd_free", None, P)
_sd_hmat = _lib._sig("tt_sd_hmat", i32, P, i32, i32, C.c_double, i32, C.c_double, C.POINTER(P))
_sd_mv = _lib._sig("tt_sd_matvec", i32, P, i32, PD, PD)
SD_MATS = {"H": 0, "Htx": 1, "Hty": 2, "Htz": 3, "Dx": 4, "Dy": 5, "Dz": 6}


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def csr_mv(ia, ja, a, x, m=None):
    """CRSxV: y = A x for a 1-based CSR with m = len(ia) - 1 rows."""
    ia, ja, a, x = _i32(ia), _i32(ja), _f64(a), _f64(x)
    m = len(ia) - 1 if m is None else m
    y = np.zeros(m)
    check(_csr_mv(m, len(x), len(ja), ia.ctypes.data_as(PI), ja.ctypes.data_as(PI), a.ctypes.data_as(PD),
                  x.ctypes.data_as(PD), y.ctypes.data_as(PD)), "tt_csr_mv")
    return y


class Lsqr:
    """prelsqr2 at construction (q and, when scale_host, the scaled values in
    self.a), slvlsqr2 in solve()."""

    def __init__(self, ia, ja, a, n, scale_host=True):
        self.ia, self.ja, self.a = _i32(ia), _i32(ja), np.array(a, dtype=np.float64)
        self.m, self.n = len(self.ia) - 1, int(n)
        self.q = np.zeros(self.n)
        h = P()
        check(_lsqr_setup(self.m, self.n, len(self.ja), self.ia.ctypes.data_as(PI), self.ja.ctypes.data_as(PI),
                          self.a.ctypes.data_as(PD), self.q.ctypes.data_as(PD), int(scale_host), C.byref(h)),
              "tt_lsqr_setup")
        self._h = h

    def solve(self, b, x0, maxit, eps4):
        """Returns (x, istop, nit) like slvlsqr2."""
        b, x = _f64(b), np.array(x0, dtype=np.float64)
        nit, istop = i32(), i32()
        check(_lsqr_solve(self._h, b.ctypes.data_as(PD), x.ctypes.data_as(PD), int(maxit), float(eps4),
                          C.byref(nit), C.byref(istop)), "tt_lsqr_solve")
        return x, istop.value, nit.value

    def solve_noalloc(self, b_dev, x_dev, maxit, eps4):
        """tt_lsqr_solve_noalloc on device pointers (integers): b_dev (m
        doubles) and x_dev (n doubles, hot start in, solution out).  Returns
        (istop, nit)."""
        nit, istop = i32(), i32()
        check(_lsqr_solve_noalloc(self._h, P(b_dev), P(x_dev), int(maxit), float(eps4), C.byref(nit),
                                  C.byref(istop)), "tt_lsqr_solve_noalloc")
        return istop.value, nit.value

    def close(self):
        if getattr(self, "_h", None):
            _lsqr_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_bsetup = _lib._sig("tt_lsqr_batch_setup", i32, i32, PI, PI, C.POINTER(C.c_int64), PI, PI, PD, PD, i32,
                    C.POINTER(P))
_bsolve = _lib._sig("tt_lsqr_batch_solve", i32, P, PD, PD, i32, C.c_float, PI, PI, C.POINTER(C.c_double))
_bwidth = _lib._sig("tt_lsqr_batch_set_width", i32, P, i32, PI)
_bprec = _lib._sig("tt_lsqr_batch_set_precision", i32, P, i32)
_btune = _lib._sig("tt_lsqr_batch_tune", i32, P, PD, PD, i32, C.c_float, PD, PD)
_bfree = _lib._sig("tt_lsqr_batch_free", None, P)


_kbsetup = _lib._sig("tt_block_batch_setup", i32, i32, PI, PI, C.POINTER(C.c_int64), PI, PI, PD, i32, C.POINTER(P))
_kbprec = _lib._sig("tt_block_batch_set_precision", i32, P, i32)
_kbsolve = _lib._sig("tt_block_batch_solve", i32, P, PD, PD, i32, C.c_float, PI, PI, C.POINTER(C.c_double))
_kbfree = _lib._sig("tt_block_batch_free", None, P)
_share = _lib._sig("tt_lsqr_set_block_sharing", i32, i32)
_share_stats = _lib._sig("tt_lsqr_block_sharing_stats", i32, C.POINTER(C.c_int64), C.POINTER(C.c_int64))


def set_block_sharing(on):
    """Batched solves whose systems have bit-identical FP32/FP16 blocks read
    one copy of the blocks, their cell-kernel workgroups interleaved
    (default on; process-wide)."""
    check(_share(1 if on else 0), "tt_lsqr_set_block_sharing")


def block_sharing_stats():
    """(interleaved shared-block launches, systems they carried) since the
    library loaded."""
    a, b = C.c_int64(), C.c_int64()
    check(_share_stats(C.byref(a), C.byref(b)), "tt_lsqr_block_sharing_stats")
    return a.value, b.value


class BlockBatch:
    """Independent subdomain systems [(ia, ja, a, n), ...] (1-based CSR, rows
    in groups of rows_per_cell sharing their columns) solved on the cell-block
    format, in groups of up to 32 systems per launch."""

    def __init__(self, systems, rows_per_cell=51):
        self.m = np.array([len(s[0]) - 1 for s in systems], np.int32)
        self.n = np.array([int(s[3]) for s in systems], np.int32)
        self.nnz = np.array([len(s[1]) for s in systems], np.int64)
        ia = np.concatenate([_i32(s[0]) for s in systems])
        ja = np.concatenate([_i32(s[1]) for s in systems])
        a = np.concatenate([np.asarray(s[2], np.float64) for s in systems])
        h = P()
        check(_kbsetup(len(systems), self.m.ctypes.data_as(PI), self.n.ctypes.data_as(PI),
                       self.nnz.ctypes.data_as(C.POINTER(C.c_int64)), ia.ctypes.data_as(PI), ja.ctypes.data_as(PI),
                       a.ctypes.data_as(PD), int(rows_per_cell), C.byref(h)), "tt_block_batch_setup")
        self._h = h
        self.noff = np.concatenate([[0], np.cumsum(self.n)])

    def set_precision(self, prec):
        """Blocks the iterations read: "fp64", "fp32" or "fp16"."""
        check(_kbprec(self._h, SD_PREC[prec]), "tt_block_batch_set_precision")

    def solve(self, b, x0, maxit, eps4):
        """b, x0: lists (or concatenated arrays).  Returns (xs, istop, nit, ms)."""
        b = _f64(np.concatenate(b) if isinstance(b, (list, tuple)) else b)
        x = np.array(np.concatenate(x0) if isinstance(x0, (list, tuple)) else x0, dtype=np.float64)
        k = len(self.m)
        nit, istop, ms = np.zeros(k, np.int32), np.zeros(k, np.int32), C.c_double()
        check(_kbsolve(self._h, b.ctypes.data_as(PD), x.ctypes.data_as(PD), int(maxit), float(eps4),
                       nit.ctypes.data_as(PI), istop.ctypes.data_as(PI), C.byref(ms)), "tt_block_batch_solve")
        return [x[self.noff[i]:self.noff[i + 1]] for i in range(k)], istop, nit, ms.value

    def close(self):
        if getattr(self, "_h", None):
            _kbfree(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LsqrBatch:
    """Independent CSR systems [(ia, ja, a, n), ...] (1-based) solved together
    in one launch; prelsqr2 of each at construction (self.q, and the scaled
    values in self.a when scale_host)."""

    def __init__(self, systems, scale_host=True):
        self.m = np.array([len(s[0]) - 1 for s in systems], np.int32)
        self.n = np.array([int(s[3]) for s in systems], np.int32)
        self.nnz = np.array([len(s[1]) for s in systems], np.int64)
        ia = np.concatenate([_i32(s[0]) for s in systems])
        ja = np.concatenate([_i32(s[1]) for s in systems])
        self.a = np.concatenate([np.asarray(s[2], np.float64) for s in systems])
        self.q = np.zeros(int(self.n.sum()))
        h = P()
        check(_bsetup(len(systems), self.m.ctypes.data_as(PI), self.n.ctypes.data_as(PI),
                      self.nnz.ctypes.data_as(C.POINTER(C.c_int64)), ia.ctypes.data_as(PI), ja.ctypes.data_as(PI),
                      self.a.ctypes.data_as(PD), self.q.ctypes.data_as(PD), int(scale_host), C.byref(h)),
              "tt_lsqr_batch_setup")
        self._h = h
        self.moff = np.concatenate([[0], np.cumsum(self.m)])
        self.noff = np.concatenate([[0], np.cumsum(self.n)])

    def set_width(self, width):
        """Gang width of every system (0 = automatic); returns the widest gang."""
        mw = C.c_int32()
        check(_bwidth(self._h, int(width), C.byref(mw)), "tt_lsqr_batch_set_width")
        return mw.value

    def tune(self, b, x0, maxit, eps4):
        """Gang-width autotuner: returns (best scale, {scale: ms})."""
        b = _f64(np.concatenate(b) if isinstance(b, (list, tuple)) else b)
        x = _f64(np.concatenate(x0) if isinstance(x0, (list, tuple)) else x0)
        best, ms = C.c_double(), np.zeros(4)
        check(_btune(self._h, b.ctypes.data_as(PD), x.ctypes.data_as(PD), int(maxit), float(eps4), C.byref(best),
                     ms.ctypes.data_as(PD)), "tt_lsqr_batch_tune")
        return best.value, dict(zip((0.5, 1.0, 2.0, 4.0), ms.tolist()))

    def set_precision(self, prec):
        """Values the LSQR products read: "fp64" or "fp32"."""
        check(_bprec(self._h, SD_PREC[prec]), "tt_lsqr_batch_set_precision")

    def solve(self, b, x0, maxit, eps4):
        """b, x0: lists (or concatenated arrays).  Returns (xs, istop, nit, ms)."""
        b = _f64(np.concatenate(b) if isinstance(b, (list, tuple)) else b)
        x = np.array(np.concatenate(x0) if isinstance(x0, (list, tuple)) else x0, dtype=np.float64)
        k = len(self.m)
        nit, istop, ms = np.zeros(k, np.int32), np.zeros(k, np.int32), C.c_double()
        check(_bsolve(self._h, b.ctypes.data_as(PD), x.ctypes.data_as(PD), int(maxit), float(eps4),
                      nit.ctypes.data_as(PI), istop.ctypes.data_as(PI), C.byref(ms)), "tt_lsqr_batch_solve")
        xs = [x[self.noff[i]:self.noff[i + 1]] for i in range(k)]
        return xs, istop, nit, ms.value

    def close(self):
        if getattr(self, "_h", None):
            _bfree(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dense_lu(a):
    """dgefa: returns (factor in LINPACK storage, pivots as float64, info)."""
    a = np.array(a, dtype=np.float64, order="F")
    n = a.shape[0]
    piv = np.zeros(n)
    info = check(_lu(n, a.ctypes.data_as(PD), piv.ctypes.data_as(PD)), "tt_dense_lu")
    return a, piv, info


def dense_lu_solve(lu, piv, b):
    """dgesl job = 0."""
    lu = np.asfortranarray(lu, dtype=np.float64)
    piv, b = _f64(piv), np.array(b, dtype=np.float64)
    check(_lu_solve(lu.shape[0], lu.ctypes.data_as(PD), piv.ctypes.data_as(PD), b.ctypes.data_as(PD)),
          "tt_dense_lu_solve")
    return b


class SdKinematics:
    """Device-resident dense single-domain kinematics of a Host's mesh."""

    def __init__(self, host, quad):
        self.host = host
        self.view = host.view()
        self._quad = quad
        h = P()
        check(_sd_create(C.byref(self.view.mesh), C.byref(quad), self.view.kode, C.byref(h)), "tt_sd_create")
        self._h = h

    def solve(self, vel, vort, urbw=(1.0, 1.0, 1.0)):
        """One boundary-vorticity update (x, y, z); returns the new vorticity."""
        vel = np.asfortranarray(vel, dtype=np.float64)
        w = np.array(vort, dtype=np.float64, order="F")
        ur = _f64(urbw)
        check(_sd_solve(self._h, vel.ctypes.data_as(PD), w.ctypes.data_as(PD), ur.ctypes.data_as(PD)), "tt_sd_solve")
        return w

    def hmat(self, which="H", leaf=64, eta=1.0, aca_type=2, aca_eps=1e-6):
        """Block matrix of one device-resident matrix, admissible blocks
        cross-approximated (tt_sd_hmat); returns a compress.HMatrix."""
        from .compress import HMatrix
        h = P()
        check(_sd_hmat(self._h, SD_MATS[which], int(leaf), float(eta), int(aca_type), float(aca_eps), C.byref(h)),
              "tt_sd_hmat")
        m = self.view.mesh
        return HMatrix.from_handle(h, m.nbnodes, m.nbnodes if which.startswith("H") else m.nnodes)

    def matvec(self, which, x):
        """Dense product of one device-resident matrix (tt_sd_matvec)."""
        x = _f64(x)
        y = np.zeros(self.view.mesh.nbnodes)
        check(_sd_mv(self._h, SD_MATS[which], x.ctypes.data_as(PD), y.ctypes.data_as(PD)), "tt_sd_matvec")
        return y

    def refine(self, eqn, prec, B, tol=1e-10, maxit=50):
        """Solves A_eqn X = B (nb x nrhs, nrhs <= 4) by FP64 iterative
        refinement with the inverse applied in prec.  Returns (X, corrections,
        relative residual, device ms of the solve, mean ms per correction)."""
        B = np.asfortranarray(B, dtype=np.float64)
        if B.ndim == 1:
            B = B[:, None]
        X = np.zeros_like(B, order="F")
        it, rr, ms = C.c_int32(), C.c_double(), np.zeros(2)
        check(_sd_refine(self._h, int(eqn), SD_PREC[prec], B.shape[1], B.ctypes.data_as(PD), X.ctypes.data_as(PD),
                         float(tol), int(maxit), C.byref(it), C.byref(rr), ms.ctypes.data_as(PD)), "tt_sd_refine")
        return X, it.value, rr.value, ms[0], ms[1]

    def close(self):
        if getattr(self, "_h", None):
            _sd_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
