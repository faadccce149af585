"""The allocation-free LSQR entry point (tt_lsqr_solve_noalloc) against
tt_lsqr_solve through the C-ABI: same kernels and order of sums, so the
solutions, iteration counts and istop must be identical (bit for bit), on
two calls in a row, the second on a smaller system; and a system larger than
the scratch is rejected with an error before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Dev:
    """Device copies through the HIP runtime the library itself loaded
    (libamdhip64.so.7 is resolved to the already-loaded runtime)."""

    def __init__(self):
        from tritok_amd import _lib  # noqa: F401  (loads the library and its runtime first)
        self.hip = C.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p.value

    def get(self, ptr, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        assert self.hip.hipDeviceSynchronize() == 0
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _random_csr(m, n, per_row, seed):
    rng = np.random.default_rng(seed)
    ja = np.concatenate([np.sort(rng.choice(n, per_row, replace=False)) for _ in range(m)]) + 1
    ia = np.arange(0, m * per_row + 1, per_row) + 1
    a = rng.standard_normal(m * per_row)
    return ia.astype(np.int32), ja.astype(np.int32), a


def test_noalloc_solve_matches_lsqr_solve():
    from tritok_amd import solvers as S
    dev = _Dev()
    for m, n, k, seed in ((3000, 1000, 11, 3), (1200, 400, 9, 4)):
        print("noalloc case m=%d n=%d per_row=%d seed=%d" % (m, n, k, seed), flush=True)
        ia, ja, a = _random_csr(m, n, k, seed)
        rng = np.random.default_rng(seed + 10)
        b = rng.standard_normal(m)
        x0 = 0.1 * rng.standard_normal(n)
        s = S.Lsqr(ia, ja, a, n)
        for eps in (1e-7, 1e-10):
            x_ref, istop_ref, nit_ref = s.solve(b, x0, 5000, eps)
            bd, xd = dev.put(b), dev.put(x0)
            istop, nit = s.solve_noalloc(bd, xd, 5000, eps)
            x = dev.get(xd, n)
            print("  eps %g: nit %d / %d, istop %d / %d, max |dx| %.3g" % (
                eps, nit, nit_ref, istop, istop_ref, np.max(np.abs(x - x_ref))), flush=True)
            assert istop == istop_ref and nit == nit_ref
            assert np.array_equal(x, x_ref)
        s.close()
    dev.free()


def test_noalloc_rejects_system_beyond_scratch():
    from tritok_amd import solvers as S
    from tritok_amd._lib import TTError
    m, n = 4_300_000, 8  # m + 3n above the 32 MiB scratch (4,194,304 doubles)
    ia = np.arange(1, m + 2, dtype=np.int32)
    ja = (np.arange(m, dtype=np.int32) % n) + 1
    a = np.ones(m)
    s = S.Lsqr(ia, ja, a, n, scale_host=False)
    dev = _Dev()
    bd, xd = dev.put(np.ones(m)), dev.put(np.zeros(n))
    with pytest.raises(TTError, match="scratch"):
        s.solve_noalloc(bd, xd, 100, 1e-7)
    s.close()
    dev.free()
