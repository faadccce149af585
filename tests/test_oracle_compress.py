"""Pins the oracle's compressed-matrix routines (cross approximation, stretched
Haar wavelets, H-matrix product) against the reference's own routines and
against constructions written here from the mathematics, independently of the
oracle's code.

compress_ref.npz holds what the reference's aca_CrossAprox, ACA_AB,
aca_RankCrossAprox (src/lib_ACA.f) and Che_WcomMatrix (src/lib_WT.f:148-210)
return for seeded matrices (tests/golden/dump_compress.f90, run by
make_golden.py): ranks, factors, residuals and CSR, which the oracle must
reproduce bit for bit, including the stopping rank of EPS_ACA_NORM's
inner-product recurrence (lib_ACA.f:438-474) and Che_Zanemari's last-column /
1e-14 rule (lib_WT.f:741-777)."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden

_REF = load_golden("compress_ref")
_ACA_CASES = [i for i in range(int(_REF["ncase"])) if int(_REF["c%02d_kind" % i]) <= 3]
_WT_CASES = [i for i in range(int(_REF["ncase"])) if int(_REF["c%02d_kind" % i]) == 4]


@pytest.mark.parametrize("i", _ACA_CASES)
def test_aca_matches_reference_routines(i):
    p = "c%02d_" % i
    kind, eps, A = int(_REF[p + "kind"]), float(_REF[p + "eps"]), _REF[p + "A"]
    vec = _REF[p + "vec"] if kind != 2 else None
    a, b, r, nrm, R = O.aca(kind, A, eps, vec=vec)
    assert r == int(_REF[p + "rank"])
    assert np.array_equal(a, _REF[p + "a"]) and np.array_equal(b, _REF[p + "b"])
    assert np.array_equal(R, _REF[p + "residual"])  # the residual the routine leaves in its argument
    if kind != 2:  # random-vector norm (the reference forms A vec with MATMUL: summation order may differ)
        assert abs(nrm - float(_REF[p + "norm"])) <= 1e-10 * float(_REF[p + "norm"])


def _textbook_aca_rank(A, eps):
    """Cross approximation stopped by the textbook test ||a_k|| ||b_k|| <=
    eps ||S_k||_F with S_k = sum of the rank-one terms, ||S_k||_F formed
    exactly: the rule the reference's EPS_ACA_NORM recurrence approximates."""
    R = A.copy()
    a, b = [], []
    for k in range(min(A.shape)):
        i, j = np.unravel_index(np.argmax(np.abs(R)), R.shape)  # row-major first maximum, as APPLAY_ACA
        av, bv = R[:, j].copy(), R[i, :] / R[i, j]
        R -= np.outer(av, bv)
        a.append(av), b.append(bv)
        S = np.array(a).T @ np.array(b)
        if np.linalg.norm(av) * np.linalg.norm(bv) <= eps * np.linalg.norm(S):
            return k + 1
    return min(A.shape)


def test_aca_type2_reference_rank_is_not_the_textbook_rank():
    """At least one fixture stops at a rank the textbook Frobenius test does
    not give, so the bit-exact rank check above pins the reference's own
    recurrence, not just the error bound."""
    diff = []
    for i in _ACA_CASES:
        p = "c%02d_" % i
        if int(_REF[p + "kind"]) == 2:
            t = _textbook_aca_rank(_REF[p + "A"], float(_REF[p + "eps"]))
            if t != int(_REF[p + "rank"]):
                diff.append((i, t, int(_REF[p + "rank"])))
    assert diff, "every type-2 fixture stops where the textbook test does"


@pytest.mark.parametrize("i", _WT_CASES)
def test_wavelet_matches_reference_routines(i):
    p = "c%02d_" % i
    A, kappa = _REF[p + "A"], float(_REF[p + "eps"])
    w = O.wt_compress(A, kappa)
    assert np.array_equal(w["zvr"], _REF[p + "zvr"]) and np.array_equal(w["c"], _REF[p + "c"])
    assert np.array_equal(w["v"], _REF[p + "v"])
    assert w["nnz_ref"] == int(_REF[p + "nnz"])  # m*n - (entries zeroed), as crsMat%nnz
    # the reference's own accuracy check: W^T C W vec against A vec
    y, y1 = O.wt_wtawx(w, _REF[p + "vec"], A.shape[0]), A @ _REF[p + "vec"]
    nrm = np.sqrt(np.sum((y1 - y) ** 2) / np.sum(y1 ** 2))
    assert abs(nrm - float(_REF[p + "norm"])) <= 1e-9 * float(_REF[p + "norm"]) + 1e-14


def test_wavelet_fixtures_exercise_the_last_column_rule():
    """Rows whose interior coefficients all fall under the threshold keep
    their last column, set to 1e-14 when it is exactly zero: the constant-row
    fixtures hold such entries."""
    counts = {i: int(np.sum(_REF["c%02d_v" % i] == 1e-14)) for i in _WT_CASES}
    assert sum(counts.values()) >= 10, counts


def _kernel_block(m, n, sep=3.0, seed=0):
    """Smooth BEM-like block: 1/|x - y| between two separated point clouds."""
    rng = np.random.default_rng(seed)
    x = rng.random((m, 3))
    y = rng.random((n, 3)) + np.array([sep, 0.0, 0.0])
    return 1.0 / np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2)


def _brute_aca_colmajor(A, rank):
    """Full-pivot cross approximation written from its definition: pick the
    largest |R| (first in column-major order), take its column divided by the
    pivot and its row, subtract the rank-one product."""
    R = A.copy()
    m, n = R.shape
    a = np.zeros((m, rank))
    b = np.zeros((rank, n))
    for k in range(rank):
        flat = np.abs(R).ravel(order="F")
        p = int(np.argmax(flat))  # argmax returns the first maximum
        i, j = p % m, p // m
        a[:, k] = R[:, j] / R[i, j]
        b[k, :] = R[i, :]
        R = R - np.outer(a[:, k], b[k, :])
    return a, b


def test_aca_type1_matches_brute_force():
    A = np.random.default_rng(3).standard_normal((23, 17))
    a, b, r, _, _ = O.aca(1, A, 0.3, vec=None)
    assert r == int(0.3 * 17)
    ra, rb = _brute_aca_colmajor(A, r)
    assert np.allclose(a, ra, rtol=1e-12, atol=1e-14) and np.allclose(b, rb, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("t", [1, 2, 3])
def test_aca_recovers_low_rank_exactly(t):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((40, 4)) @ rng.standard_normal((4, 31))
    eps = {1: 0.13, 2: 1e-12, 3: 1e-13}[t]
    a, b, r, nrm, R = O.aca(t, A, eps, vec=rng.random(31))
    assert r >= 4 and np.abs(a @ b - A).max() < 1e-12 * np.abs(A).max()
    assert np.abs(R).max() < 1e-12 * np.abs(A).max()  # the residual left in place


def test_aca_type2_error_follows_tolerance():
    """On a smooth separated block the Frobenius-type stopping test gives an
    approximation error of the order of eps, with more rank for smaller eps."""
    A = _kernel_block(120, 90)
    ranks = []
    for eps in (1e-3, 1e-6, 1e-9):
        a, b, r, _, _ = O.aca(2, A, eps)
        err = np.linalg.norm(A - a @ b) / np.linalg.norm(A)
        assert err < 30 * eps, (eps, err)
        ranks.append(r)
    assert ranks[0] < ranks[1] < ranks[2] < 40


def test_aca_type3_passes_its_vector_test():
    A = _kernel_block(80, 70, seed=1)
    vec = np.random.default_rng(9).random(70)
    a, b, r, nrm, _ = O.aca(3, A, 1e-8, vec=vec)
    ref = A @ vec
    got = a @ (b @ vec)
    assert np.sqrt(np.sum((ref - got) ** 2) / np.sum(ref ** 2)) < 1e-8 and nrm < 1e-8 and r < 40
    assert np.allclose(O.aca_abxv(a, b, vec), got, rtol=1e-13, atol=0)


def _haar(N):
    """Orthonormal Haar matrix, coarse-to-fine rows, by the Kronecker recursion
    H_2N = [H_N (x) (1,1); I_N (x) (1,-1)] / sqrt(2)."""
    H = np.array([[1.0]])
    while H.shape[0] < N:
        n = H.shape[0]
        H = np.vstack([np.kron(H, [1.0, 1.0]), np.kron(np.eye(n), [1.0, -1.0])]) / np.sqrt(2.0)
    return H


@pytest.mark.parametrize("N", [2, 8, 64, 256])
def test_haar_transforms_equal_the_haar_matrix(N):
    H = _haar(N)
    assert np.allclose(H @ H.T, np.eye(N), atol=1e-14)
    x = np.random.default_rng(N).standard_normal(N)
    assert np.allclose(O.wt_wxv(x), H @ x, rtol=0, atol=1e-13)
    assert np.allclose(O.wt_wtxv(x), H.T @ x, rtol=0, atol=1e-13)


def test_dvanak():
    for n, k in [(1, 0), (2, 0), (3, 1), (4, 1), (5, 2), (33, 5), (64, 5), (65, 6), (9602, 13)]:
        assert O.dvanak(n) == k, n


@pytest.mark.parametrize("m,n", [(16, 16), (13, 21), (5, 7), (100, 37), (1, 9)])
def test_wavelet_product_is_exact_without_threshold(m, n):
    """kappa = 0 keeps every coefficient: W^T (W A W^T) (W x) = A x for any
    size, which fixes the padding by duplication / zero insertion and the
    compaction back to the original length."""
    rng = np.random.default_rng(m * 100 + n)
    A = rng.standard_normal((m, n))
    w = O.wt_compress(A, 0.0)
    x = rng.standard_normal(n)
    assert np.abs(O.wt_wtawx(w, x, m) - A @ x).max() < 1e-13 * np.abs(A).sum(1).max()
    assert len(w["v"]) == w["nnz_ref"] == m * n


def test_wavelet_threshold_keeps_scaling_entries_and_bounds_error():
    A = _kernel_block(150, 120, sep=1.5)
    x = np.random.default_rng(2).standard_normal(120)
    errs, nnz = [], []
    for kappa in (1e-4, 1e-2, 1.0):
        w = O.wt_compress(A, kappa)
        T = w["T"]
        assert np.all(T[0, :] != 0) and np.all(T[:, 0] != 0)  # scaling row / column kept
        dense = np.zeros_like(T)
        for r in range(T.shape[0]):
            for k in range(w["zvr"][r] - 1, w["zvr"][r + 1] - 1):
                dense[r, w["c"][k] - 1] = w["v"][k]
        assert np.array_equal(dense, T)  # the CSR holds exactly the kept entries
        errs.append(np.linalg.norm(O.wt_wtawx(w, x, 150) - A @ x) / np.linalg.norm(A @ x))
        nnz.append(len(w["v"]))
    assert errs[0] < errs[1] < errs[2] and nnz[0] > nnz[1] > nnz[2]
    assert errs[0] < 1e-5


def test_hmat_product_sums_parts():
    """A 2x2 block partition of a random matrix with one dense, one wavelet and
    two cross-approximated parts (full rank, so exact): y = A x."""
    rng = np.random.default_rng(4)
    A = rng.standard_normal((30, 26))
    perm_r = rng.permutation(30) + 1
    perm_c = rng.permutation(26) + 1
    rs, cs = [perm_r[:13], perm_r[13:]], [perm_c[:11], perm_c[11:]]
    parts = []
    for i, r in enumerate(rs):
        for j, c in enumerate(cs):
            blk = A[np.ix_(r - 1, c - 1)]
            if (i, j) == (0, 0):
                parts.append(dict(rows=r, cols=c, kind="dense", dense=blk))
            elif (i, j) == (0, 1):
                parts.append(dict(rows=r, cols=c, kind="wt", wt=O.wt_compress(blk, 0.0)))
            else:
                a, b, _, _, _ = O.aca(1, blk, 1.0)
                parts.append(dict(rows=r, cols=c, kind="aca", aca=(a, b)))
    x = rng.standard_normal(26)
    assert np.abs(O.hmat_mxv(30, parts, x) - A @ x).max() < 1e-12


def _sphere(n, seed=5):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    d = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
    np.fill_diagonal(d, 1.0)
    A = 1.0 / (4 * np.pi * d)
    np.fill_diagonal(A, 1.0)
    return p, A


def test_library_partition_tiles_and_is_admissible():
    """The library's host-side partition (tt_hmat_partition): every entry of
    the matrix in exactly one part, dense parts within the leaf size, and the
    admissibility test min(diam) < eta * dist on the parts' bounding boxes."""
    from tritok_amd import compress as K
    p, _ = _sphere(900)
    leaf, eta = 32, 1.5
    rptr, rn, cptr, cn, adm = K.partition(p, p, leaf, eta)
    cover = np.zeros((900, 900), np.int32)
    for k in range(len(adm)):
        r, c = rn[rptr[k]:rptr[k + 1]] - 1, cn[cptr[k]:cptr[k + 1]] - 1
        cover[np.ix_(r, c)] += 1
        br, bc = p[r], p[c]
        dr = np.linalg.norm(br.max(0) - br.min(0))
        dc = np.linalg.norm(bc.max(0) - bc.min(0))
        gap = np.maximum(0, np.maximum(br.min(0) - bc.max(0), bc.min(0) - br.max(0)))
        if adm[k]:
            assert min(dr, dc) < eta * np.linalg.norm(gap)
        else:
            assert len(r) <= leaf and len(c) <= leaf
    assert np.all(cover == 1)
    assert adm.sum() > 0


def test_hmat_with_cross_approximated_parts_approximates_dense():
    """The oracle's block product over the library's partition, admissible
    parts cross-approximated to 1e-8: relative error of the order of eps."""
    from tritok_amd import compress as K
    p, A = _sphere(1200)
    rptr, rn, cptr, cn, adm = K.partition(p, p, 32, 2.0)
    parts, stored = [], 0
    for k in range(len(adm)):
        r, c = rn[rptr[k]:rptr[k + 1]], cn[cptr[k]:cptr[k + 1]]
        blk = A[np.ix_(r - 1, c - 1)]
        if adm[k]:
            a, b, rk, _, _ = O.aca(2, blk, 1e-8)
            parts.append(dict(rows=r, cols=c, kind="aca", aca=(a, b)))
            stored += (len(r) + len(c)) * rk
        else:
            parts.append(dict(rows=r, cols=c, kind="dense", dense=blk))
            stored += len(r) * len(c)
    x = np.random.default_rng(1).standard_normal(1200)
    y = O.hmat_mxv(1200, parts, x)
    assert np.linalg.norm(y - A @ x) / np.linalg.norm(A @ x) < 1e-6
    assert stored < 0.9 * 1200 * 1200
