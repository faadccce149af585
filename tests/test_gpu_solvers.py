"""GPU solver entry points against the oracle.

Parity bars: column scaling q and the scaled values within 5e-16 / 1e-15
relative (prelsqr2's sums are formed in the same order; the device square
root and division round differently in the last bits); LSQR solutions within 1e-8 relative of the oracle's with
iteration counts within 2 (different summation order in the products);
LU factors within 1e-12 relative with identical pivots; the dense
boundary-vorticity update within 1e-10 relative.
"""
import numpy as np
import pytest

import oracle as O
from conftest import host_for, load_golden, rel_l2

pytestmark = pytest.mark.gpu


def _random_csr(m, n, per_row, seed):
    rng = np.random.default_rng(seed)
    ja = np.concatenate([np.sort(rng.choice(n, per_row, replace=False)) for _ in range(m)]) + 1
    ia = np.arange(0, m * per_row + 1, per_row) + 1
    a = rng.standard_normal(m * per_row)
    return ia.astype(np.int32), ja.astype(np.int32), a


def test_csr_mv_matches_dense_product():
    from tritok_amd import solvers as S
    ia, ja, a = _random_csr(1000, 700, 13, 1)
    x = np.random.default_rng(2).standard_normal(700)
    dense = np.zeros((1000, 700))
    for i in range(1000):
        dense[i, ja[ia[i] - 1:ia[i + 1] - 1] - 1] = a[ia[i] - 1:ia[i + 1] - 1]
    y = S.csr_mv(ia, ja, a, x)
    assert np.allclose(y, dense @ x, rtol=1e-13, atol=1e-13)


def _vte_system(g, e=0):
    """The reference's VTE system of a fixture case (sMat2crsSysRhsB_w_fill):
    H + beta*Re*B on unknown node columns, -G on unknown flux columns, columns
    in increasing unknown order."""
    h = host_for(g)
    v = h.view()
    m = v.mesh
    nc, nn, nq = m.nicell, m.nnodes, m.nq
    arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
    idc = arr(m.idc, nc * 27).reshape(27, nc).T - 1
    ibf = arr(m.ibf, nc * 24).reshape(4, 6, nc).transpose(2, 1, 0).reshape(nc, 24) - 1
    wk = arr(v.wkode, nn * 3).reshape(3, nn)[e]
    wq = arr(v.wqkode, nq * 3).reshape(3, nq)[e]
    beta = 1.5 / v.tstep
    H, G, B = g["macro_H"], g["macro_G"], g["macro_B"]
    rows_i, rows_j, vals = [1], [], []
    for c in range(nc):
        codes = np.concatenate([wk[idc[c]], wq[ibf[c]]])
        for r in range(51):
            row = c * 51 + r
            ent = []
            for col in range(51):
                if codes[col] < 0:
                    val = H[row, col] + B[row, col] * beta * v.ReN if col < 27 else -G[row, col - 27]
                    ent.append((-codes[col], val))
            ent.sort()
            rows_j += [j for j, _ in ent]
            vals += [x for _, x in ent]
            rows_i.append(len(rows_j) + 1)
    return np.array(rows_i, np.int32), np.array(rows_j, np.int32), np.array(vals), int(v.nunk[e])


@pytest.mark.parametrize("kind", ["random", "vte_dc2", "vte_jit2"])
def test_lsqr_matches_oracle(kind):
    from tritok_amd import solvers as S
    if kind == "random":
        ia, ja, a = _random_csr(3000, 1000, 11, 3)
        n = 1000
    else:
        ia, ja, a, n = _vte_system(load_golden(kind.split("_")[1]))
    m = len(ia) - 1
    rng = np.random.default_rng(5)
    b = rng.standard_normal(m)
    x0 = 0.1 * rng.standard_normal(n)
    ao = a.copy()
    q_ref = O.prelsqr2(m, n, ia, ja, ao)  # scales ao in place
    s = S.Lsqr(ia, ja, a, n)
    dq = np.max(np.abs(s.q - q_ref) / q_ref)
    da = np.max(np.abs(s.a - ao) / np.maximum(np.abs(ao), 1e-300))
    print(kind, "q max rel diff", dq, "scaled values", da, flush=True)
    assert dq <= 5e-16 and da <= 1e-15  # a few roundings: the device square root and division
    for eps in (1e-7, 1e-10):
        x_ref, istop_ref, nit_ref = O.slvlsqr2(m, n, ia, ja, ao, q_ref, b, x0, 5000, eps)
        x, istop, nit = s.solve(b, x0, 5000, eps)
        print(kind, eps, "nit", nit, nit_ref, "istop", istop, istop_ref, "rel", rel_l2(x, x_ref), flush=True)
        assert istop == istop_ref
        assert abs(nit - nit_ref) <= 2
        assert rel_l2(x, x_ref) < (1e-8 if eps < 1e-9 else 1e-5)
    s.close()


def test_dense_lu_and_solve_match_linpack():
    from tritok_amd import solvers as S
    rng = np.random.default_rng(11)
    n = 300
    a = rng.standard_normal((n, n))
    lu_ref, piv_ref, info_ref = O.dgefa(a)
    lu, piv, info = S.dense_lu(a)
    assert info == info_ref == 0
    assert np.array_equal(piv, piv_ref)
    assert rel_l2(lu, lu_ref) < 1e-12
    b = rng.standard_normal(n)
    x_ref = O.dgesl(lu_ref, piv_ref, b)
    x = S.dense_lu_solve(lu_ref, piv_ref, b)
    assert rel_l2(x, x_ref) < 1e-13
    assert rel_l2(a @ x, b) < 1e-10


def _oracle_bw_step(h, g, vel, vort, urbw):
    v = h.view()
    m = v.mesh
    nb, nn = m.nbnodes, m.nnodes
    arr = lambda p, n, dt=np.int32: np.ctypeslib.as_array(p, shape=(n,)).copy()
    lbn, gbn = arr(m.lbn, nn), arr(m.gbn, nb)
    kode = arr(v.kode, nb * 3)
    nx, ny, nz = (np.ctypeslib.as_array(p, shape=(nb,)).copy() for p in (m.nx, m.ny, m.nz))
    mats = [np.asfortranarray(g["sd_" + k]) for k in ("H", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz")]
    vel = np.asfortranarray(vel)
    w = np.array(vort, order="F")
    for e in range(3):
        kd = kode.reshape(3, nb)[e]
        A = np.zeros((nb, nb))
        for i in range(nb):
            if kd[i] in (0, 3):
                A[i] = nx[i] * mats[4][i, gbn - 1] + ny[i] * mats[5][i, gbn - 1] + nz[i] * mats[6][i, gbn - 1]
            else:
                A[i, i] = 1.0
        lu, piv, _ = O.dgefa(A)
        dfi = np.zeros(nb)
        O._sdkm_rhs(e + 1, nn, nb, O._i(lbn), O._i(gbn), O._i(kode), O._d(nx), O._d(ny), O._d(nz),
                    *[O._d(x) for x in mats], nb, O._d(vel), O._d(w), O._d(dfi), 1)
        dfi = np.zeros(nb) if O.dnrm2(dfi) == 0.0 else O.dgesl(lu, piv, dfi)
        rows = np.isin(kd, (0, 3))
        w[gbn[rows] - 1, e] = w[gbn[rows] - 1, e] * (1.0 - urbw[e]) + urbw[e] * dfi[rows]
    return w


@pytest.mark.parametrize("case", ["dc2", "jit2"])
def test_sd_kinematics_step_matches_oracle(case, gauss_table):
    from tritok_amd import integrals as I, solvers as S
    g = load_golden(case)
    h = host_for(g)
    nn = h.view().mesh.nnodes
    rng = np.random.default_rng(21)
    vel, vort = rng.standard_normal((nn, 3)), rng.standard_normal((nn, 3))
    urbw = (1.0, 0.7, 0.4)
    sd = S.SdKinematics(h, I.quad(4, 4, 4, 4, 2, table=gauss_table))
    w = sd.solve(vel, vort, urbw)
    w_ref = _oracle_bw_step(h, g, vel, vort, urbw)
    e = rel_l2(w, w_ref)
    print(case, "bw step rel", e, flush=True)
    assert e < 1e-10
    sd.close()


def _oracle_sys(h, g, e):
    """A_e of sdKm_sysm (lib_sdkm.f:1171-1210) from the reference's sdkm matrices."""
    v = h.view()
    m = v.mesh
    nb = m.nbnodes
    gbn = np.ctypeslib.as_array(m.gbn, shape=(nb,)).copy()
    kd = np.ctypeslib.as_array(v.kode, shape=(3 * nb,)).copy().reshape(3, nb)[e]
    nx, ny, nz = (np.ctypeslib.as_array(p, shape=(nb,)).copy() for p in (m.nx, m.ny, m.nz))
    D = [np.asfortranarray(g["sd_" + k]) for k in ("Dx", "Dy", "Dz")]
    A = np.zeros((nb, nb))
    for i in range(nb):
        if kd[i] in (0, 3):
            A[i] = nx[i] * D[0][i, gbn - 1] + ny[i] * D[1][i, gbn - 1] + nz[i] * D[2][i, gbn - 1]
        else:
            A[i, i] = 1.0
    return A


@pytest.mark.parametrize("case", ["dc2", "jit2"])
def test_sd_refine_every_precision_reaches_fp64(case, gauss_table):
    """Four right-hand sides solved by FP64 iterative refinement with the
    inverse in FP64, FP32, BF16 and FP16 on the matrix cores: every precision
    reaches the 1e-10 relative residual (checked here in numpy) and the
    oracle's dgefa/dgesl solution."""
    from tritok_amd import integrals as I, solvers as S
    g = load_golden(case)
    h = host_for(g)
    sd = S.SdKinematics(h, I.quad(4, 4, 4, 4, 2, table=gauss_table))
    rng = np.random.default_rng(31)
    for e in range(3):
        A = _oracle_sys(h, g, e)
        B = rng.standard_normal((A.shape[0], 4))
        lu, piv, _ = O.dgefa(A)
        X_ref = np.stack([O.dgesl(lu, piv, B[:, c].copy()) for c in range(4)], axis=1)
        its = {}
        for prec in ("fp64", "fp32", "bf16", "fp16"):
            X, it, rr, ms, ms_mv = sd.refine(e + 1, prec, B, tol=1e-10, maxit=60)
            res = np.linalg.norm(B - A @ X) / np.linalg.norm(B)
            err = rel_l2(X, X_ref)
            print(case, e, prec, "corrections", it, "relres", rr, res, "err", err, flush=True)
            assert rr <= 1e-10 and res <= 2e-10
            assert err < 1e-8
            its[prec] = it
        assert its["fp64"] <= 2 and its["fp16"] >= its["fp32"]
    sd.close()


def test_lsqr_batch_matches_oracle_per_system():
    """Systems of different sizes and kinds in one launch: each agrees with
    the oracle's slvlsqr2 as the single-system solver does."""
    from tritok_amd import solvers as S
    systems = [_vte_system(load_golden("dc2")), _vte_system(load_golden("jit2"), e=1)]
    for m, n, k, seed in [(3000, 1000, 11, 3), (400, 120, 7, 4), (2500, 1200, 9, 6)]:
        ia, ja, a = _random_csr(m, n, k, seed)
        systems.append((ia, ja, a, n))
    for ia, ja, a, n in systems:
        assert len(np.unique(ja)) == n  # every unknown in some equation
    rng = np.random.default_rng(9)
    bs = [rng.standard_normal(len(s[0]) - 1) for s in systems]
    x0s = [0.1 * rng.standard_normal(s[3]) for s in systems]
    B = S.LsqrBatch(systems)
    for eps in (1e-7, 1e-10):
        xs, istop, nit, ms = B.solve(bs, x0s, 5000, eps)
        print("batch eps", eps, "device ms", ms, flush=True)
        for k, (ia, ja, a, n) in enumerate(systems):
            m = len(ia) - 1
            ao = np.array(a, np.float64)
            q = O.prelsqr2(m, n, ia, ja, ao)
            assert np.max(np.abs(B.q[B.noff[k]:B.noff[k + 1]] - q) / q) <= 5e-16
            x_ref, istop_ref, nit_ref = O.slvlsqr2(m, n, ia, ja, ao, q, bs[k], x0s[k], 5000, eps)
            print(k, "nit", nit[k], nit_ref, "istop", istop[k], istop_ref, "rel", rel_l2(xs[k], x_ref), flush=True)
            assert istop[k] == istop_ref and abs(nit[k] - nit_ref) <= 2
            assert rel_l2(xs[k], x_ref) < (1e-8 if eps < 1e-9 else 1e-5)
    # repeatable
    xs2, _, nit2, _ = B.solve(bs, x0s, 5000, 1e-10)
    assert all(np.array_equal(u, w) for u, w in zip(xs, xs2)) and np.array_equal(nit, nit2)
    # every system on a gang of several workgroups (barriers and sums across
    # workgroups): the same solutions to the same bar, repeatable
    for width in (3, 8):
        assert B.set_width(width) == width
        xg, istop_g, nit_g, ms = B.solve(bs, x0s, 5000, 1e-10)
        print("gang width", width, "device ms", ms, "nit", nit_g, flush=True)
        for k, (ia, ja, a, n) in enumerate(systems):
            assert istop_g[k] == istop[k] and abs(nit_g[k] - nit[k]) <= 2
            assert rel_l2(xg[k], xs[k]) < 1e-8
        xg2, _, nit_g2, _ = B.solve(bs, x0s, 5000, 1e-10)
        assert all(np.array_equal(u, w) for u, w in zip(xg, xg2)) and np.array_equal(nit_g, nit_g2)
    assert B.set_width(0) >= 1
    B.close()


def test_lsqr_batch_fp32_values_match_oracle_on_rounded_values():
    """FP32 values in the batched products: from x0 = 0 the hot-start
    residual is b in both, so each solve is the oracle's slvlsqr2 on the
    scaled values rounded to FP32 (the same iteration to the order of the
    sums: counts within 2, solutions within 1e-8 at eps 1e-10); against the
    FP64 solve the difference is the rounding of the values (reported,
    within 1e-4)."""
    from tritok_amd import solvers as S
    systems = [_vte_system(load_golden("dc2")), _vte_system(load_golden("jit2"), e=1)]
    for m, n, k, seed in [(3000, 1000, 11, 3), (2500, 1200, 9, 6)]:
        ia, ja, a = _random_csr(m, n, k, seed)
        systems.append((ia, ja, a, n))
    rng = np.random.default_rng(19)
    bs = [rng.standard_normal(len(s[0]) - 1) for s in systems]
    x0s = [np.zeros(s[3]) for s in systems]
    B = S.LsqrBatch(systems)
    x64, _, _, _ = B.solve(bs, x0s, 5000, 1e-10)
    B.set_precision("fp32")
    for eps in (1e-7, 1e-10):
        xs, istop, nit, ms = B.solve(bs, x0s, 5000, eps)
        print("fp32 batch eps", eps, "device ms", ms, flush=True)
        for k, (ia, ja, a, n) in enumerate(systems):
            m = len(ia) - 1
            ao = np.array(a, np.float64)
            q = O.prelsqr2(m, n, ia, ja, ao)
            a32 = ao.astype(np.float32).astype(np.float64)
            x_ref, istop_ref, nit_ref = O.slvlsqr2(m, n, ia, ja, a32, q, bs[k], x0s[k], 5000, eps)
            e = rel_l2(xs[k], x_ref)
            print(k, "nit", nit[k], nit_ref, "istop", istop[k], istop_ref, "rel vs rounded-value oracle", e,
                  "vs fp64 batch", rel_l2(xs[k], x64[k]), flush=True)
            assert istop[k] == istop_ref and abs(nit[k] - nit_ref) <= 2
            assert e < (1e-8 if eps < 1e-9 else 1e-5)
            if eps < 1e-9:
                assert rel_l2(xs[k], x64[k]) < 1e-4
    B.set_precision("fp64")
    x64b, _, _, _ = B.solve(bs, x0s, 5000, 1e-10)
    assert all(np.array_equal(u, w) for u, w in zip(x64, x64b))
    B.close()


def test_lsqr_batch_gang_width_tuner_keeps_oracle_parity():
    from tritok_amd import solvers as S
    systems = [_vte_system(load_golden("dc2")), _vte_system(load_golden("jit2"), e=1)]
    for m, n, k, seed in [(3000, 1000, 11, 3), (400, 120, 7, 4)]:
        ia, ja, a = _random_csr(m, n, k, seed)
        systems.append((ia, ja, a, n))
    rng = np.random.default_rng(29)
    bs = [rng.standard_normal(len(s[0]) - 1) for s in systems]
    x0s = [0.1 * rng.standard_normal(s[3]) for s in systems]
    B = S.LsqrBatch(systems)
    scale, ms = B.tune(bs, x0s, 5000, 1e-10)
    print("gang width scale", scale, ms, flush=True)
    assert scale in ms and all(t > 0 for t in ms.values())
    xs, istop, nit, _ = B.solve(bs, x0s, 5000, 1e-10)
    for k, (ia, ja, a, n) in enumerate(systems):
        m = len(ia) - 1
        ao = np.array(a, np.float64)
        q = O.prelsqr2(m, n, ia, ja, ao)
        x_ref, istop_ref, nit_ref = O.slvlsqr2(m, n, ia, ja, ao, q, bs[k], x0s[k], 5000, 1e-10)
        assert istop[k] == istop_ref and abs(nit[k] - nit_ref) <= 2
        assert rel_l2(xs[k], x_ref) < 1e-8
    B.close()


def test_block_batch_matches_oracle_per_system():
    """Subdomain systems of different meshes and components solved on the
    cell-block format in groups: each agrees with the oracle's slvlsqr2 as the
    single-system solvers do, and repeats bit for bit; FP32 and FP16 blocks
    stay close to the FP64 solutions."""
    from tritok_amd import solvers as S
    systems = [_vte_system(load_golden("dc2")), _vte_system(load_golden("jit2"), e=1),
               _vte_system(load_golden("dc2"), e=2), _vte_system(load_golden("nc2"), e=0)]
    # more systems than one launch group holds (16): the groups are formed by size
    systems = systems * 5
    rng = np.random.default_rng(11)
    bs = [rng.standard_normal(len(s[0]) - 1) for s in systems]
    x0s = [0.1 * rng.standard_normal(s[3]) for s in systems]
    B = S.BlockBatch(systems, rows_per_cell=51)
    for eps in (1e-7, 1e-10):
        xs, istop, nit, ms = B.solve(bs, x0s, 5000, eps)
        print("block batch eps", eps, "device ms", ms, flush=True)
        for k, (ia, ja, a, n) in enumerate(systems[:4]):
            m = len(ia) - 1
            ao = np.array(a, np.float64)
            q = O.prelsqr2(m, n, ia, ja, ao)
            x_ref, istop_ref, nit_ref = O.slvlsqr2(m, n, ia, ja, ao, q, bs[k], x0s[k], 5000, eps)
            print(k, "nit", nit[k], nit_ref, "istop", istop[k], istop_ref, "rel", rel_l2(xs[k], x_ref), flush=True)
            assert istop[k] == istop_ref and abs(nit[k] - nit_ref) <= 2
            assert rel_l2(xs[k], x_ref) < (1e-8 if eps < 1e-9 else 1e-5)
    xs2, _, nit2, _ = B.solve(bs, x0s, 5000, 1e-10)
    assert all(np.array_equal(u, w) for u, w in zip(xs, xs2)) and np.array_equal(nit, nit2)
    for prec, tol in (("fp32", 1e-4), ("fp16", 3e-2)):
        B.set_precision(prec)
        xp, istop_p, nit_p, ms = B.solve(bs, x0s, 5000, 1e-10)
        err = max(rel_l2(u, w) for u, w in zip(xp, xs))
        print(prec, "device ms", ms, "max rel vs FP64", err, "nit", nit_p.min(), nit_p.max(), flush=True)
        assert np.all(istop_p > 0) and err < tol
    B.close()


def test_block_batch_rejects_systems_without_cell_structure():
    from tritok_amd import solvers as S
    from tritok_amd._lib import TTError
    ia, ja, a = _random_csr(510, 300, 9, 5)  # rows spread over 300 columns: no 51-row cell structure
    with pytest.raises(TTError):
        S.BlockBatch([(ia, ja, a, 300)], rows_per_cell=51)
