"""Pins the CPU oracle (oracle/) before anything is compared against it.

* against the reference binary's own outputs (tests/golden/*.npz, written by
  tests/golden/make_golden.py from oracle/_ref/Tritok): integral matrices,
  per-iteration LSQR counts and converged fields must agree exactly when the
  reference's Gauss table is used;
* against mathematics that fixes a result independently of both codes:
  Gauss-Legendre exactness, shape-function identities, rigid-body and
  linear-field identities of the integrals, LSQR/LU against scipy.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg

import oracle as O
from conftest import host_for, load_golden, rel_l2

MACRO = ["H", "G", "B", "Abdx", "Abdy", "Abdz", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]
SDKM = ["H", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("rule", range(1, 10))
def test_gauss_matches_closed_form(rule):
    gi, om = O.gauss(rule)
    x, w = np.polynomial.legendre.leggauss(len(gi))
    assert np.abs(gi - x).max() < 1e-15 and np.abs(om - w).max() < 1e-14
    # exact for polynomials of degree 2N-1; the nodes are the roots of P_N, so
    # the rule gives 0 for int P_N^2 = 2/(2N+1) (degree 2N is not integrated)
    n = len(gi)
    assert abs(np.sum(om * gi ** (2 * n - 2)) - 2.0 / (2 * n - 1)) < 1e-13
    pn = np.polynomial.legendre.legval(gi, [0] * n + [1])
    assert np.abs(pn).max() < 1e-12 and abs(np.sum(om) - 2.0) < 1e-14


def test_gauss_table_of_reference_is_gauss_legendre(gauss_table):
    for rule in range(1, 10):
        a, b = gauss_table["ng1"][rule - 1], gauss_table["ng2"][rule - 1]
        gi, om = O.gauss(rule)
        assert np.abs(gauss_table["gi"][a - 1:b] - gi).max() < 5e-15
        assert np.abs(gauss_table["ome"][a - 1:b] - om).max() < 5e-15


# ------------------------------------------------------------ shape functions
LC = np.array([[0, 0, 1, 1, 0, 0, 1, 1, 0, .5, 1, .5, 0, 0, 1, 1, 0, .5, 1, .5, .5, 0, .5, 1, .5, .5, .5],
               [1, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, .5, .5, .5, .5, 0, 0, 0, 0, 1, .5, .5, .5, .5, 0, .5],
               [1, 0, 0, 1, 1, 0, 0, 1, .5, 0, .5, 1, 1, 0, 0, 1, .5, 0, .5, 1, .5, .5, 0, .5, 1, .5, .5]]).T * 2 - 1


def test_shape27_kronecker_and_unity():
    for k in range(27):
        f = O.shape27(*LC[k])
        assert np.allclose(f, np.eye(27)[k], atol=1e-15)
    rng = np.random.default_rng(1)
    for p in rng.uniform(-1, 1, (20, 3)):
        assert abs(O.shape27(*p).sum() - 1.0) < 1e-14


def test_shape27_mass_matrix_integral():
    """sum_ij int phi_i phi_j over (-1,1)^3 = 8 (partition of unity), and the
    reference's TestIfifj value 21/4 (lib_int.f:3461) for the entry it checks
    is reproduced by tensor Gauss with 8 points."""
    gi, om = O.gauss(4)
    tot = 0.0
    for a, wa in zip(gi, om):
        for b, wb in zip(gi, om):
            for c, wc in zip(gi, om):
                f = O.shape27(a, b, c)
                tot += wa * wb * wc * f.sum() ** 2
    assert abs(tot - 8.0) < 1e-13


def test_shape9_and_flux_shapes():
    nodes9 = [(-1, -1), (0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (0, 0)]
    for k, (a, b) in enumerate(nodes9):
        assert np.allclose(O.shape9(a, b), np.eye(9)[k], atol=1e-15)
    t = 0.75
    for k, (a, b) in enumerate([(-t, -t), (t, -t), (t, t), (-t, t)]):
        assert np.allclose(O.dl34(a, b), np.eye(4)[k], atol=1e-15)


# ------------------------------------------------------- integrals vs reference
def _mesh(g):
    from tritok_amd import Host
    return Host.from_arrays(g["mesh_x"], g["mesh_ibc"], g["mesh_idc"], g["mesh_bewn"],
                            int(g["mesh_nbnodes"])).mesh_arrays()


@pytest.mark.parametrize("case", ["dc2", "jit2"])
def test_macro_integrals_match_reference(case, gauss_table):
    g = load_golden(case)
    m = _mesh(g)
    r = O.macro_integrals(O.quad(4, 4, 4, 4, 2, table=gauss_table), m)
    for k in MACRO:
        assert np.array_equal(r[k], g["macro_" + k]), k
    assert np.allclose(r["ont"], g["macro_ont"], rtol=1e-6, atol=1e-13)
    # with exact Gauss-Legendre values instead of the reference's typed table
    r2 = O.macro_integrals(O.quad(4, 4, 4, 4, 2), m)
    for k in MACRO:
        assert rel_l2(r2[k], g["macro_" + k]) < 1e-13, k


@pytest.mark.parametrize("case", ["dc2", "jit2"])
def test_sd_integrals_match_reference(case, gauss_table):
    g = load_golden(case)
    m = _mesh(g)
    r = O.sd_integrals(O.quad(4, 4, 4, 4, 2, table=gauss_table), m)
    for k in SDKM:
        assert np.array_equal(r[k], g["sd_" + k]), k


def test_dc8_rows_match_reference(gauss_table):
    """8^3 mesh: first/last cell blocks of the macro matrices and five sdkm rows."""
    g = load_golden("dc8")
    m = _mesh(g)
    q = O.quad(4, 4, 4, 4, 2, table=gauss_table)
    for row in g["sd_rows"]:
        r = O.sd_integrals(q, m, int(row), int(row) + 1)
        k = list(g["sd_rows"]).index(row)
        for name in SDKM:
            assert np.array_equal(r[name][0], g["sd_" + name][k]), (row, name)


def test_macro_rigid_body_and_linear_identities(gauss_table):
    """Independent of the reference: after the corrections, H rows sum to 0
    (rigid body), Ht rows of u-sources sum to 0, and sum_j H_ij x_j = sum_j D_ij
    for each coordinate (lib_fmat.f:359-387)."""
    g = load_golden("jit2")
    m = _mesh(g)
    r = O.macro_integrals(O.quad(4, 4, 4, 4, 2, table=gauss_table), m)
    nc = m["nc"]
    assert np.abs(r["H"].sum(1)).max() < 1e-13
    for k in ["Htx", "Hty", "Htz"]:
        assert np.abs(r[k].sum(1)).max() < 1e-13
    X = m["x"][m["idc"] - 1]  # (nc, 27, 3)
    rows_u = (np.arange(nc)[:, None] * 51 + np.arange(27)[None, :]).ravel()
    for d, D in [(0, "Dx"), (1, "Dy"), (2, "Dz")]:
        hx = np.einsum("cj,cj->c", r["H"][rows_u].reshape(nc * 27, 27), np.repeat(X[:, :, d], 27, axis=0))
        assert np.abs(hx - r[D][rows_u].sum(1)).max() < 1e-13
    # q-source rows: the free term is 1/2 (source on a face)
    G = r["G"]
    assert np.isfinite(G).all()


def test_domain_regular_integral_brute_force():
    """A regular domain integral against a 3x finer tensor Gauss rule on a
    distorted cell (independent quadrature)."""
    q = O.quad(4, 4, 4, 4, 2)
    c = np.array([[0, 0, 0], [1.1, 0, 0], [0, 0.9, 0], [1.0, 1.0, 0.1], [0, 0.1, 1], [1, 0, 1.2], [0.1, 1, 1],
                  [1, 1.1, 1]], float)
    p = np.array([2.3, -0.7, 0.4])
    be, dx, dy, dz, n = O.inted27(q, p, c, 0)
    gi, om = O.gauss(8)  # 48 points
    acc = np.zeros((4, 27))
    for k, wk in zip(gi, om):
        for j, wj in zip(gi, om):
            for i, wi in zip(gi, om):
                e = np.array([i, j, k])
                N = np.array([(1 + s0 * i) * (1 + s1 * j) * (1 + s2 * k) / 8 for s2 in (-1, 1) for s1 in (-1, 1)
                              for s0 in (-1, 1)])
                dN = np.array([[s0 * (1 + s1 * j) * (1 + s2 * k) / 8, s1 * (1 + s0 * i) * (1 + s2 * k) / 8,
                                s2 * (1 + s0 * i) * (1 + s1 * j) / 8] for s2 in (-1, 1) for s1 in (-1, 1)
                               for s0 in (-1, 1)])
                xc = N @ c
                J = dN.T @ c
                r = p - xc
                ra = np.linalg.norm(r)
                f = O.shape27(*e)
                w = wi * wj * wk * np.linalg.det(J)
                acc += w * np.outer([1 / (4 * np.pi * ra), *(r / (4 * np.pi * ra ** 3))], f)
    assert np.abs(be - acc[0]).max() < 1e-12 * np.abs(acc[0]).max() * 1e3
    assert np.abs(dx - acc[1]).max() < 1e-9 * np.abs(acc[1]).max()


# ----------------------------------------------------------- linear algebra
# ------------------------------------------------- modified-Helmholtz integrals
@pytest.mark.parametrize("case", ["mh2u", "mh2v", "mh2h"])
def test_macro_mh_match_reference(case, gauss_table):
    """FMATmodHelm against the reference's tri.modHelm.macro.int: uniform mu = 20
    (2^3 cube), mu = 1 + xyz and uniform mu = 100 (jittered 2^3 cube)."""
    g = load_golden(case)
    m = _mesh(g)
    r = O.macro_integrals_mh(O.quad(4, 4, 4, 4, 2, table=gauss_table), m, g["mh_diff"], g["mh_qdiff"],
                             float(g["mh_beta"]))
    for k in MACRO:
        assert np.array_equal(r[k], g["mh_" + k]), k
    assert np.array_equal(r["ont"], g["mh_ont"])
    assert r["mumax"] == g["mh_mumax"]


def test_macro_mh_mu_zero_is_laplace(gauss_table):
    """beta = 0 gives mu = 0 at every source: the modified-Helmholtz assembly
    (its other correction order included) must reproduce FMATkmdcSteady bit for bit."""
    g = load_golden("jit2")
    m = _mesh(g)
    q = O.quad(4, 4, 4, 4, 2, table=gauss_table)
    r = O.macro_integrals_mh(q, m, np.ones(m["nn"]), np.ones(m["nq"]), 0.0)
    for k in MACRO:
        assert np.array_equal(r[k], g["macro_" + k]), k
    assert r["mumax"] == 0.0


def test_domain_regular_mh_integral_brute_force():
    """A regular modified-Helmholtz domain integral against a 48^3 tensor Gauss
    rule written here from the fundamental solution exp(-mu r)/(4 pi r) and its
    gradient -(1 + mu r) exp(-mu r) r_vec/(4 pi r^3) (r_vec = x - p)."""
    q = O.quad(4, 4, 4, 4, 2)
    c = np.array([[0, 0, 0], [1.1, 0, 0], [0, 0.9, 0], [1.0, 1.0, 0.1], [0, 0.1, 1], [1, 0, 1.2], [0.1, 1, 1],
                  [1, 1.1, 1]], float)
    p = np.array([1.9, -0.4, 0.4])
    mu = 1.7
    be, dx, dy, dz, n = O.inted27_mh(q, p, c, 0, mu)
    gi, om = O.gauss(8)
    acc = np.zeros((4, 27))
    for k, wk in zip(gi, om):
        for j, wj in zip(gi, om):
            for i, wi in zip(gi, om):
                N = np.array([(1 + s0 * i) * (1 + s1 * j) * (1 + s2 * k) / 8 for s2 in (-1, 1) for s1 in (-1, 1)
                              for s0 in (-1, 1)])
                dN = np.array([[s0 * (1 + s1 * j) * (1 + s2 * k) / 8, s1 * (1 + s0 * i) * (1 + s2 * k) / 8,
                                s2 * (1 + s0 * i) * (1 + s1 * j) / 8] for s2 in (-1, 1) for s1 in (-1, 1)
                               for s0 in (-1, 1)])
                xc = N @ c
                rv = xc - p
                ra = np.linalg.norm(rv)
                us = np.exp(-mu * ra) / (4 * np.pi * ra)
                grad_x = -(1 + mu * ra) * np.exp(-mu * ra) * rv / (4 * np.pi * ra ** 3)  # d u*/d x (field point)
                w = wi * wj * wk * np.linalg.det(dN.T @ c)
                acc += w * np.outer([us, *grad_x], O.shape27(i, j, k))
    for got, want in zip((be, dx, dy, dz), acc):
        assert np.abs(got - want).max() < 1e-7 * np.abs(want).max()
    # mu enters: the Laplace values differ by far more than that tolerance
    be0 = O.inted27(q, p, c, 0)[0]
    assert np.abs(be0 - acc[0]).max() > 1e-2 * np.abs(acc[0]).max()


def _mh_field_residual(n, mu_kernel, mu_field=1.0):
    """Max over macro rows of |H u - G q| for u = exp(mu_field a.x), which solves
    lap u = mu_field^2 u: with the kernel's mu equal to mu_field the domain terms
    cancel and the boundary equation holds up to interpolation error."""
    from tritok_amd import Host
    m = Host.cube(n).mesh_arrays()
    x, xq, idc, ibf, flpm = m["x"], m["xq"], m["idc"], m["ibf"], m["flpm"]
    a = np.array([0.6, 0.0, 0.8])
    u = lambda pts: np.exp(mu_field * pts @ a)
    r = O.macro_integrals_mh(O.quad(4, 4, 4, 4, 2), m, np.full(m["nn"], 1.0 / mu_kernel ** 2),
                             np.full(m["nq"], 1.0 / mu_kernel ** 2), 1.0)
    worst = 0.0
    for ic in range(m["nc"]):
        un = u(x[idc[ic] - 1])
        cen = x[idc[ic] - 1].mean(0)
        qf = np.zeros(24)
        for s in range(6):
            pts = xq[ibf[ic, s] - 1]
            nrm = pts.mean(0) - cen
            nrm /= np.linalg.norm(nrm)
            qf[4 * s:4 * s + 4] = mu_field * (a @ nrm) * u(pts) * flpm[ic, s]
        rows = slice(51 * ic, 51 * ic + 51)
        worst = max(worst, np.abs(r["H"][rows] @ un - r["G"][rows] @ qf).max())
    return worst


def test_macro_mh_closed_form_field():
    """Closed-form modified-Helmholtz solution on the unit cube: the residual of
    the macro boundary equation falls at the interpolation order (h^3) under
    refinement, and a kernel with a 10 % wrong mu leaves a residual >10x larger."""
    r2, r4 = _mh_field_residual(2, 1.0), _mh_field_residual(4, 1.0)
    assert r2 < 3e-3 and r4 < 4e-4 and r2 / r4 > 5.0
    assert _mh_field_residual(4, 1.1) > 10 * r4


def test_dgefa_dgesl_vs_lapack():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((40, 40))
    b = rng.standard_normal(40)
    lu, piv, info = O.dgefa(a)
    assert info == 0
    x = O.dgesl(lu, piv, b)
    assert np.allclose(x, scipy.linalg.solve(a, b), rtol=1e-12, atol=1e-12)
    # same pivot sequence as LAPACK getrf (first maximum, partial pivoting)
    _, lp = scipy.linalg.lu_factor(a)
    assert np.array_equal(piv[:-1].astype(int) - 1, lp[:-1])


def test_lsqr_vs_scipy():
    rng = np.random.default_rng(4)
    A = scipy.sparse.random(300, 120, density=0.05, random_state=5, format="csr") + scipy.sparse.eye(300, 120)
    A = A.tocsr()
    b = rng.standard_normal(300)
    ia = (A.indptr + 1).astype(np.int32)
    ja = (A.indices + 1).astype(np.int32)
    a = A.data.copy()
    q = O.prelsqr2(300, 120, ia, ja, a)
    assert np.allclose(q, 1 / np.sqrt(np.asarray(A.multiply(A).sum(0)).ravel()))
    x, istop, nit = O.slvlsqr2(300, 120, ia, ja, a, q, b, np.zeros(120), 5000, np.float32(1e-12))
    ref = scipy.sparse.linalg.lsqr(A, b, atol=1e-14, btol=1e-14, iter_lim=5000)[0]
    assert rel_l2(x, ref) < 1e-9
    assert istop in (1, 2)


def test_dnrm2_scaling():
    x = np.array([3e200, 4e200, 0.0])
    assert O.dnrm2(x) == pytest.approx(5e200)
    assert O.dnrm2(np.array([3.0, 4.0])) == 5.0


# ----------------------------------------------------- Picard vs reference
@pytest.mark.parametrize("case", ["dc2", "nc2", "jit2", "dc2t"])
def test_picard_matches_reference_exactly(case, gauss_table):
    g = load_golden(case)
    h = host_for(g)
    c = O.case_from_view(h.view())
    c.quad = O.quad(4, 4, 4, 4, 2, table=gauss_table)
    f, sta, ite, _ = O.picard(c, h.fields())
    assert len(sta) == len(g["sta"])
    assert np.array_equal(ite[:, :10], g["ite"][:, 2:12])
    for k in ["velocity", "vorticity", "qvorticity", "temp", "qtemp"]:
        assert np.array_equal(f[k].reshape(-1), g["rst_" + k].reshape(-1)), k


@pytest.mark.slow
def test_picard_dc8_matches_reference_exactly(gauss_table):
    g = load_golden("dc8")
    h = host_for(g)
    c = O.case_from_view(h.view())
    c.quad = O.quad(4, 4, 4, 4, 2, table=gauss_table)
    f, sta, ite, _ = O.picard(c, h.fields())
    assert len(sta) == len(g["sta"]) == 96
    assert np.array_equal(ite[:, :10], g["ite"][:, 2:12])
    for k in ["velocity", "vorticity"]:
        assert np.array_equal(f[k].reshape(-1), g["rst_" + k].reshape(-1)), k
