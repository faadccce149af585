"""The oracle's single-domain modified-Helmholtz path (or_mhsd.c) against the
reference's own runs (tests/golden/mhsd*.npz, EQNS imH 3, 5, 6):

* the integrals with corrections and B' (FMATsdmHfs + MakeBprime) against the
  reference's tri.modHelm.full.int, entry by entry;
* the whole time loop (fmmHSysM, dgefa, fmSolveModHelmDC / DCwawt / DCaca,
  dgesl, relaxation; modHelm_SetProblem, setGRAD and Int2MacroNodes from the
  package's host side) against the reference's mH field after each time step
  (tri.dat, 9 significant digits) and its RMS error (tri.mH-rms.dat).
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden


def mhsd_host(g):
    from tritok_amd import Host
    h = Host.from_arrays(g["mesh_x"], g["mesh_ibc"], g["mesh_idc"], g["mesh_bewn"], int(g["mesh_nbnodes"]))
    return h.read_bic(g["bic_text"].tobytes().decode())


def qdiff_sequential(m, diff):
    """Int2MacroNodes (lib_util.f:1377-1475) with the oracle's cshape27 and the
    reference's summation order"""
    from tritok_amd import mhsd
    xsi = [-.75, -.75, .75, .75, -.75, .75, .75, -.75, 1, 1, 1, 1, .75, -.75, -.75, .75, -1, -1, -1, -1, -.75, .75,
           .75, -.75]
    eta = [-.75, .75, .75, -.75, -1, -1, -1, -1, -.75, .75, .75, -.75, 1, 1, 1, 1, .75, -.75, -.75, .75, -.75, -.75,
           .75, .75]
    ceta = [-1, -1, -1, -1, -.75, -.75, .75, .75, -.75, -.75, .75, .75, -.75, -.75, .75, .75, -.75, -.75, .75, .75, 1,
            1, 1, 1]
    fis = [O.shape27(xsi[k], eta[k], ceta[k]) for k in range(24)]
    q = np.zeros(m["nq"])
    for ic in range(m["nc"]):
        for ii in range(6):
            for jj in range(4):
                fi = fis[ii * 4 + jj]
                acc = 0.0
                for kk in range(27):
                    acc = acc + fi[kk] * diff[m["idc"][ic, kk] - 1]
                q[m["ibf"][ic, ii, jj] - 1] = acc
    np.testing.assert_allclose(q, mhsd.int2macro(m, diff), rtol=1e-14, atol=1e-15)
    return q


def setup_case(g):
    from tritok_amd import mhsd
    host = mhsd_host(g)
    m = host.mesh_arrays()
    cd = mhsd.codes(host)
    tstep = float(g["tstep"])
    beta, beta2, beta3 = mhsd.time_scheme(tstep)
    p = mhsd.set_problem(m, cd, int(g["mhf"]), 0.0, np.zeros(m["nn"]))
    diff = p["diff"]
    qdiff = qdiff_sequential(m, diff)
    return host, m, cd, beta, diff, qdiff


def test_codes_number_unknowns_like_the_reference():
    from tritok_amd import mhsd
    g = load_golden("mhsd2")
    host = mhsd_host(g)
    cd = mhsd.codes(host)
    m = host.mesh_arrays()
    kode, qkode = cd["kode"], cd["qkode"]
    # every node and flux point numbered once, unknowns nodes first then flux points
    unk = np.concatenate([-kode[kode < 0], -qkode[qkode < 0]])
    known = np.concatenate([kode[kode > 0], qkode[qkode > 0]])
    assert np.array_equal(unk, np.arange(1, cd["nunk"] + 1))
    assert np.array_equal(np.sort(known), np.arange(1, cd["nknown"] + 1))
    # interior nodes are unknown; walls 1, 3, 5 prescribe u, walls 2, 4, 6 q
    assert np.all(kode[m["lbn"] == 0] < 0)
    walls = cd["bewn"][np.arange(4 * m["ne"]) // 4]
    assert np.all((qkode > 0) == np.isin(walls, (2, 4, 6)))
    iwn_node = np.zeros(m["nn"], np.int32)
    iwn_node[m["gbn"] - 1] = cd["iwn"]
    assert np.all((kode > 0) == np.isin(iwn_node, (1, 3, 5)))
    # flux-node ids: the element's four flux points
    assert cd["ibcf"].min() >= 1 and cd["ibcf"].max() <= m["nq"] and len(np.unique(cd["ibcf"])) == 4 * m["ne"]


def test_integrals_match_reference_file(gauss_table):
    g = load_golden("mhsd2")
    host, m, cd, beta, diff, qdiff = setup_case(g)
    assert beta == g["int_beta"]
    s = 0.0
    for v in diff:
        s = s + v
    assert s == g["int_sumdiff"]
    q = O.quad(table=gauss_table)
    mats = O.mhsd_integrals(q, m, cd["ibcf"], diff, qdiff, beta)
    assert mats["mumax"] == g["int_mumax"]
    np.testing.assert_allclose(mats["ont"], g["int_ont"], rtol=1e-9, atol=1e-16)
    worst = 0.0
    for pre in ("U", "Q"):
        for k in O.MHSD_NAMES:
            ref = g["mhsd_" + pre + k]
            got = mats[pre][k]
            assert got.shape == ref.shape, (pre, k)
            d = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
            worst = max(worst, d)
            assert d <= 1e-13, (pre, k, d)
    print("oracle vs reference integrals: max rel", worst)


def _compressed_products(kind, g, mats, m):
    """per-matrix products of the reference's compressed forms, built by the
    oracle's ACA / wavelet restatements (pinned by compress_ref.npz)"""
    extra = g["inp_text"].tobytes().decode()
    prods = {}
    for pre in ("U", "Q"):
        for k in ("B", "Bp", "Dx", "Dy", "Dz", "Abdx", "Abdy", "Abdz"):
            A = mats[pre][k]
            if kind == "aca":
                line = [ln for ln in extra.splitlines() if ln.startswith("ACAC")][0].split()
                a, b, r, _, _ = O.aca(int(line[1]), A, float(line[2]), vec=np.full(A.shape[1], 0.5))
                prods[pre, k] = ("aca", a, b)
            else:
                line = [ln for ln in extra.splitlines() if ln.startswith("WAWT")][0].split()
                prods[pre, k] = ("wt", O.wt_compress(A, float(line[1])), A.shape[0])
    return prods


def _apply(p, x):
    if p[0] == "aca":
        return O.aca_abxv(p[1], p[2], x)
    return O.wt_wtawx(p[1], x, p[2])


def run_oracle(g, gauss_table, kind="dense"):
    from tritok_amd import mhsd
    host, m, cd, beta, diff, qdiff = setup_case(g)
    _, beta2, beta3 = mhsd.time_scheme(float(g["tstep"]))
    q = O.quad(table=gauss_table)
    mats = O.mhsd_integrals(q, m, cd["ibcf"], diff, qdiff, beta)
    nunk, nknown = cd["nunk"], cd["nknown"]
    A, RHS = O.mhsd_sysm(m, cd["ibcf"], cd["kode"], cd["qkode"], nunk, nknown, diff, qdiff, mats)
    assert not np.any(A == -99999.999) and not np.any(RHS[:, :nknown] == -99999.999)
    lu, piv, info = O.dgefa(A)
    assert info == 0
    prods = _compressed_products(kind, g, mats, m) if kind != "dense" else None
    dgrad = mhsd.set_grad(m, diff)
    f = int(g["mhf"])
    u = cd["u"].copy()
    qu = cd["q"].copy()
    p = mhsd.set_problem(m, cd, f, 0.0, np.zeros(m["nn"]))
    u[cd["kode"] > 0] = p["u_known"][cd["kode"] > 0]
    qu[cd["qkode"] > 0] = p["q_known"][cd["qkode"] > 0]
    pts, ppts = u.copy(), u.copy()
    t, steps = 0.0, []
    for step in range(int(g["nstep"])):
        t = t + float(g["tstep"])
        ppts, pts = pts, u.copy()
        mhtime = -pts * beta2 + ppts * beta3
        for it in range(int(g["nnlit"])):
            p = mhsd.set_problem(m, cd, f, t, mhtime)
            u[cd["kode"] > 0] = p["u_known"][cd["kode"] > 0]
            qu[cd["qkode"] > 0] = p["q_known"][cd["qkode"] > 0]
            if kind == "dense":
                r = O.mhsd_rhs(m, cd["ibcf"], cd["kode"], cd["qkode"], nunk, nknown, diff, qdiff, beta, u, qu, p["vel"],
                               p["rhs"], dgrad, mats, RHS)
            else:
                v = O.mhsd_vectors(m, u, p["vel"], p["rhs"], dgrad)
                names = ("B", "Bp", "Dx", "Dy", "Dz", "Abdx", "Abdy", "Abdz")
                yu = [_apply(prods["U", k], v[i]) for i, k in enumerate(names)]
                yq = [_apply(prods["Q", k], v[i]) for i, k in enumerate(names)]
                r = O.mhsd_rhs_mv(m, cd["kode"], cd["qkode"], nunk, nknown, u, qu, RHS, yu, yq)
            x = O.dgesl(lu, piv, r) if np.linalg.norm(r) != 0 else np.zeros(nunk)
            u[cd["kode"] < 0] = x[-cd["kode"][cd["kode"] < 0] - 1]
            qu[cd["qkode"] < 0] = x[-cd["qkode"][cd["qkode"] < 0] - 1]
        steps.append((u.copy(), mhsd.rms_error(m, u, p["anal"])))
    return steps


@pytest.mark.parametrize("name,kind", [("mhsd2", "dense"), ("mhsd2a", "aca"), ("mhsd2w", "wt"),
                                       pytest.param("mhsd4", "dense", marks=pytest.mark.slow),
                                       pytest.param("mhsd4a", "aca", marks=pytest.mark.slow),
                                       pytest.param("mhsd4w", "wt", marks=pytest.mark.slow)])
def test_time_loop_matches_reference_field(name, kind, gauss_table):
    g = load_golden(name)
    steps = run_oracle(g, gauss_table, kind)
    for k, (u, rms) in enumerate(steps):
        ref = g["mh_u"][k]
        d = np.abs(u - ref).max() / np.abs(ref).max()
        print(name, "step", k + 1, "max rel vs tri.dat", d, "rms", rms, g["mh_rms"][k, 1])
        assert d < 2e-8, (k, d)  # tri.dat holds 9 significant digits
        assert abs(rms - g["mh_rms"][k, 1]) <= 2e-9 * g["mh_rms"][k, 1], (rms, g["mh_rms"][k])


def test_compression_changes_the_reference_result():
    """the compressed runs differ from the dense one by more than tri.dat's
    rounding, so the compressed-path comparisons pin the compressed products"""
    d, a = load_golden("mhsd4"), load_golden("mhsd4a")
    assert np.abs(d["mh_u"] - a["mh_u"]).max() / np.abs(d["mh_u"]).max() > 1e-6


def test_integral_file_writer_reproduces_reference_bytes(tmp_path):
    """tt_write_mhsd_int on the reference's own matrices writes the reference's
    tri.modHelm.full.int byte for byte; tt_read_mhsd_int reads it back."""
    import ctypes as C
    import hashlib

    from tritok_amd import _lib
    from tritok_amd.mhsd import _write_int
    g = load_golden("mhsd2")
    nn, nb = (int(v) for v in g["int_nn_nb"])
    U = [np.asfortranarray(g["mhsd_U" + k]) for k in O.MHSD_NAMES]
    Q = [np.asfortranarray(g["mhsd_Q" + k]) for k in O.MHSD_NAMES]
    nf = Q[0].shape[0]
    ont = np.ascontiguousarray(g["int_ont"])
    PD = _lib.PD
    path = str(tmp_path / "tri.modHelm.full.int")
    _lib.check(_write_int(path.encode(), int(g["int_version"]), nn, nb, nf, float(g["int_beta"]),
                          float(g["int_sumdiff"]), float(g["int_mumax"]), (PD * 10)(*[a.ctypes.data_as(PD) for a in U]),
                          (PD * 10)(*[a.ctypes.data_as(PD) for a in Q]), ont.ctypes.data_as(PD)), "write")
    raw = open(path, "rb").read()
    assert len(raw) == int(g["int_bytes"])
    assert hashlib.sha256(raw).digest() == g["int_sha256"].tobytes()
    rd = _lib._sig("tt_read_mhsd_int", _lib.i32, C.c_char_p, _lib.i32, _lib.PI, PD, C.POINTER(PD), C.POINTER(PD), PD)
    U2 = [np.zeros_like(a, order="F") for a in U]
    Q2 = [np.zeros_like(a, order="F") for a in Q]
    hdr, hd, o2 = np.zeros(3, np.int32), np.zeros(3), np.zeros(5)
    _lib.check(rd(path.encode(), nf, hdr.ctypes.data_as(_lib.PI), hd.ctypes.data_as(PD),
                  (PD * 10)(*[a.ctypes.data_as(PD) for a in U2]), (PD * 10)(*[a.ctypes.data_as(PD) for a in Q2]),
                  o2.ctypes.data_as(PD)), "read")
    assert list(hdr) == [int(g["int_version"]), nn, nb]
    assert list(hd) == [g["int_beta"], g["int_sumdiff"], g["int_mumax"]]
    assert all(np.array_equal(a, b) for a, b in zip(U + Q, U2 + Q2)) and np.array_equal(o2, ont)
