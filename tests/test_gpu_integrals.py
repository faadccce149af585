"""GPU integral assembly against the oracle and the reference fixtures.

Parity bar: FP64 integral matrices within 1e-12 relative (Frobenius per matrix)
of the oracle (which equals the reference bit for bit with the same Gauss
table), on uniform, jittered and 8^3 meshes; for the modified-Helmholtz
matrices also per cell, since mu spans three decades across a batch."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _host(g):
    from tritok_amd import Host
    return Host.from_arrays(g["mesh_x"], g["mesh_ibc"], g["mesh_idc"], g["mesh_bewn"], int(g["mesh_nbnodes"]))


@pytest.mark.parametrize("case", ["dc2", "jit2"])
def test_macro_integrals_gpu_vs_reference(case, gauss_table):
    from tritok_amd import integrals as I
    g = load_golden(case)
    h = _host(g)
    r = I.macro_integrals(h.view(), I.quad(4, 4, 4, 4, 2, table=gauss_table))
    for k in I.MACRO:
        e = rel_l2(r[k], g["macro_" + k])
        print(case, k, "rel", e, flush=True)
        assert e < TOL, k
    assert np.allclose(r["ont"], g["macro_ont"], rtol=1e-3, atol=1e-12)


@pytest.mark.parametrize("case", ["dc2", "jit2"])
def test_sd_integrals_gpu_vs_reference(case, gauss_table):
    from tritok_amd import integrals as I
    g = load_golden(case)
    h = _host(g)
    r = I.sd_integrals(h.view(), I.quad(4, 4, 4, 4, 2, table=gauss_table))
    for k in I.SDKM:
        e = rel_l2(r[k], g["sd_" + k])
        print(case, k, "rel", e, flush=True)
        assert e < TOL, k


def test_sd_integrals_gpu_dc8_rows(gauss_table):
    from tritok_amd import integrals as I
    g = load_golden("dc8")
    h = _host(g)
    q = I.quad(4, 4, 4, 4, 2, table=gauss_table)
    for k, row in enumerate(g["sd_rows"]):
        r = I.sd_integrals(h.view(), q, int(row), int(row) + 1)
        for name in I.SDKM:
            assert rel_l2(r[name][0], g["sd_" + name][k]) < TOL, (row, name)


def test_macro_gpu_vs_oracle_copy_and_computed_gauss():
    """COPY=1 replication and the library's own Gauss values (no table)."""
    from tritok_amd import integrals as I
    from tritok_amd import Host
    h = Host.cube(3)
    m = h.mesh_arrays()
    for copy in (0, 1):
        r = I.macro_integrals(h.view(), I.quad(4, 4, 4, 4, 2), copy=copy)
        o = O.macro_integrals(O.quad(4, 4, 4, 4, 2), m, copy=copy)
        for k in I.MACRO:
            assert rel_l2(r[k], o[k]) < TOL, (copy, k)


@pytest.mark.parametrize("case", ["mh2u", "mh2v", "mh2h"])
def test_macro_mh_gpu_vs_reference(case, gauss_table):
    """FMATmodHelm on the device against the reference's tri.modHelm.macro.int."""
    from tritok_amd import integrals as I
    g = load_golden(case)
    h = _host(g)
    r = I.macro_integrals_mh(h.view(), I.quad(4, 4, 4, 4, 2, table=gauss_table), g["mh_diff"], g["mh_qdiff"],
                             float(g["mh_beta"]))
    for k in I.MACRO:
        e = rel_l2(r[k], g["mh_" + k])
        print(case, k, "rel", e, flush=True)
        assert e < TOL, k
    assert np.allclose(r["ont"], g["mh_ont"], rtol=1e-3, atol=1e-12)
    assert r["mumax"] == g["mh_mumax"]


def test_macro_mh_gpu_batch_vs_oracle():
    """Independent jittered cells with mu log-uniform in [0.1, 100] per cell
    (the batched-integrals workload, 27 cells): every cell block of every
    matrix within 1e-12 of the oracle."""
    from tritok_amd import integrals as I
    from tritok_amd import workloads as W
    b = W.cell_batch(3, jitter=0.1, seed=7, mu_range=(0.1, 100.0))
    am = I.ArrayMesh(b["arrays"])
    r = I.macro_integrals_mh(am, I.quad(4, 4, 4, 4, 2), b["diff"], b["qdiff"], b["beta"])
    a = dict(b["arrays"], nn=am.mesh.nnodes, nc=am.mesh.nicell, nq=am.mesh.nq)
    o = O.macro_integrals_mh(O.quad(4, 4, 4, 4, 2), a, b["diff"], b["qdiff"], b["beta"])
    nc = am.mesh.nicell
    for k in I.MACRO:
        per = r[k].shape[0] // nc
        worst = max(rel_l2(r[k][c * per:(c + 1) * per], o[k][c * per:(c + 1) * per]) for c in range(nc))
        print("batch", k, "worst cell rel", worst, flush=True)
        assert worst < TOL, k
    assert r["mumax"] == o["mumax"]


def test_macro_mh_gpu_mu_zero_is_laplace(gauss_table):
    """beta = 0 (mu = 0): the modified-Helmholtz kernel reduces to Laplace."""
    from tritok_amd import integrals as I
    g = load_golden("jit2")
    h = _host(g)
    m = h.mesh_arrays()
    r = I.macro_integrals_mh(h.view(), I.quad(4, 4, 4, 4, 2, table=gauss_table), np.ones(m["nn"]),
                             np.ones(m["nq"]), 0.0)
    for k in I.MACRO:
        assert rel_l2(r[k], g["macro_" + k]) < TOL, k


def test_macro_kernel_evaluation_counts_match_oracle():
    """The device's count of fundamental-solution evaluations (the unit of the
    integral roofline) equals the oracle's replay of the subdivision rules."""
    from tritok_amd import integrals as I
    from tritok_amd import workloads as W
    b = W.cell_batch(2, jitter=0.1, seed=3, mu_range=(0.5, 50.0))
    am = I.ArrayMesh(b["arrays"])
    I.macro_integrals_mh(am, I.quad(4, 4, 4, 4, 2), b["diff"], b["qdiff"], b["beta"])
    st = I.last_integral_stats()
    a = dict(b["arrays"], nn=am.mesh.nnodes, nc=am.mesh.nicell, nq=am.mesh.nq)
    O.macro_integrals_mh(O.quad(4, 4, 4, 4, 2), a, b["diff"], b["qdiff"], b["beta"])
    print("device", st, "oracle", O.last_points(), flush=True)
    assert (st["domain_points"], st["boundary_points"]) == O.last_points()
    assert st["ms"] > 0
