"""A Fortran host driving the library through the ISO_C_BINDING module
(tritok_amd/fortran/tt_fortran_driver, built with the library):

  * prelsqr2_gpu / slvlsqr2_gpu, called with the reference's argument lists,
    against the oracle's prelsqr2 / slvlsqr2 (bars as for the C entry points);
  * the whole steady loop from a .geo / .bic / tri.inp, with the restart and
    history files written through the library: the restart's fields within
    1e-8 of the reference's converged fields (dc2t)."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, REPO, load_golden, rel_l2

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(REPO, "tritok_amd", "fortran", "tt_fortran_driver")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail("tt_fortran_driver not built (tritok_amd/build.py builds it with amdflang)")
    return DRIVER


def test_fortran_caller_lsqr_matches_oracle(tmp_path):
    from test_gpu_solvers import _vte_system
    ia, ja, a, n = _vte_system(load_golden("jit2"), e=1)
    m, nnz = len(ia) - 1, len(ja)
    rng = np.random.default_rng(41)
    b = rng.standard_normal(m)
    x0 = 0.1 * rng.standard_normal(n)
    eps = np.float32(1e-10)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(struct.pack("=iiiif", m, n, nnz, 5000, eps))
        f.write(ia.astype("<i4").tobytes() + ja.astype("<i4").tobytes())
        f.write(np.asarray(a, "<f8").tobytes() + b.astype("<f8").tobytes() + x0.astype("<f8").tobytes())
    print("fortran lsqr m=%d n=%d nnz=%d" % (m, n, nnz), flush=True)
    subprocess.run([_driver(), "lsqr", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True,
                   timeout=120)
    out = open(tmp_path / "out.bin", "rb").read()
    q = np.frombuffer(out, "<f8", n, 0)
    a_s = np.frombuffer(out, "<f8", nnz, 8 * n)
    x = np.frombuffer(out, "<f8", n, 8 * (n + nnz))
    nit, ierr = np.frombuffer(out, "<i4", 2, 8 * (2 * n + nnz))
    ao = np.array(a, np.float64)
    q_ref = O.prelsqr2(m, n, ia, ja, ao)
    x_ref, istop_ref, nit_ref = O.slvlsqr2(m, n, ia, ja, ao, q_ref, b, x0, 5000, float(eps))
    print("nit", nit, nit_ref, "ierr", ierr, "rel", rel_l2(x, x_ref), flush=True)
    assert np.max(np.abs(q - q_ref) / q_ref) <= 5e-16
    assert np.max(np.abs(a_s - ao) / np.maximum(np.abs(ao), 1e-300)) <= 1e-15
    assert ierr == 0 and istop_ref <= 3 and abs(int(nit) - nit_ref) <= 2
    assert rel_l2(x, x_ref) < 1e-8


def test_fortran_caller_runs_the_case_and_writes_the_reference_restart(tmp_path):
    from tritok_amd import files as F
    sys.path.insert(0, GOLDEN)
    import make_golden as MG
    g = load_golden("dc2t")
    geo = dict(x=g["mesh_x"], ibc=g["mesh_ibc"], idc=g["mesh_idc"], bewn=g["mesh_bewn"],
               nbnodes=int(g["mesh_nbnodes"]))
    MG.write_geo(str(tmp_path / "mesh.geo"), geo, title="dc2t")
    (tmp_path / "case.bic").write_bytes(g["bic_text"].tobytes())
    (tmp_path / "tri.inp").write_bytes(g["inp_text"].tobytes())
    print("fortran case dc2t", flush=True)
    r = subprocess.run([_driver(), "case", str(tmp_path / "mesh.geo"), str(tmp_path / "case.bic"),
                        str(tmp_path / "tri.inp"), str(tmp_path)], check=True, timeout=300, capture_output=True,
                       text=True)
    print(r.stdout.strip(), flush=True)
    nn, nq = len(g["mesh_x"]), len(g["rst_qvorticity"])
    fields, ver, ts, gl, rt = F.read_restart(tmp_path / "tri.rst", nn, nq)
    assert ver == 4 and ts == 1 and abs(gl - len(g["sta"])) <= 3
    for k in ("velocity", "vorticity", "qvorticity"):
        e = rel_l2(fields[k].reshape(-1), g["rst_" + k].reshape(-1))
        print("fortran case", k, e, flush=True)
        assert e < 1e-8, (k, e)
    sta = (tmp_path / "tri.sta").read_text().splitlines()
    assert len(sta) == gl + 2 and sta[1].startswith(" tstep  tsite      vx")
