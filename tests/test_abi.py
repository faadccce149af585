"""C-ABI boundary and host-side preprocessing on the CPU (no GPU calls).

* libtritok_amd.so loads and exports every function include/*.h declares;
* the cube generator and the .bic writer reproduce the reference's own inputs:
  on the reference mesh kva2x2x2.geo (fixture arrays), the generated
  lid-driven-cavity and heated-cube .bic texts give the same codes, unknown
  numbering and initial fields as mesh/3Dldc-X.bic and
  mesh/NaturalConvection-X.bic (fixture texts);
* tri.inp parsing and the unit cube's sizes (SURVEY table of n^3 cubes).
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from conftest import REPO, load_golden

DECL = re.compile(r"^\s*(?:int|void|int64_t|const\s+char\s*\*)\s*\*?\s*(tt_\w+)\s*\(", re.M)


def _declared():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        names |= set(DECL.findall(open(h).read()))
    return sorted(names)


def test_library_exports_every_declared_function():
    from tritok_amd import _lib
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_last_error_reports_bad_arguments():
    from tritok_amd import _lib
    f = _lib._sig("tt_lsqr_setup", _lib.i32, _lib.i32, _lib.i32, _lib.i32, _lib.PI, _lib.PI, _lib.PD, _lib.PD,
                  _lib.i32, C.POINTER(_lib.P))
    assert f(1, 1, 1, None, None, None, None, 0, None) < 0  # NULL arrays: rejected before any device call
    assert b"NULL" in _lib.tt_last_error()


def test_macro_mh_rejects_nonpositive_diffusivity():
    """mu = sqrt(beta/diff) must be finite: checked on the host before any
    device call."""
    from tritok_amd import integrals as I
    from tritok_amd import Host
    from tritok_amd._lib import TTError
    h = Host.cube(1)
    m = h.mesh_arrays()
    with pytest.raises(TTError, match="diff must be positive"):
        I.macro_integrals_mh(h.view(), I.quad(), np.zeros(m["nn"]), np.ones(m["nq"]), 1.0)
    with pytest.raises(TTError, match="beta"):
        I.macro_integrals_mh(h.view(), I.quad(), np.ones(m["nn"]), np.ones(m["nq"]), -1.0)


def test_compressed_entry_points_reject_bad_arguments():
    """Argument checks of the compressed-matrix entry points run before any
    device call."""
    from tritok_amd import compress as K
    from tritok_amd._lib import TTError
    with pytest.raises(TTError, match="kappa"):
        K.Wavelet(np.ones((3, 3)), kappa=-1.0)
    with pytest.raises(TTError, match="bad sizes"):
        K.HMatrix(np.ones((4, 4)), np.zeros((4, 3)), np.zeros((4, 3)), leaf=0)
    with pytest.raises(TTError, match="zvr"):
        K.wt_wtawx(np.array([1, 2], np.int32), np.array([1, 1], np.int32), np.ones(2), np.ones(2), 1)


def _codes(h):
    v = h.view()
    nb, nn, nq = v.mesh.nbnodes, v.mesh.nnodes, v.mesh.nq
    arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
    out = {name: arr(getattr(v, name), n) for name, n in
           [("kode", nb * 3), ("wkode", nn * 3), ("wqkode", nq * 3), ("Tkode", nn), ("Tqkode", nq), ("kmu", nn)]}
    if v.kobc:
        out["kobc"] = arr(v.kobc, nb)
    out["nunk"] = np.array(list(v.nunk) + [v.Tnunk, v.nsmcol])
    return out


@pytest.mark.parametrize("case,kind", [("dc2", "cavity"), ("nc2", "heated")])
def test_generated_bic_matches_reference_bic(case, kind):
    from tritok_amd import Host, workloads as W
    g = load_golden(case)
    inp = g["inp_text"].tobytes().decode()
    mk = lambda t: Host.from_arrays(g["mesh_x"], g["mesh_ibc"], g["mesh_idc"], g["mesh_bewn"],
                                    int(g["mesh_nbnodes"])).read_bic(t).read_inp(inp)
    ref = mk(g["bic_text"].tobytes().decode())
    gen = mk(W.bic_cavity() if kind == "cavity" else W.bic_heated(t_x0=-0.5, t_x1=0.5))
    a, b = _codes(ref), _codes(gen)
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    fa, fb = ref.fields(), gen.fields()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k


@pytest.mark.parametrize("n", [2, 3, 8, 20])
def test_cube_sizes(n):
    from tritok_amd import Host
    h = Host.cube(n)
    m = h.view().mesh
    assert (m.nnodes, m.nbnodes, m.nbelem, m.nicell, m.nq) == (
        (2 * n + 1) ** 3, 6 * (2 * n) ** 2 + 2, 6 * n * n, n ** 3, 12 * n * n * (n + 1))


def test_cube_matches_reference_mesh_geometry():
    """Host.cube(2) has the node set, cells and boundary of kva2x2x2.geo (up to numbering)."""
    from tritok_amd import Host
    g = load_golden("dc2")
    a = Host.cube(2).mesh_arrays()
    x_ref = g["mesh_x"]
    key = lambda p: tuple(np.round(p, 9))
    pos = {key(p): i for i, p in enumerate(a["x"])}
    perm = np.array([pos[key(p)] for p in x_ref])  # reference node -> cube node
    assert len(set(perm)) == len(x_ref)
    cells_ref = {tuple(sorted(perm[c - 1])) for c in g["mesh_idc"]}
    cells_cube = {tuple(sorted(c - 1)) for c in a["idc"]}
    assert cells_ref == cells_cube
    faces_ref = {tuple(sorted(perm[f - 1])) for f in g["mesh_ibc"]}
    faces_cube = {tuple(sorted(f - 1)) for f in a["ibc"]}
    assert faces_ref == faces_cube


def test_read_inp_parses_hot_path_keywords():
    from tritok_amd import workloads as W
    h = W.natural_convection(2, ra=1e4, pr=0.7, eps=1e-7, lsqs_eps=1e-9)
    v = h.view()
    assert v.iDT == 1 and v.RaN == 1e4 and v.PrN == pytest.approx(0.7)
    assert v.eps == pytest.approx(1e-7) and v.lsqs_eps == pytest.approx(1e-9, rel=1e-6)
    assert list(v.urDw) == [0.1, 0.1, 0.1] and v.urDT == pytest.approx(0.05)
    with pytest.raises(Exception):
        W.Host.cube(2).read_bic(W.bic_heated()).read_inp("EQNS 1 1 1 1 1 0\n")  # diffusion-convection: not on this path


AMDFLANG = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(AMDFLANG), reason="amdflang not installed")
def test_fortran_binding_compiles_and_matches_c_layout(tmp_path):
    """tritok_amd/fortran/tritok_amd_bind.f90 compiles, and its BIND(C) types
    have the sizes of the C structs (so a Fortran host passes them as-is)."""
    import subprocess
    src = os.path.join(REPO, "tritok_amd", "fortran", "tritok_amd_bind.f90")
    prog = tmp_path / "sizes.f90"
    prog.write_text("PROGRAM sizes\n  USE tritok_amd_bind\n  TYPE(tt_mesh_view) :: m\n  TYPE(tt_quad) :: q\n"
                    "  TYPE(tt_case_view) :: v\n  PRINT '(3I6)', C_SIZEOF(m), C_SIZEOF(q), C_SIZEOF(v)\nEND PROGRAM\n")
    subprocess.run([AMDFLANG, "-c", src, "-o", str(tmp_path / "b.o"), "-module-dir", str(tmp_path)], check=True,
                   cwd=tmp_path)
    libdir = os.path.join(REPO, "tritok_amd")
    subprocess.run([AMDFLANG, str(prog), str(tmp_path / "b.o"), "-I", str(tmp_path), "-o", str(tmp_path / "sizes"),
                    "-L" + libdir, "-ltritok_amd", "-Wl,-rpath," + libdir], check=True, cwd=tmp_path)
    fsz = subprocess.run([str(tmp_path / "sizes")], check=True, capture_output=True, text=True).stdout.split()
    csrc = tmp_path / "sizes.c"
    csrc.write_text('#include <stdio.h>\n#include "tritok_bem.h"\nint main(void){printf("%zu %zu %zu\\n",'
                    'sizeof(tt_mesh_view),sizeof(tt_quad),sizeof(tt_case_view));return 0;}\n')
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(csrc), "-o", str(tmp_path / "csizes")], check=True)
    csz = subprocess.run([str(tmp_path / "csizes")], check=True, capture_output=True, text=True).stdout.split()
    assert fsz == csz, (fsz, csz)
