"""The reference's file formats written and read by the library (host
code): tri.macro.int, tri.sdkm.int.N, tri.rst version 4, tri.sta, tri.ite.

files_ref.npz holds the reference's own files of the dc2 run (tri.rst,
tri.sta and tri.ite as bytes; the SHA-256 of tri.macro.int and
tri.sdkm.int.1, whose values dc2.npz holds): written from the same values,
the library's files must be byte-for-byte the reference's."""
import hashlib
import os
import sys

import numpy as np

from conftest import GOLDEN, load_golden

sys.path.insert(0, GOLDEN)
import make_golden as MG  # noqa: E402  (the fixtures' own record parsers)

REF = load_golden("files_ref")
DC2 = load_golden("dc2")
NN, NB, NC, NQ = 125, 98, 8, 144


def _sha(path):
    return np.frombuffer(hashlib.sha256(open(path, "rb").read()).digest(), np.uint8)


def test_macro_cache_is_the_reference_file(tmp_path):
    from tritok_amd import files as F
    m = {k: DC2["macro_" + k] for k in F.MACRO}
    m["ont"] = DC2["macro_ont"]
    p = tmp_path / "tri.macro.int"
    F.write_macro_int(p, m, NN, NB)
    assert os.path.getsize(p) == int(REF["macro_bytes"])
    assert len(MG.read_records(str(p))) == int(REF["macro_records"])
    assert np.array_equal(_sha(p), REF["macro_sha256"])
    back = F.read_macro_int(p, NC)
    assert back["nnodes"] == NN and back["nbnodes"] == NB
    for k in F.MACRO:
        assert np.array_equal(back[k], m[k]), k
    assert np.array_equal(back["ont"], m["ont"])


def test_sdkm_cache_is_the_reference_file(tmp_path):
    from tritok_amd import files as F
    sd = {k: DC2["sd_" + k] for k in F.SDKM}
    sd["ont"] = DC2["sd_ont"]
    p = tmp_path / "tri.sdkm.int.1"
    F.write_sdkm_int(p, sd, NN, NB, nproc=1)
    assert os.path.getsize(p) == int(REF["sdkm_bytes"])
    assert np.array_equal(_sha(p), REF["sdkm_sha256"])
    back = F.read_sdkm_int(p, NB, NN, NB)
    assert list(back["header"]) == [NN, NB, 1]
    for k in F.SDKM:
        assert np.array_equal(back[k], sd[k]), k


def _records_to_fields(raw_path):
    recs = MG.read_records(raw_path)
    v = MG.doubles(recs[3:])
    out, p = {}, 0
    from tritok_amd import files as F
    for k in F.RESTART:
        rows = NQ if k in ("qvorticity", "qtemp", "qdc") else NN
        cols = 3 if k in ("velocity", "pnlsvelocity", "vorticity", "ptsvorticity", "pptsvorticity", "pnlsvorticity",
                          "qvorticity") else 1
        out[k] = v[p:p + rows * cols].reshape((cols, rows)).T
        p += rows * cols
    assert p == len(v)
    hdr = (np.frombuffer(recs[0], "<i4")[0], np.frombuffer(recs[1], "<i4"), np.frombuffer(recs[2], "<f8")[0])
    return out, hdr


def test_restart_is_the_reference_file_and_reads_back(tmp_path):
    from tritok_amd import files as F
    raw = tmp_path / "ref.rst"
    raw.write_bytes(REF["rst_raw"].tobytes())
    fields, (ver, cnt, rtime) = _records_to_fields(str(raw))
    assert ver == 4 and int(cnt[1]) == len(DC2["sta"])
    # the converged fields of dc2.npz are these records
    assert np.array_equal(fields["velocity"], DC2["rst_velocity"])
    p = tmp_path / "tri.rst"
    F.write_restart(p, fields, NN, NQ, tstep=int(cnt[0]), glnnlit=int(cnt[1]), rtime=rtime)
    assert p.read_bytes() == raw.read_bytes()
    back, v, ts, gl, rt = F.read_restart(p, NN, NQ)
    assert (v, ts, gl, rt) == (4, int(cnt[0]), int(cnt[1]), rtime)
    for k in F.RESTART:
        assert np.array_equal(back[k], fields[k]) or (np.isnan(back[k]).all() and np.isnan(fields[k]).all()), k


def test_sta_and_ite_lines_are_the_reference_lines(tmp_path):
    from tritok_amd import files as F
    sta, ite = DC2["sta"], DC2["ite"]
    F.write_sta_ite(tmp_path / "tri.sta", tmp_path / "tri.ite", err=sta[:, 2:11], nit=ite[:, 2:14],
                    tstep=sta[:, 0].astype(np.int32), tsnnlit=sta[:, 1].astype(np.int32))
    for name in ("sta", "ite"):
        ref = REF[name + "_raw"].tobytes().decode().splitlines()
        got = (tmp_path / ("tri." + name)).read_text().splitlines()
        assert got[1:] == [line for line in ref[1:]], name  # the first line names the program


def test_g10_5_edge_values(tmp_path):
    """Fortran G10.5 around the F/E boundaries (0.1, 1e5), zero and the cap."""
    from tritok_amd import files as F
    vals = [0.0, 0.09999949, 0.0999995, 0.1, 0.99999, 0.999996, 1.0, 9.999999999, 12345.4, 99999.4, 99999.6,
            1.5e-5, 3.2e5, 1e-120]
    F.write_sta_ite(tmp_path / "s", None, err=np.array(vals[:9]), nit=np.zeros(12, np.int32))
    F.write_sta_ite(tmp_path / "t", None, err=np.array(vals[5:]), nit=np.zeros(12, np.int32))
    line = (tmp_path / "s").read_text().splitlines()[2]
    items = [line[15 + 12 * k:15 + 12 * k + 10] for k in range(9)]
    assert items == ["0.0000    ", ".99999E-01", ".10000    ", ".10000    ", ".99999    ", "1.0000    ",
                     "1.0000    ", "10.000    ", "12345.    "]
    line = (tmp_path / "t").read_text().splitlines()[2]
    items = [line[15 + 12 * k:15 + 12 * k + 10] for k in range(9)]
    assert items[3:] == ["12345.    ", "99999.    ", ".10000E+06", ".15000E-04", ".32000E+06", ".10000-119"]


def test_reference_log_and_history_parsers(tmp_path):
    """The CPU-baseline leg reads the reference's Picard-loop CPU time from the
    table in tri.log and its per-iteration changes from tri.sta."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import cpu_sample
    (tmp_path / "tri.log").write_text(
        "   CPU time information [s]\n|-------------------------------|-------------|\n"
        "| z Dw solve                    |       1.002 |\n| solve                         |       4.044 |\n"
        "| TOTAL                         |       8.390 |\n")
    assert cpu_sample._ref_solve_seconds(str(tmp_path / "tri.log")) == 4.044
    (tmp_path / "tri.sta").write_text(
        "TriTok v8.0, Februar 2025\n tstep  tsite          vx          vy\n"
        "     1      1  .10000E+01  .20000E-01  .30000E-02  .40000E-03  .50000E-04  .60000E-05  0.0000      0.0000      0.0000    \n"
        "     1      2  .11000E+01  .21000E-01  .31000E-02  .41000E-03  .51000E-04  .61000E-05  0.0000      0.0000      0.0000    \n")
    sta = cpu_sample._ref_sta(str(tmp_path / "tri.sta"))
    assert sta.shape == (2, 9) and sta[1, 0] == 1.1 and sta[0, 5] == 6e-6
