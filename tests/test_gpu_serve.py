"""The request loop: each case's converged fields are the same when it runs
alone and when it runs concurrently with others (slots sharing the device)
or in a batch with others (continuous batching: shared LSQR launches), in
FP64 and in mixed precision; admission accounting matches the device memory
the cases hold; two requests of the serving stream (6^3 cavity and
convection, tight tolerances) match the reference's converged fields within
1e-8 relative L2 whether served alone, among other stream requests or in a
batch with them."""
import numpy as np
import pytest

from conftest import host_for, load_golden, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("velocity", "vorticity", "qvorticity", "temp", "qtemp")


def _cases():
    from tritok_amd import workloads as W
    return [(W.cavity(2, re=100.0), "fp64"), (W.natural_convection(2, ra=1e4), "fp64"),
            (W.cavity(3, re=50.0), "fp16"), (W.natural_convection(3, ra=2e4), "fp32")]


def test_results_alone_equal_results_in_the_mix():
    from tritok_amd.serve import Server
    cases = _cases()
    alone = []
    with Server(slots=1) as s:
        for host, prec in cases:
            rid = s.submit(host, prec)
            info = s.wait(rid)
            assert info["status"] == 0, info
            alone.append((s.fields(rid), info))
    with Server(slots=4) as s:
        ids = [s.submit(host, prec) for host, prec in cases]
        infos = [s.wait(i) for i in ids]
        mixed = [(s.fields(i), info) for i, info in zip(ids, infos)]
        starts = sorted(s.wait(i)["t_start"] for i in ids)
        ends = sorted(s.wait(i)["t_end"] for i in ids)
    assert starts[-1] < ends[0]  # all four were admitted before the first finished: they overlapped
    for (fa, ia), (fm, im), (_, prec) in zip(alone, mixed, cases):
        assert im["status"] == 0 and im["picard_iterations"] == ia["picard_iterations"], (ia, im)
        for k in FIELDS:
            assert np.isfinite(fa[k]).all() and np.isfinite(fm[k]).all(), (prec, k)
            scale = max(np.abs(fa[k]).max(), 1e-300)
            d = np.abs(fm[k] - fa[k]).max() / scale
            print(prec, k, "alone vs mix max rel", d, flush=True)
            assert d <= 1e-12, (prec, k, d)
        assert ia["device_bytes"] <= ia["estimated_bytes"], ia


def test_server_matches_direct_case_run():
    from tritok_amd import workloads as W
    from tritok_amd.case import Case
    from tritok_amd.serve import Server
    host = W.cavity(2, re=100.0)
    with Case(host, prec="fp64") as c:
        sta, _ = c.run()
        ref = c.fields()
    with Server(slots=2) as s:
        rid = s.submit(host, "fp64")
        info = s.wait(rid)
        got = s.fields(rid)
    assert info["picard_iterations"] == len(sta)
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_stream_requests_match_reference_alone_and_in_the_mix(prec):
    from tritok_amd import workloads as W
    from tritok_amd.serve import Server
    fx = {k: load_golden(k) for k in ("sv6c", "sv6n")}
    others = [W.serving_case(sp) for sp in W.serving_stream(count=60) if sp["n"] in (6, 8)][:4]
    alone = {}
    with Server(slots=1) as s:
        for k, g in fx.items():
            print("alone", k, prec, flush=True)
            rid = s.submit(host_for(g), prec)
            info = s.wait(rid)
            assert info["status"] == 0, info
            alone[k] = s.fields(rid)
    hosts = {k: host_for(g) for k, g in fx.items()}
    order = [("sv6c", hosts["sv6c"]), (0, others[0]), (1, others[1]), ("sv6n", hosts["sv6n"]), (2, others[2]),
             (3, others[3])]
    with Server(slots=6) as s:
        print("mix", prec, flush=True)
        ids = {k: s.submit(h, prec) for k, h in order}
        infos = {k: s.wait(i) for k, i in ids.items()}
        mixed = {k: s.fields(ids[k]) for k in fx}
    for k in fx:  # each fixture request ran while others were on the device
        a = infos[k]
        over = [j for j, b in infos.items() if j != k and b["t_start"] < a["t_end"] and a["t_start"] < b["t_end"]]
        assert len(over) >= 2, (k, over)
    for k, g in fx.items():
        assert infos[k]["status"] == 0, infos[k]
        for f in FIELDS:
            assert np.array_equal(mixed[k][f], alone[k][f]), (k, prec, f)
            ref = g["rst_" + f].reshape(-1)
            if np.linalg.norm(ref) == 0:
                assert np.abs(mixed[k][f]).max() < 1e-12, (k, f)
                continue
            e = rel_l2(np.asarray(mixed[k][f]).reshape(-1), ref)
            print(k, prec, f, "vs reference rel L2", e, flush=True)
            assert e < 1e-8, (k, prec, f, e)


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_batched_serving_matches_alone_and_reference(prec):
    from tritok_amd import workloads as W
    from tritok_amd.serve import Server
    fx = {k: load_golden(k) for k in ("sv6c", "sv6n")}
    hosts = {k: host_for(g) for k, g in fx.items()}
    others = [W.serving_case(sp) for sp in W.serving_stream(count=80) if sp["n"] == 6][:4]
    order = [("sv6c", hosts["sv6c"]), (0, others[0]), ("sv6n", hosts["sv6n"]), (1, others[1]), (2, others[2]),
             (3, others[3])]
    alone = {}
    with Server(slots=1) as s:
        for k, h in order:
            print("alone", k, prec, flush=True)
            rid = s.submit(h, prec)
            info = s.wait(rid)
            assert info["status"] == 0 and info["batch_peak"] == 1, info
            alone[k] = (s.fields(rid), info)
    with Server(slots=1, max_batch=8) as s:
        print("batched", prec, flush=True)
        ids = {k: s.submit(h, prec) for k, h in order}
        infos = {k: s.wait(i) for k, i in ids.items()}
        got = {k: s.fields(i) for k, i in ids.items()}
    for k, _ in order:
        a, b = alone[k][1], infos[k]
        print(k, prec, "picard", b["picard_iterations"], "batch peak", b["batch_peak"], flush=True)
        assert b["status"] == 0 and b["batch_peak"] >= 2, b
        assert b["picard_iterations"] == a["picard_iterations"] and b["lsqr_iterations"] == a["lsqr_iterations"]
        for f in FIELDS:
            assert np.array_equal(got[k][f], alone[k][0][f]), (k, prec, f)
    for k, g in fx.items():
        for f in FIELDS:
            ref = g["rst_" + f].reshape(-1)
            if np.linalg.norm(ref) == 0:
                assert np.abs(got[k][f]).max() < 1e-12, (k, f)
                continue
            e = rel_l2(np.asarray(got[k][f]).reshape(-1), ref)
            print(k, prec, f, "batched vs reference rel L2", e, flush=True)
            assert e < 1e-8, (k, prec, f, e)

