"""The request loop: each case's converged fields are the same when it runs
alone and when it runs concurrently with others (slots sharing the device),
in FP64 and in mixed precision; admission accounting matches the device
memory the cases hold."""
import numpy as np
import pytest

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
