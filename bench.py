"""Picard iterations/s of the steady natural-convection case on one MI355X.

Workload: differentially heated unit cube at n^3 = 20^3 subdomains (27-node
cells, 41^3 = 68,921 nodes, 9,602 boundary nodes), Ra = 1e5, Pr = 0.71, the
reference's iteration parameters (GAUS 4 4 4 4, IIDI 2, LSQS 5000 1e-7,
DLSE 1 10 0.001) with the relaxation that converges at this Rayleigh number
from rest (URDW 0.1, URDT 0.05).  Setup (integrals, factorisations,
least-squares systems) is done once; the matrices stay resident in HBM.
The CPU baseline is the oracle timed on a bounded sample of the same
workload (oracle/cpu_sample.py).

A step is one Picard iteration (main.f90:1200-1589): boundary vorticities,
kinematics x,y,z, energy, vorticity transport x,y,z, convergence test and
dynamic LSQR tolerance.  The FP64 leg and the mixed leg (FP16 least-squares
blocks with FP32 accumulation and FP64 hot-start residuals, scalars and
vectors; FP16 kinematics products with FP64 anchor steps) start from the same
initial fields; `value` is the mixed leg's rate, the FP64 leg's is reported
beside it, with the converged mixed-vs-FP64 difference of a tight-tolerance
run (accuracy leg).  Further legs: the batched modified-Helmholtz integrals,
the batched least-squares solves and the serving request stream.

    python bench.py [--steps K] [--warmup W] [--n 20] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "Picard iters/s at 20³ subdomains, mixed vs FP64 at 1e-8 rel; p99 solve latency"
HBM_PEAK_GBS = 8000.0
FIELDS = ("velocity", "vorticity", "qvorticity", "temp", "qtemp")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / nb) if nb > 0 else float(np.linalg.norm(a))


def run_leg(case, prec, steps, warmup, probe_stride):
    case.set_precision(prec)
    case.reset()
    for _ in range(warmup):
        case.step()
    case.probe(probe_stride)
    lat, nits, phases = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        rc, err, nit = case.step()
        lat.append(time.perf_counter() - ts)
        nits.append(nit.copy())
        phases.append(case.timings())
    total = time.perf_counter() - t0
    stats51 = case.probe_stats(51)
    stats27 = case.probe_stats(27)
    case.probe(0)
    return dict(total_s=total, lat=np.array(lat), nits=np.array(nits), phases=phases, probe51=stats51,
                probe27=stats27, fields=case.fields())


def copy_for_calibration(nbytes):
    """One device-to-device copy of nbytes (the runtime's copy kernel), so that
    a counter run sees a kernel of known traffic: nbytes read, nbytes written."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a), hip.hipFree(b)


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/*traffic*.md: a line 'traffic_bytes_per_launch <kernel> <bytes>')."""
    import glob
    for f in sorted(glob.glob(os.path.join(REPO, "profiles", "*traffic*.md"))):
        for line in open(f):
            t = line.split()
            if len(t) == 3 and t[0] == "traffic_bytes_per_launch" and t[1] == kernel:
                return float(t[2])
    return None


def accuracy_leg(n, mixed, anchor_period=16):
    """Converged fields of the heated cube at tight tolerances (Picard change
    1e-10; the reference's dynamic LSQR tolerance, change/10 capped at 1e-3,
    down to 1e-12), FP64 and mixed, from the same start.  (With a fixed LSQR
    tolerance of 1e-12 from the first step the 20^3 Picard iteration diverges
    from rest; the dynamic tolerance damps the first steps.)"""
    from tritok_amd import workloads
    from tritok_amd.case import Case
    host = workloads.natural_convection(n, eps=1e-10, lsqs_eps=1e-12, dlse=(1, 10.0, 0.001), nnlit=20000)
    out = {}
    with Case(host, prec=mixed) as c:
        c.set_anchor_period(anchor_period)
        for prec in ("fp64", mixed):
            c.set_precision(prec)
            c.reset()
            t0 = time.perf_counter()
            sta, ite = c.run()
            out[prec] = (c.fields(), len(sta), time.perf_counter() - t0)
            log("accuracy %d^3 %s: %d Picard iterations to 1e-10 in %.1f s" % (n, prec, len(sta), out[prec][2]))
    diff = {k: rel_l2(out[mixed][0][k], out["fp64"][0][k]) for k in FIELDS}
    return {"workload": "heated cube %d^3, Ra=1e5, Picard to 1e-10, dynamic LSQR tolerance down to 1e-12" % n,
            "rel_l2": diff, "max_rel_l2": max(diff.values()), "iterations": {p: out[p][1] for p in out},
            "seconds": {p: out[p][2] for p in out}}


def serving_leg(requests, slots, max_iter, precs, max_batch=8, seed=20251203):
    """The serving workload's request stream (Poisson arrivals at 20/s, cubes
    of 6^3..16^3 subdomains, half convection / half cavity, each to a Picard
    change of 1e-6), first `requests` arrivals submitted in real time to the
    request loop, once per precision; each slot steps a batch of up to
    max_batch cases of one kind together.  Latency = finish - submit."""
    from tritok_amd import workloads
    from tritok_amd.serve import Server
    stream = workloads.serving_stream(count=requests, rate=20.0, seed=seed)
    hosts = [workloads.serving_case(s) for s in stream]  # the host side prepares inputs before they arrive
    out = {"workload": "first %d requests of the seeded 2000-request stream: Poisson 20/s, n in {6,8,12,16}, "
                       "50%% convection Ra log-U[1e3,1e6] / 50%% cavity Re U[10,400], Picard to 1e-6 "
                       "(cap %d iterations)" % (requests, max_iter),
           "requests": requests, "slots": slots, "max_batch": max_batch, "arrival_rate_per_s": 20.0}
    for prec in precs:
        with Server(slots=slots, max_iter=max_iter, max_batch=max_batch) as srv:
            t0 = time.perf_counter()
            ids = []
            for spec, h in zip(stream, hosts):
                dt = spec["t"] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(h, prec))
            infos = []
            for k, i in enumerate(ids):
                infos.append(srv.wait(i))
                if (k + 1) % 20 == 0:
                    log("serving %s: %d of %d finished after %.1f s" % (prec, k + 1, len(ids),
                                                                         time.perf_counter() - t0))
            wall = time.perf_counter() - t0
        lat = np.array([i["t_end"] - i["t_submit"] for i in infos])
        run = np.array([i["t_end"] - i["t_start"] for i in infos])
        setup = np.array([i["t_setup"] - i["t_start"] for i in infos])
        per_n = {}
        for sp, i in zip(stream, infos):
            d = per_n.setdefault(sp["n"], {"count": 0, "setup_s": [], "picard_s": [], "its": []})
            d["count"] += 1
            d["setup_s"].append(i["t_setup"] - i["t_start"])
            d["picard_s"].append(i["t_end"] - i["t_setup"])
            d["its"].append(i["picard_iterations"])
        per_n = {str(k): {"count": d["count"], "setup_s_mean": float(np.mean(d["setup_s"])),
                          "picard_s_mean": float(np.mean(d["picard_s"])), "picard_its_mean": float(np.mean(d["its"]))}
                 for k, d in sorted(per_n.items())}
        ok = [i["status"] == 0 for i in infos]
        log("serving %s: %d requests in %.1f s, latency p50 %.2f s p99 %.2f s, %d converged, %d capped, %d failed" % (
            prec, requests, wall, np.percentile(lat, 50), np.percentile(lat, 99), sum(ok),
            sum(i["status"] == 1 for i in infos), sum(i["status"] < 0 for i in infos)))
        out[prec] = {"latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                   "mean": float(lat.mean()), "max": float(lat.max())},
                     "service_s": {"p50": float(np.percentile(run, 50)), "p99": float(np.percentile(run, 99))},
                     "solves_per_s": requests / wall, "wall_s": wall, "converged": int(sum(ok)),
                     "keeps_up_with_arrivals": bool(requests / wall >= 20.0),
                     "setup_share_of_service": float(setup.sum() / run.sum()), "per_size": per_n,
                     "capped": int(sum(i["status"] == 1 for i in infos)),
                     "failed": int(sum(i["status"] < 0 for i in infos)),
                     "mean_picard_iterations": float(np.mean([i["picard_iterations"] for i in infos])),
                     "batch_peak_mean": float(np.mean([i["batch_peak"] for i in infos])),
                     "per_request": [{"n": s["n"], "kind": s["kind"], "its": i["picard_iterations"],
                                      "status": i["status"], "latency_s": round(float(l), 3)}
                                     for s, i, l in zip(stream, infos, lat)]}
    return out


FP64_PEAK_TFLOPS = 78.6
# algorithmic flops per fundamental-solution evaluation, modified-Helmholtz
# kernel: 27 shapes x 4 outputs x 2 + 15 (+3 for mu) over cells, (7 x 9 + 4) x 2
# + 25 (+3) over faces; square roots, divisions, exp and the singular
# transforms' trigonometry are not credited
FLOP_DOMAIN_MH, FLOP_BOUNDARY_MH = 234.0, 162.0


def batched_integrals_leg(n_side, seed=20251203):
    """Modified-Helmholtz macro integrals of n_side^3 independent jittered
    cells (mu log-uniform in [0.1, 100] per cell) in one launch."""
    from tritok_amd import integrals as I
    from tritok_amd import workloads
    b = workloads.cell_batch(n_side, jitter=0.1, seed=seed, mu_range=(0.1, 100.0))
    am = I.ArrayMesh(b["arrays"])
    q = I.quad(4, 4, 4, 4, 2)
    I.macro_integrals_mh(am, q, b["diff"], b["qdiff"], b["beta"])  # warm-up (module load)
    t0 = time.perf_counter()
    I.macro_integrals_mh(am, q, b["diff"], b["qdiff"], b["beta"])
    wall = time.perf_counter() - t0
    st = I.last_integral_stats()
    flops = st["domain_points"] * FLOP_DOMAIN_MH + st["boundary_points"] * FLOP_BOUNDARY_MH
    tflops = flops / (st["ms"] * 1e-3) / 1e12
    nc = am.mesh.nicell
    log("batched integrals: %d cells in %.1f ms on the device (%.2f s with transfers), %.3g domain + %.3g boundary "
        "evaluations, %.2f TFLOP/s" % (nc, st["ms"], wall, st["domain_points"], st["boundary_points"], tflops))
    return {"workload": "%d independent 27-node cells (%d^3 arrangement, corners jittered +-10%% of the edge), "
                        "mu log-uniform in [0.1, 100] per cell, modified-Helmholtz macro matrices" % (nc, n_side),
            "cells": nc, "kernel_ms": st["ms"], "cells_per_s": nc / (st["ms"] * 1e-3), "wall_s_with_transfers": wall,
            "domain_points": st["domain_points"], "boundary_points": st["boundary_points"],
            "roofline": {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tflops / FP64_PEAK_TFLOPS,
                         "flop_per_point": {"domain": FLOP_DOMAIN_MH, "boundary": FLOP_BOUNDARY_MH}}}


def batched_solves_leg(count, seed=20251203):
    """`count` independent lid-driven-cavity vorticity systems (n^3 cells, n
    uniform in 2..10, Re uniform in [10, 400], seeded; right-hand sides
    synthetic), solved to the reference's LSQR tolerance 1e-7 in one launch."""
    from tritok_amd import integrals as I
    from tritok_amd import solvers as S
    from tritok_amd import workloads
    rng = np.random.default_rng(seed)
    ns = rng.integers(2, 11, size=count)
    res = rng.uniform(10.0, 400.0, size=count)
    parts = {}
    systems, bs = [], []
    for n, re in zip(ns, res):
        n = int(n)
        if n not in parts:
            host = workloads.cavity(n, re=100.0)
            parts[n] = workloads.vte_csr_parts(host, I.macro_integrals(host.view(), I.quad()))
        ia, ja, ah, ab, nu, beta = parts[n]
        systems.append((ia, ja, ah + ab * beta * re, nu))
        bs.append(rng.standard_normal(len(ia) - 1))
    x0 = [np.zeros(s_[3]) for s_ in systems]
    t0 = time.perf_counter()
    B = S.LsqrBatch(systems)
    setup = time.perf_counter() - t0
    m = np.array([len(s_[0]) - 1 for s_ in systems], float)
    nnz = np.array([len(s_[1]) for s_ in systems], float)
    nunk = np.array([s_[3] for s_ in systems], float)
    out = {"workload": "%d lid-driven-cavity vorticity systems, n^3 cells with n uniform in 2..10, Re uniform in "
                       "[10, 400], synthetic right-hand sides, LSQR eps 1e-7, FP64 and FP32 values" % count,
           "systems": count, "nonzeros": float(nnz.sum()), "setup_s": setup}
    x64 = None
    scale, tune_ms = B.tune(bs, x0, 5000, 1e-7)  # gang-width autotuner (FP64 values)
    out["gang_width_tuning"] = {"best_scale": scale, "ms_per_scale": tune_ms}
    log("batched solves: gang widths scaled by %g (ms per scale %s)" % (scale, tune_ms))
    for prec, vb in (("fp64", 8), ("fp32", 4)):
        B.set_precision(prec)
        B.solve(bs, x0, 5000, 1e-7)  # warm-up (FP32: makes the value copies)
        xs, istop, nit, ms = B.solve(bs, x0, 5000, 1e-7)
        alg = float(np.sum(nit * (2 * ((vb + 4) * nnz + 4 * (m + 1)) + 8 * (3 * m + 6 * nunk))))
        gbs = alg / (ms * 1e-3) / 1e9
        if prec == "fp64":
            x64 = xs
        err = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(xs, x64))
        log("batched solves %s: %d systems (%.3g nonzeros) in %.1f ms, %d..%d iterations, %.0f GB/s algorithmic, "
            "max rel diff vs FP64 %.2e" % (prec, count, nnz.sum(), ms, nit.min(), nit.max(), gbs, err))
        out[prec] = {"kernel_ms": ms, "systems_per_s": count / (ms * 1e-3),
                     "iterations": {"min": int(nit.min()), "mean": float(nit.mean()), "max": int(nit.max())},
                     "istop": {int(k): int(v) for k, v in zip(*np.unique(istop, return_counts=True))},
                     "max_rel_l2_vs_fp64": err,
                     "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": gbs / HBM_PEAK_GBS,
                                  "bytes": "per system iteration 2*((%d+4)*nnz + 4*(m+1)) + 8*(3m + 6n)" % vb}}
    B.close()
    # the same systems on the cell-block format (groups of up to 32 systems
    # per launch of each LSQR kernel)
    t0 = time.perf_counter()
    KB = S.BlockBatch(systems, rows_per_cell=51)
    out["blocks"] = {"setup_s": time.perf_counter() - t0,
                     "form": "per-cell 51 x 51 blocks of the scaled matrix, fused A v / A^T u' cell kernel, groups "
                             "of up to 32 systems of similar size per launch"}
    cells = m / 51.0
    for prec, vb in (("fp64", 8), ("fp32", 4), ("fp16", 2)):
        KB.set_precision(prec)
        KB.solve(bs, x0, 5000, 1e-7)  # warm-up: tuning and the solve histories that size the launch batches
        xs, istop, nit, ms = KB.solve(bs, x0, 5000, 1e-7)
        alg = float(np.sum(nit * (cells * (51 * 51 * vb + 51 * 4 + 51 * 8 + 51 * 4) + 16 * m + 8 * 6 * nunk)))
        gbs = alg / (ms * 1e-3) / 1e9
        err = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(xs, x64))
        log("batched solves on blocks %s: %.1f ms, %d..%d iterations, %.0f GB/s algorithmic, max rel diff vs CSR "
            "FP64 %.2e" % (prec, ms, nit.min(), nit.max(), gbs, err))
        out["blocks"][prec] = {
            "kernel_ms": ms, "systems_per_s": count / (ms * 1e-3),
            "iterations": {"min": int(nit.min()), "mean": float(nit.mean()), "max": int(nit.max())},
            "istop": {int(k): int(v) for k, v in zip(*np.unique(istop, return_counts=True))},
            "max_rel_l2_vs_csr_fp64": err,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS,
                         "bytes": "per system iteration cells*(51*51*%d (blocks) + 51*16 (column map, v, A^T u' "
                                  "partials)) + 16m + 48n" % vb}}
    KB.close()
    return out


def mixed_kinematics_leg(n, tol=1e-10):
    """Mixed-precision kinematics on an n^3 cube: the x equation's system
    (nbnodes x nbnodes) solved for four right-hand sides by FP64 iterative
    refinement with its inverse held in FP32, BF16 and FP16 (per-row 2^k
    scales, matrix cores, FP32 accumulation), against the FP64 solve."""
    from tritok_amd import integrals as I
    from tritok_amd import solvers as S
    from tritok_amd import workloads
    host = workloads.natural_convection(n, ra=1e5, pr=0.71)
    m = host.mesh_arrays()
    nb = len(m["gbn"])
    t0 = time.perf_counter()
    sd = S.SdKinematics(host, I.quad(4, 4, 4, 4, 2))
    setup = time.perf_counter() - t0
    # the boundary values of a smooth velocity and temperature field as the
    # four right-hand-side columns (u_x, u_y, u_z, T)
    p = m["x"][m["gbn"] - 1]
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    B = np.stack([np.sin(np.pi * y) * np.sin(np.pi * z), np.sin(np.pi * z) * np.sin(np.pi * x),
                  np.sin(np.pi * x) * np.sin(np.pi * y), x - 0.5], axis=1)
    # each solve timed three times, the median kept (single runs of a
    # millisecond-long solve pick up the odd stall of the box)
    def timed(prec, **kw):
        runs = [sd.refine(1, prec, B, **kw) for _ in range(3)]
        order = sorted(range(3), key=lambda i: runs[i][3])
        return runs[order[1]]

    sd.refine(1, "fp64", B, tol=1e-14, maxit=4)
    X64, it64, rr64, ms64, mv64 = timed("fp64", tol=1e-14, maxit=4)
    out = {"workload": "kinematics x equation of the %d^3 heated cube (%d x %d), right-hand sides u_x, u_y, u_z, T "
                       "(smooth fields at the boundary nodes), FP64 iterative refinement to %g" % (n, nb, nb, tol),
           "nbnodes": nb, "setup_s": setup, "timing": "median of three solves after a first untimed one",
           "fp64": {"corrections": it64, "relres": rr64, "ms": ms64, "ms_product": mv64}}
    bytes_per = {"fp32": 4, "bf16": 2, "fp16": 2}
    for prec in ("fp32", "bf16", "fp16"):
        sd.refine(1, prec, B, tol=tol, maxit=60)  # first use builds the tile copy
        X, it, rr, ms, mv = timed(prec, tol=tol, maxit=60)
        err = float(np.linalg.norm(X - X64) / np.linalg.norm(X64))
        gbs = nb * nb * bytes_per[prec] / (mv * 1e-3) / 1e9
        tflops = 2.0 * nb * nb * 16 / (mv * 1e-3) / 1e12  # the instructions' 16 columns
        log("mixed kinematics %s: %d corrections to relres %.2e, error vs FP64 %.2e, %.2f ms (product %.3f ms, "
            "%.0f GB/s)" % (prec, it, rr, err, ms, mv, gbs))
        out[prec] = {"corrections": it, "relres": rr, "rel_err_vs_fp64": err, "ms": ms, "ms_product": mv,
                     "product_GBps": gbs, "product_hbm_frac": gbs / HBM_PEAK_GBS, "mfma_tflops": tflops}
    log("mixed kinematics fp64: %d corrections, relres %.2e, %.2f ms (product %.3f ms)" % (it64, rr64, ms64, mv64))
    sd.close()
    return out


def aca_kinematics_leg(n, leaf=64, eta=1.0, eps=1e-6, reps=5):
    """The heated cube's single-domain H (nbnodes x nbnodes) compressed where
    it lies on the device: geometric clusters of at most `leaf` boundary
    nodes, admissible blocks cross-approximated (ACA type 2, tolerance eps),
    the rest dense; compression, build time, the four-column product
    (u_x, u_y, u_z, T) against the dense device product."""
    from tritok_amd import integrals as I
    from tritok_amd import solvers as S
    from tritok_amd import workloads
    host = workloads.natural_convection(n, ra=1e5, pr=0.71)
    nb = host.view().mesh.nbnodes
    sd = S.SdKinematics(host, I.quad(4, 4, 4, 4, 2))
    t0 = time.perf_counter()
    hm = sd.hmat("H", leaf=leaf, eta=eta, aca_type=2, aca_eps=eps)
    build = time.perf_counter() - t0
    X = np.random.default_rng(20251204).standard_normal((nb, 4))
    hm.mv(X)  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        Y = hm.mv(X)
    mv_ms = 1e3 * (time.perf_counter() - t0) / reps
    errs = []
    for c in range(4):
        ref = sd.matvec("H", X[:, c])
        errs.append(float(np.linalg.norm(Y[:, c] - ref) / np.linalg.norm(ref)))
    out = {"workload": "single-domain H of the %d^3 heated cube (%d x %d), leaf %d, eta %g, ACA type 2 tolerance %g"
                       % (n, nb, nb, leaf, eta, eps),
           "nbnodes": nb, "parts": hm.nparts, "max_rank": hm.max_rank, "stored_doubles": hm.stored,
           "compression": hm.stored / float(nb * nb), "build_s": build,
           "mv4_ms_with_transfers": mv_ms, "mv4_GBps_stored": hm.stored * 8 / (mv_ms * 1e-3) / 1e9,
           "rel_err_vs_dense": max(errs)}
    log("aca kinematics %d^3: %d parts, max rank %d, %.3f of dense, build %.2f s, 4-column product %.2f ms, "
        "error vs dense %.2e" % (n, hm.nparts, hm.max_rank, out["compression"], build, mv_ms, max(errs)))
    hm.close()
    sd.close()
    return out


def reference_leg(n=8, nnlit=10):
    """The reference itself (oracle/_ref/Tritok, compiled from its own sources)
    on the CPU-runnable configuration: the n^3 lid-driven cavity, Re = 100,
    URDW 0.1, from rest, nnlit Picard iterations on one host core (one MPI
    rank).  Its integral caches are written from this library's device
    integrals, so the reference's set-up reads them instead of integrating
    (about a minute on one core); the rate is its own CPU time of the Picard
    loop (tri.log).  The GPU runs the same case in FP64 and mixed precision,
    and its per-iteration changes are compared with the reference's tri.sta."""
    import shutil
    import tempfile
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import cpu_sample  # the reference binary: CPU baseline leg only
    from tritok_amd import files as F
    from tritok_amd import integrals as I
    from tritok_amd import workloads
    from tritok_amd.case import Case
    host = workloads.cavity(n, re=100.0, urdw=0.1, nnlit=nnlit)
    view = host.view()
    m = view.mesh
    q = I.quad(4, 4, 4, 4, 2)
    work = tempfile.mkdtemp(prefix="tt_reference_")
    try:
        mac = I.macro_integrals(view, q)
        F.write_macro_int(os.path.join(work, "tri.macro.int"), {**{k: mac[k] for k in F.MACRO}, "ont": mac["ont"]},
                          m.nnodes, m.nbnodes)
        del mac
        sd = I.sd_integrals(view, q)
        F.write_sdkm_int(os.path.join(work, "tri.sdkm.int.1"), {**{k: sd[k] for k in F.SDKM}, "ont": sd["ont"]},
                         m.nnodes, m.nbnodes)
        del sd
        host.write_geo(os.path.join(work, "case.geo"))
        log("reference: %d^3 cavity, %d Picard iterations on one core ..." % (n, nnlit))
        secs, sta_ref = cpu_sample.reference_run(work, None, workloads.bic_cavity(),
                                                 workloads.inp_text(re=100.0, urdw=0.1, nnlit=nnlit))
    finally:
        shutil.rmtree(work, ignore_errors=True)
    out = {"case": "lid-driven cavity %d^3 cells (%d nodes), Re = 100, URDW 0.1, from rest, %d Picard iterations"
                   % (n, m.nnodes, len(sta_ref)),
           "reference": {"binary": "oracle/_ref/Tritok (amdflang -O2, MPICH, 1 rank)", "cores": 1,
                         "picard_loop_s": secs, "iters_per_s": len(sta_ref) / secs}}
    for prec in ("fp64", "fp16"):
        with Case(host, prec=prec, quad=q) as c:
            sta = []
            for _ in range(2):
                rc, err, _nit = c.step()
                sta.append(err)
            t0 = time.perf_counter()
            k = 0
            for _ in range(len(sta_ref) - 2):
                rc, err, _nit = c.step()
                sta.append(err)
                k += 1
            dt = time.perf_counter() - t0
        sta = np.array(sta)[:len(sta_ref)]
        cols = [j for j in range(6) if np.abs(sta_ref[:, j]).max() > 0]
        rel = float(max(np.max(np.abs(sta[:, j] - sta_ref[:, j]) / np.abs(sta_ref[:, j])) for j in cols))
        out[prec] = {"iters_per_s": k / dt, "over_reference": (k / dt) / out["reference"]["iters_per_s"],
                     "max_rel_diff_per_iteration_changes": rel}
    log("reference: %.2f Picard its/s on one core; GPU fp64 %.1f, mixed %.1f its/s" % (
        out["reference"]["iters_per_s"], out["fp64"]["iters_per_s"], out["fp16"]["iters_per_s"]))
    return out


def mhsd_leg(n, f=7, tstep=0.1, nstep=2, nnlit=3):
    """The single-domain modified-Helmholtz path (EQNS imH 3, 5, 6 of the
    reference) on the n^3 cube: u prescribed on walls 1, 3, 5, the normal flux
    on 2, 4, 6, manufactured problem f (diff = 2 + sin(pi/2 xyz)); integrals,
    system and LU inverse once, then per compression mode (dense, wavelet,
    cross approximation) two time steps of three fixed-point solves each,
    with the device time of every solve and the RMS error against the
    analytic solution.  Dense solve roofline: the eight right-hand-side
    matrices of the equation rows, RHS and A^-1 streamed once per solve."""
    from tritok_amd import Host, mhsd, workloads
    from tritok_amd import integrals as I
    host = Host.cube(n).read_bic(workloads.bic_mh())
    t0 = time.perf_counter()
    pr = mhsd.prepare(host, f, tstep)
    prep_s = time.perf_counter() - t0
    cd = pr["cd"]
    m = pr["m"]
    nn, nb, nf, nunk, nknown = m["nn"], m["nb"], 4 * m["ne"], cd["nunk"], cd["nknown"]
    out = {"workload": "%d^3 cube, %d nodes + %d boundary flux points as sources, %d unknowns, manufactured problem "
                       "%d, %d steps x %d solves" % (n, nn, nf, nunk, f, nstep, nnlit),
           "host_prepare_s": prep_s}
    with mhsd.Operator(host.view(), I.quad(4, 4, 4, 4, 2), cd, pr["diff"], pr["qdiff"], pr["beta"]) as op:
        op.system()
        st = op.stats()
        out.update(integrals_ms=st["ms_integrals"], system_ms=st["ms_system"],
                   matrices_GB=8.0 * (nn + nf) * (5 * nn + 4 * nb + nf) / 1e9)
        dense_bytes = 8.0 * (nunk * (5 * nn + 3 * nb) + nunk * nknown + nunk * nunk)
        for mode, kw in (("dense", {}), ("wavelet", dict(wt_kappa=1e-5)), ("aca", dict(aca_type=2, aca_eps=1e-6))):
            ratio = op.compress(mode, **kw)
            cms = op.stats()["ms_compress"] if mode != "dense" else 0.0
            steps = mhsd.time_loop(op, pr, nstep, nnlit)
            ms = np.array([x for r in steps for x in r["solve_ms"]])
            rec = {"compress_ms": cms, "stored_over_full": ratio, "solve_ms_mean": float(ms.mean()),
                   "solve_ms_min": float(ms.min()), "rms_error": [r["rms"] for r in steps]}
            if mode == "dense":
                gbs = dense_bytes / (ms.min() * 1e-3) / 1e9
                rec["roofline"] = {"bound": "hbm", "bytes": dense_bytes, "achieved": gbs, "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
                u_dense = steps[-1]["u"]
            else:
                rec["max_rel_vs_dense"] = float(np.abs(steps[-1]["u"] - u_dense).max() / np.abs(u_dense).max())
            out[mode] = rec
            log("mhsd %d^3 %s: ratio %.3f, compress %.1f ms, solve %.3f ms (min %.3f), rms %s" % (
                n, mode, ratio, cms, rec["solve_ms_mean"], rec["solve_ms_min"], rec["rms_error"]))
    log("mhsd %d^3: integrals %.1f ms, system %.1f ms, matrices %.2f GB" % (n, out["integrals_ms"], out["system_ms"],
                                                                           out["matrices_GB"]))
    return out


FLOP_SDKM_DOMAIN = 3 * 27 * 2 + 15  # three gradient outputs x 27 shape functions, plus the kernel


def sdkm_regular_points(host, points_per_cell=512):
    """Regular (row, cell, point) evaluations of the single-domain domain
    integrals: every boundary row against every cell that does not hold its
    node, at the 8-point rule's 512 points per cell (GAUS 4)."""
    m = host.mesh_arrays()
    nb, nc = len(m["gbn"]), m["idc"].shape[0]
    cells_of = np.bincount(m["idc"].ravel() - 1, minlength=m["nn"])  # cells holding each node
    singular = int(cells_of[m["gbn"] - 1].sum())
    return (nb * nc - singular) * points_per_cell


def assembly_entry(host, tm, macro_stats):
    """Integral-assembly throughput of the case setup: device times of the
    macro and single-domain matrices, credited flop per evaluation as in
    the batched-integrals leg (singular transforms and geometry uncredited)."""
    pts = sdkm_regular_points(host)
    sd_tf = pts * FLOP_SDKM_DOMAIN / (tm["setup_sdkm"] * 1e-3) / 1e12
    mac = macro_stats["domain_points"] * FLOP_DOMAIN_MH + macro_stats["boundary_points"] * FLOP_BOUNDARY_MH
    mac_tf = mac / (tm["setup_macro"] * 1e-3) / 1e12 if tm["setup_macro"] > 0 else None
    log("assembly: macro %.0f ms (%.2f TFLOP/s credited), sdkm %.0f ms (%.3g regular points, %.1f TFLOP/s credited "
        "over the whole sdkm setup)" % (tm["setup_macro"], mac_tf or 0.0, tm["setup_sdkm"], pts, sd_tf))
    return {"macro_ms": tm["setup_macro"], "macro_points": {"domain": macro_stats["domain_points"],
                                                            "boundary": macro_stats["boundary_points"]},
            "macro_tflops": mac_tf, "sdkm_ms": tm["setup_sdkm"], "sdkm_regular_points": pts,
            "sdkm_tflops": sd_tf, "sdkm_frac_fp64_peak": sd_tf / FP64_PEAK_TFLOPS, "lu_ms": tm["setup_lu"],
            "flop_per_point": {"sdkm_domain": FLOP_SDKM_DOMAIN, "macro_domain": FLOP_DOMAIN_MH,
                               "macro_boundary": FLOP_BOUNDARY_MH}}


def assembly_only(n):
    """Macro and single-domain integrals of the n^3 heated cube through the
    routine-level entry points (no factorisation), for counter runs."""
    from tritok_amd import integrals as I
    from tritok_amd import workloads
    host = workloads.natural_convection(n, ra=1e5, pr=0.71)
    q = I.quad(4, 4, 4, 4, 2)
    t0 = time.perf_counter()
    I.macro_integrals(host.view(), q)
    t1 = time.perf_counter()
    I.sd_integrals(host.view(), q)
    t2 = time.perf_counter()
    log("assembly only %d^3: macro %.2f s, sdkm %.2f s (with transfers), %.3g sdkm regular points" % (
        n, t1 - t0, t2 - t1, sdkm_regular_points(host)))


def lsqr_matrix_bytes(info, value_bytes):
    """Bytes the two sparse products of one LSQR iteration would move in the
    reference's CSR form (values + int32 column indices, row pointers):
    2 * ((value_bytes + 4) * nnz + 4 * (m + 1)).  Reported beside the roofline
    as a CSR-equivalent rate; the block format moves less."""
    return 2.0 * ((value_bytes + 4) * info["nnz"] + 4 * (info["m"] + 1))


def cell_kernel_bytes(info, value_bytes, partial_bytes):
    """Algorithmic HBM bytes of one launch of the least-squares cell kernel on
    block storage: every cell's R x L block once (value_bytes per entry,
    padding not counted), its column map (int32), u read and u' written (FP64,
    m each), v read once (FP64, n) and the A^T u' partials written
    (partial_bytes per cell column)."""
    c, R, L = info["cells"], info["rows_per_cell"], info["cols_per_cell"]
    return float(c * R * L * value_bytes + c * L * 4 + 16 * info["m"] + 8 * info["n"] + c * L * partial_bytes)


def unknown_kernel_bytes(info, partial_bytes):
    """Algorithmic HBM bytes of one launch of the unknown kernel: the cell
    partials gathered (partial_bytes per cell column), the gather table (8
    int32 per unknown), x and w read and written, v' written (FP64)."""
    c, L, n = info["cells"], info["cols_per_cell"], info["n"]
    return float(c * L * partial_bytes + 32 * n + 40 * n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 16 timed steps after 2 warm-up steps hold exactly one FP64 anchor step of
    # the kinematics (period 16): the steady-state share of that work
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=20)
    ap.add_argument("--probe-stride", type=int, default=32)
    ap.add_argument("--anchor-period", type=int, default=16,
                    help="Picard iterations per FP64 anchor step of the reduced-precision kinematics")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the least-squares iterations from captured HIP graph chunks instead of launching "
                         "every kernel on the stream")
    ap.add_argument("--mixed", default="fp16", choices=["fp32", "fp16", "bf16"],
                    help="reduced precision of the mixed leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--reference-n", type=int, default=8,
                    help="n: time the reference binary on the n^3 cavity beside the GPU (0: skip)")
    ap.add_argument("--accuracy-n", type=int, default=20,
                    help="cube size of the converged mixed-vs-FP64 comparison (0: skip)")
    ap.add_argument("--no-kinematics", action="store_true",
                    help="EQNS iBw = 0: skip the single-domain kinematics (least-squares systems only)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the serving stream replayed in real time (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-batch", type=int, default=8,
                    help="cases of one kind a slot steps together (1: one case per slot)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving leg (mixed precision), then exit")
    ap.add_argument("--serve-max-iter", type=int, default=400)
    ap.add_argument("--batch-cells", type=int, default=16,
                    help="n: batched modified-Helmholtz integrals of n^3 independent cells (0: skip)")
    ap.add_argument("--batch-solves", type=int, default=1024,
                    help="independent cavity systems solved in one launch (0: skip)")
    ap.add_argument("--aca-kin", type=int, default=20,
                    help="n: ACA-compressed single-domain H of an n^3 cube (0: skip)")
    ap.add_argument("--mhsd", type=int, default=8,
                    help="n: single-domain modified-Helmholtz solves on an n^3 cube (0: skip)")
    ap.add_argument("--mhsd-only", action="store_true", help="only the --mhsd leg, then exit")
    ap.add_argument("--mixed-kin", type=int, default=24,
                    help="n: mixed-precision kinematics solve on an n^3 cube (0: skip)")
    ap.add_argument("--assembly-only", type=int, default=0,
                    help="n: only the macro and single-domain integrals of an n^3 cube, then exit (counter runs)")
    ap.add_argument("--mixed-kin-only", action="store_true",
                    help="only the mixed-precision kinematics leg (--mixed-kin n), then exit (profiler runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB of a device-to-device copy issued first (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process drives one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.assembly_only:
        assembly_only(args.assembly_only)
        return
    if args.serve_only:
        sv = serving_leg(args.serve_requests, args.serve_slots, args.serve_max_iter, (args.mixed,), args.serve_batch)
        for p in sv:
            if isinstance(sv[p], dict) and "per_request" in sv[p]:
                sv[p].pop("per_request")
        log("serving only: " + json.dumps(sv))
        return
    if args.mhsd_only:
        log("mhsd only: " + json.dumps(mhsd_leg(args.mhsd)))
        return
    if args.mixed_kin_only:
        mk = mixed_kinematics_leg(args.mixed_kin)
        log("mixed kinematics only: " + json.dumps({p: mk[p] for p in ("fp64", "fp32", "bf16", "fp16")}))
        return

    from tritok_amd import workloads
    from tritok_amd.case import Case
    from tritok_amd import device_info
    from tritok_amd import integrals as I

    dev = device_info()
    log("device", dev)
    if args.calibrate_copy:
        copy_for_calibration(args.calibrate_copy << 20)
    n = args.n
    t0 = time.perf_counter()
    host = workloads.natural_convection(n, ra=1e5, pr=0.71, ibw=0 if args.no_kinematics else 1)
    view = host.view()
    log("host preprocessing %.2fs  nnodes=%d nbnodes=%d cells=%d" % (time.perf_counter() - t0, view.mesh.nnodes,
                                                                     view.mesh.nbnodes, view.mesh.nicell))
    t0 = time.perf_counter()
    case = Case(host, prec=args.mixed)
    setup_s = time.perf_counter() - t0
    case.set_graphs(args.graphs)
    case.set_anchor_period(args.anchor_period)
    tm = case.timings()
    log("setup %.2fs (macro %.0f ms, sdkm %.0f ms, LU %.0f ms, systems %.0f ms), device bytes %.2f GB" % (
        setup_s, tm["setup_macro"], tm["setup_sdkm"], tm["setup_lu"], tm["setup_systems"],
        case.device_bytes() / 1e9))
    assembly = assembly_entry(host, tm, I.last_integral_stats()) if not args.no_kinematics else None

    legs = {}
    for prec in ("fp64", args.mixed):
        legs[prec] = run_leg(case, prec, args.steps, args.warmup, args.probe_stride)
        L = legs[prec]
        log("%s: %.3f s for %d steps (%.1f ms/step), LSQR iterations per step %s" % (
            prec, L["total_s"], args.steps, 1e3 * L["total_s"] / args.steps, L["nits"].sum(0)[3:10] / args.steps))
    f64, f32 = legs["fp64"]["fields"], legs[args.mixed]["fields"]
    diff = {k: rel_l2(f32[k], f64[k]) for k in FIELDS}
    log("mixed vs fp64 rel L2 after %d iterations: %s" % (args.warmup + args.steps, diff))

    mixed = legs[args.mixed]
    value = args.steps / mixed["total_s"]
    fp64_rate = args.steps / legs["fp64"]["total_s"]
    lat = np.sort(mixed["lat"]) * 1e3

    # roofline of the dominant kernel: lsqr_A on the 51-row systems (VTE x, y, z and T)
    half = args.mixed == "fp16"
    kname = "lsqr_Aw<_Float16,51,51,4>" if half else "lsqr_Aw<float,51,51,4>"
    vbytes = 2 if half else 4
    infos = [case.system_info(s) for s in range(4)]
    tuned = {"vorticity_energy": {k: infos[0][k] for k in ("grid_fp64", "grid_fp32", "grid_fp16")},
             "kinematics": {k: case.system_info(4)[k] for k in ("grid_fp64", "grid_fp32", "grid_fp16")}}
    nit = mixed["nits"].sum(0)
    w = np.array([nit[6], nit[7], nit[8], nit[9]], float)
    pbytes = 4  # FP16 and FP32 blocks store the A^T u' partials in FP32
    cell_bytes = float((w * np.array([cell_kernel_bytes(i, vbytes, pbytes) for i in infos])).sum() / max(w.sum(), 1.0))
    unk_bytes = float((w * np.array([unknown_kernel_bytes(i, pbytes) for i in infos])).sum() / max(w.sum(), 1.0))
    csr_bytes = float((w * np.array([lsqr_matrix_bytes(i, vbytes) for i in infos])).sum() / max(w.sum(), 1.0))
    avg_ms = mixed["probe51"]["avg_ms"]
    traffic = read_traffic(kname)  # measured HBM bytes per launch (profiles/)
    achieved = cell_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    # the whole LSQR iteration of the 51-row systems: the vorticity and energy
    # phases' device time over their iterations (includes their right-hand
    # sides, hot-start residuals and the host's batch synchronisations)
    ph_mean = {k: float(np.mean([p[k] for p in mixed["phases"]])) for k in ("vte", "temperature")}
    it_ms = (ph_mean["vte"] + ph_mean["temperature"]) / max(w.sum() / args.steps, 1.0)
    it_gbs = (cell_bytes + unk_bytes) / (it_ms * 1e-3) / 1e9
    roofline = {"kernel": "%s (51-row blocks)" % kname, "bound": "hbm",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None,
                "traffic": traffic,
                "traffic_GBps": traffic / (avg_ms * 1e-3) / 1e9 if traffic and avg_ms > 0 else None,
                "traffic_frac": traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic and avg_ms > 0 else None,
                "algorithmic_bytes_per_launch": cell_bytes,
                "bytes": "cells*R*L*2 (FP16 block values, unpadded) + cells*L*4 (column maps) + 16m (u read, u' "
                         "written) + 8n (v) + cells*L*4 (FP32 A^T u' partials)",
                "avg_launch_ms": avg_ms, "launches_timed": mixed["probe51"]["sampled"], "tuned_grid": tuned,
                "launches": mixed["probe51"]["launches"],
                "block_bytes_per_launch": float(infos[0]["cells"] * (2608 * 2 if half else 2604 * 4)),
                "csr_equivalent_bytes_per_launch": csr_bytes,
                "csr_equivalent_GBps": csr_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None,
                "lsqr_iteration": {"algorithmic_bytes": cell_bytes + unk_bytes,
                                   "unknown_kernel_bytes": unk_bytes, "ms": it_ms, "GBps": it_gbs,
                                   "frac": it_gbs / HBM_PEAK_GBS,
                                   "time": "vorticity + energy phase device time per step / their LSQR "
                                           "iterations per step"}}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import cpu_sample  # the oracle: CPU baseline leg only
        cpu = cpu_sample.picard_rate(host, case, legs["fp64"]["nits"], budget_s=args.cpu_seconds, log=log)
    ref_cpu = None
    if not args.no_cpu and args.reference_n > 0:
        ref_cpu = reference_leg(args.reference_n)

    case.close()  # free the 20^3 case before the other legs
    acc = accuracy_leg(args.accuracy_n, args.mixed, args.anchor_period) if args.accuracy_n > 0 else None
    mixkin = mixed_kinematics_leg(args.mixed_kin) if args.mixed_kin > 0 else None
    acakin = aca_kinematics_leg(args.aca_kin) if args.aca_kin > 0 else None
    mh_sd = mhsd_leg(args.mhsd) if args.mhsd > 0 else None
    batched = batched_integrals_leg(args.batch_cells) if args.batch_cells > 0 else None
    bsolves = batched_solves_leg(args.batch_solves) if args.batch_solves > 0 else None
    serving = (serving_leg(args.serve_requests, args.serve_slots, args.serve_max_iter, ("fp64", args.mixed),
                           args.serve_batch)
               if args.serve_requests > 0 else None)

    ph = mixed["phases"]
    out = {
        "metric": METRIC, "value": value, "unit": "Picard iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * mixed["total_s"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "%s/f32/f64" % {"fp16": "f16", "bf16": "bf16", "fp32": "f32"}[args.mixed],
        "data": "synthetic",
        "config": {"workload": "natural convection, differentially heated cube, %d^3 subdomains (%d nodes, %d "
                               "boundary nodes), Ra=1e5, Pr=0.71" % (n, view.mesh.nnodes, view.mesh.nbnodes),
                   "subdomains": n ** 3,
                   "precision": "%s least-squares blocks (FP32 accumulation) with FP64 residual anchors, "
                                "scalars and vectors; kinematics products in %s with per-row 2^k scales, FP64 anchor "
                                "step every %d iterations" % ("FP16" if half else "FP32", args.mixed.upper(),
                                                              args.anchor_period),
                   "anchor_steps_timed": sum(1 for i in range(args.warmup, args.warmup + args.steps)
                                             if i % max(1, args.anchor_period) == 0),
                   "lsqr_launch": "chunks of 2..64 iterations replayed from captured HIP graphs" if args.graphs
                                  else "kernel by kernel on the stream"},
        "fp64_iters_per_s": fp64_rate, "mixed_over_fp64": value / fp64_rate,
        "mixed_vs_fp64_rel_l2_converged": acc["max_rel_l2"] if acc else None, "accuracy": acc,
        "mixed_vs_fp64_rel_l2_after_steps": max(diff.values()),
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                            "max": float(lat[-1])},
        "phase_ms_mean": {k: float(np.mean([p[k] for p in ph])) for k in
                          ("boundary_vorticity", "km", "temperature", "vte", "convergence", "step_total")},
        "lsqr_iters_per_step": {"fp64": (legs["fp64"]["nits"].sum(0)[3:10] / args.steps).tolist(),
                                "mixed": (nit[3:10] / args.steps).tolist()},
        "setup_s": setup_s, "device": dev, "roofline": roofline, "cpu_baseline": cpu, "reference_cpu": ref_cpu,
        "p99_solve_latency_s": serving[args.mixed]["latency_s"]["p99"] if serving else None,
        "serving": serving, "batched_integrals": batched, "batched_solves": bsolves, "mixed_kinematics": mixkin,
        "aca_kinematics": acakin, "single_domain_mh": mh_sd,
        "assembly": assembly,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
