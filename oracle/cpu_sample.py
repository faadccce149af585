"""CPU baseline of one Picard iteration, timed on a bounded sample.

TEST INFRASTRUCTURE ONLY (the oracle): used by the cpu_baseline leg of
bench.py, never by the product package.

One steady Picard iteration of the reference (main.f90:1200-1589) spends its
time in three places, all of fixed cost per call (no data-dependent work):

  * sdkm_rhs (lib_sdkm.f:349-519): one pass over the seven dense kinematics
    matrices per equation, 3 equations;
  * dgesl (lib_Solve_v11.f:905-1037): one triangular solve pair per equation;
  * LSQR iterations (lib_Solve_v11.f:3128-3887) on the kinematics,
    vorticity-transport and energy systems: one CSR product and one
    transposed CSR product per iteration, plus a hot-start residual per solve.

The sample times the oracle's restatement of each on the full-size operands of
the workload (the real sparsity patterns of the subdomain systems, the real
dense dimensions) with synthetic values, for a subset of the dense rows and a
run of 600 LSQR iterations, and scales each to one Picard iteration with the
LSQR iteration counts the GPU run measured on the same workload.  The
right-hand-side builders (per-cell 51 x 27 products, under 2% of the work) are
not included.
"""
import ctypes as C
import os
import time

import numpy as np

import oracle as O


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _arr(ptr, n, dtype=np.int32):
    ctype = C.c_int32 if dtype == np.int32 else C.c_double
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()


def _csr_pattern(view, kind, e=0):
    """1-based CSR (ia, ja) of a subdomain system: kind 'vte' (51 rows per cell,
    node and flux columns), 'km' (27 rows, node columns)."""
    m = view.mesh
    nc, nn, nq = m.nicell, m.nnodes, m.nq
    idc = _arr(m.idc, nc * 27).reshape(27, nc).T - 1
    if kind == "km":
        code = _arr(view.kmu, nn)
        cols = code[idc]  # (nc, 27)
        rows_per_cell = 27
    else:
        wk = _arr(view.wkode, nn * 3).reshape(3, nn)[e]
        wq = _arr(view.wqkode, nq * 3).reshape(3, nq)[e]
        ibf = _arr(m.ibf, nc * 24).reshape(4, 6, nc).transpose(2, 1, 0).reshape(nc, 24) - 1
        cols = np.concatenate([wk[idc], wq[ibf]], axis=1)  # (nc, 51)
        rows_per_cell = 51
    unk = np.where(cols < 0, -cols, 0)  # 1-based unknown number or 0
    per_cell = [np.sort(r[r > 0]) for r in unk]
    ja_cell = [np.tile(p, rows_per_cell) for p in per_cell]
    counts = np.repeat([len(p) for p in per_cell], rows_per_cell)
    ia = np.concatenate([[1], 1 + np.cumsum(counts)]).astype(np.int32)
    ja = np.concatenate(ja_cell).astype(np.int32)
    n = int(unk.max())
    return ia, ja, nc * rows_per_cell, n


LSQR_ITS = (20, 6020)  # two runs of the sample solve; their difference is timed per iteration


def _lsqr_iter_seconds(ia, ja, m, n, its=LSQR_ITS):
    rng = np.random.default_rng(7)
    a = rng.uniform(-1.0, 1.0, len(ja))
    q = O.prelsqr2(m, n, ia, ja, a)
    b = rng.uniform(-1.0, 1.0, m)
    x0 = np.zeros(n)
    t = []
    for k in its:
        t0 = time.perf_counter()
        O.slvlsqr2(m, n, ia, ja, a, q, b, x0, k, 1e-30)  # tolerance out of reach: exactly k iterations
        t.append(time.perf_counter() - t0)
    per_iter = (t[1] - t[0]) / (its[1] - its[0])
    setup = max(t[0] - its[0] * per_iter, 0.0)  # hot-start residual and scaling per solve
    return per_iter, setup


def picard_rate(host, case, nits, budget_s=20.0, log=print):
    """CPU Picard iterations/s of the workload `host` (oracle, this machine's
    cores); nits: (steps, 12) LSQR counts per Picard step from the GPU run."""
    view = host.view()
    m = view.mesh
    nb, nn = m.nbnodes, m.nnodes
    cores = _cores()
    t_start = time.perf_counter()
    # ---- dense right-hand side: a row sample of the seven matrices
    rows = max(64, min(nb, int(2.4e9 // (4 * nb + 3 * nn))))  # up to about 19 GB of FP64 operands (all rows at 20^3)
    H = [np.full((rows, nb), 1e-3, order="F") for _ in range(4)]
    D = [np.full((rows, nn), 1e-3, order="F") for _ in range(3)]
    lbn, gbn = _arr(m.lbn, nn), _arr(m.gbn, nb)
    kode = np.zeros(nb * 3, np.int32)
    nx = np.ones(nb)
    vel, vort = np.ones((nn, 3), order="F"), np.ones((nn, 3), order="F")
    dfi = np.zeros(rows)
    d = O._d
    t0 = time.perf_counter()
    for eqn in (1, 2, 3):
        O._sdkm_rhs(eqn, nn, rows, O._i(lbn), O._i(gbn), O._i(kode), d(nx), d(nx), d(nx), *[d(x) for x in H],
                    *[d(x) for x in D], rows, d(vel), d(vort), d(dfi), cores)
    t_rhs = (time.perf_counter() - t0) / 3 * nb / rows
    del H, D
    log("cpu sample: sdkm_rhs %d of %d rows on %d threads -> %.3f s per equation" % (rows, nb, cores, t_rhs))
    # ---- dgesl on an nb x nb factor
    lu = np.full((nb, nb), 1e-6, order="F")
    np.fill_diagonal(lu, 1.0)
    piv = np.arange(1, nb + 1, dtype=np.float64)
    t0 = time.perf_counter()
    for _ in range(3):
        O.dgesl(lu, piv, np.ones(nb))
    t_sl = (time.perf_counter() - t0) / 3
    del lu
    log("cpu sample: dgesl n=%d -> %.3f s" % (nb, t_sl))
    # ---- LSQR iterations on the real patterns
    ia, ja, mrows, n = _csr_pattern(view, "vte", 0)
    t_vte, s_vte = _lsqr_iter_seconds(ia, ja, mrows, n)
    log("cpu sample: VTE LSQR m=%d n=%d nnz=%d -> %.4f s/iteration" % (mrows, n, len(ja), t_vte))
    ia, ja, mrows, n = _csr_pattern(view, "km")
    t_km, s_km = _lsqr_iter_seconds(ia, ja, mrows, n)
    log("cpu sample: KM LSQR m=%d n=%d nnz=%d -> %.4f s/iteration" % (mrows, n, len(ja), t_km))
    its = np.asarray(nits, float).mean(0)
    km_its, vte_its, t_its = its[3:6].sum(), its[6:9].sum(), its[9]
    per_iter = (3 * t_rhs + 3 * t_sl + km_its * t_km + 3 * s_km + (vte_its + t_its) * t_vte +
                (4 if view.iDT == 1 else 3) * s_vte)
    sample_s = time.perf_counter() - t_start
    log("cpu sample: %.2f s of sampling; %.2f s per Picard iteration" % (sample_s, per_iter))
    return {"value": 1.0 / per_iter, "unit": "Picard iterations/s", "cores": cores, "kind": "port",
            "sample": ("oracle (C restatement of the reference, -O2) on this host: sdkm_rhs over %d of %d dense "
                       "rows on %d OpenMP threads, three dgesl, %d LSQR iterations each on the vorticity and "
                       "kinematics systems' real CSR patterns (serial, as the reference), scaled by the GPU run's "
                       "mean LSQR counts per step (%.0f kinematics, %.0f vorticity, %.0f energy); %.1f s of CPU "
                       "work" % (rows, nb, cores, LSQR_ITS[1] - LSQR_ITS[0], km_its, vte_its, t_its, sample_s)),
            "seconds_per_iteration": per_iter}


REFBIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "Tritok")


def _ref_solve_seconds(log_path):
    """The 'solve' line of the reference's CPU-time table (tri.log): the
    seconds spent in its Picard loop (main.f90's timing, lib_Time.f90)."""
    with open(log_path) as f:
        for line in f:
            p = [x.strip() for x in line.split("|")]
            if len(p) >= 3 and p[1] == "solve":
                return float(p[2])
    raise RuntimeError("no solve time in " + log_path)


def _ref_sta(path):
    """Rows of the reference's tri.sta (step, iteration, nine relative changes)."""
    rows = []
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) >= 11 and p[0].isdigit() and p[1].isdigit():
                rows.append([float(x) for x in p[2:11]])
    return np.array(rows)


def reference_run(workdir, geo_text, bic_text, inp_text, timeout_s=120.0):
    """Runs the reference binary (compiled from the reference's own sources by
    oracle/Makefile.ref) in workdir, which must hold the integral caches
    tri.macro.int and tri.sdkm.int.1 (so that its set-up reads them instead of
    integrating); returns its Picard-loop seconds and its tri.sta rows.
    One MPI rank, serial Fortran: one host core."""
    import subprocess
    if not os.path.exists(REFBIN):
        raise RuntimeError("reference binary not built (oracle/_ref/Tritok)")
    if geo_text is not None:
        with open(os.path.join(workdir, "case.geo"), "w") as f:
            f.write(geo_text)
    with open(os.path.join(workdir, "case.bic"), "w") as f:
        f.write(bic_text)
    with open(os.path.join(workdir, "tri.inp"), "w") as f:
        f.write(inp_text.replace("MDIR ./\n", "MDIR ./\nLGEO case.geo\nLBIC case.bic\n"))
    subprocess.run([REFBIN], cwd=workdir, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                   timeout=timeout_s)
    return _ref_solve_seconds(os.path.join(workdir, "tri.log")), _ref_sta(os.path.join(workdir, "tri.sta"))
