"""Benchmark: single-source BFS GTEPS on a symmetrised RMAT-22 graph (Int32
CSR resident in HBM), with the step kernel's HBM roofline, plus exact
betweenness TEPS on RMAT-20 with 1024 explicit sources, each beside the
CPU oracle timed on this host.

A "step" is one full BFS (distances + parents) from one of 64 seeded roots
of degree >= 1 (cycled).  value = sum over timed BFS runs of m_comp divided
by the wall time of the timed region, where m_comp = (sum of degrees of the
reached vertices) / 2.

Further legs of the default run, each its own key of the JSON line: the
uneven 4096-graph betweenness batch ("batch"), the learned direction switch
against the alpha/beta rule on its three graphs ("learned_switch") and the
per-call drop-in cost on RMAT-22 ("dropin").

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-bc] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "lightgraphs.jl_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "BFS GTEPS (RMAT-22) + exact-BC TEPS; learned vs α/β switch, same GPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    """Threads of the threaded CPU legs: the CPUs this process may run on
    (its affinity mask), capped by OMP_NUM_THREADS where the host sets it
    (a GPU host that shares its cores between GPUs states each GPU's share
    there)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def cpu_median(fn, runs=3):
    """One untimed warm-up run of fn, then the median wall time of `runs`
    timed runs (seconds) and the list of those times."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def step_bytes(st, nv):
    """Algorithmic HBM bytes of one BFS step (level L -> L+1).

    bottom-up: V/8 visited + V/8 frontier words + V/8 next-frontier write
               + 8 per examined vertex (row_ptr pair) + 4 per adjacency entry
               read until the first frontier hit + 8 per new vertex
               (distance + parent)
    top-down:  12 per frontier vertex (queue entry + row_ptr pair) + 4 per
               expanded edge + V/8 visited bitmap + 12 per new vertex
               (distance, parent, queue entry)
    """
    if st.dir == 1:
        return 3 * nv / 8 + 8 * st.cand + 4 * st.scanned + 8 * st.n_next
    return 12 * st.n_front + 4 * st.m_front + nv / 8 + 12 * st.n_next


def bfs_bench(args, lgmi, _lib, orc):
    lib = _lib.load()
    t0 = time.time()
    g = lgmi.generators.rmat(22, 16, A=0.57, B=0.19, C=0.19, seed=1)
    rp, col = g.csr()
    nv, nnz = g.nv(), len(col)
    deg = np.diff(rp).astype(np.int64)
    log(f"RMAT-22 built: nv={nv} nnz={nnz} ({time.time() - t0:.1f}s)")
    h = g.handle()
    rng = np.random.default_rng(20260101)
    roots = rng.choice(np.nonzero(deg > 0)[0], size=64, replace=False).astype(np.int32)

    # m_comp per root (untimed): sum of degrees of reached vertices / 2
    mcomp = np.zeros(len(roots), dtype=np.float64)
    dist = np.empty(nv, dtype=np.int32)
    st = _lib.BfsStats()
    for i, r in enumerate(roots):
        _lib.check(lib.lgmi_bfs_run(h, int(r), 0, 1, _lib.ctypes.byref(st)))
        _lib.check(lib.lgmi_bfs_fetch(h, _lib.i32p(dist), None))
        mcomp[i] = deg[dist >= 0].sum() / 2.0
    log(f"m_comp per root: min {mcomp.min():.3e} max {mcomp.max():.3e}")

    # The K timed searches go through lgmi_bfs_run_many: eight searches at a
    # time, each on its own stream and search state, each writing distances
    # and parents into its own rows of caller-owned device arrays (up to 256
    # searches per call, the arrays reused from call to call).
    K = args.steps
    order = np.array([roots[i % len(roots)] for i in range(max(K, args.warmup))], dtype=np.int32)
    chunk = max(1, min(K, 256))
    dbuf = _lib.DeviceBuffer(4 * nv * chunk)
    pbuf = _lib.DeviceBuffer(4 * nv * chunk)

    def run_many(n):
        for s0 in range(0, n, chunk):
            part = np.ascontiguousarray(order[s0:min(n, s0 + chunk)])
            _lib.check(lib.lgmi_bfs_run_many(h, _lib.i32p(part), len(part), 0, dbuf.ptr, pbuf.ptr))

    run_many(args.warmup)
    _lib.check(lib.lgmi_synchronize())
    edges = float(sum(mcomp[i % len(roots)] for i in range(K)))
    t_start = time.perf_counter()
    run_many(K)
    _lib.check(lib.lgmi_synchronize())
    elapsed = time.perf_counter() - t_start
    gteps = edges / elapsed / 1e9
    log(f"timed {K} BFS (lgmi_bfs_run_many): {elapsed * 1e3:.2f} ms total, {gteps:.2f} GTEPS")
    # the rows of the last call against the per-call search (untimed)
    n_last = K - (K - 1) // chunk * chunk
    rows_d = dbuf.to_host(np.int32, (n_last, nv))
    rows_p = pbuf.to_host(np.int32, (n_last, nv))
    par_one = np.empty(nv, dtype=np.int32)
    for j in sorted({0, n_last - 1}):
        _lib.check(lib.lgmi_bfs_run(h, int(order[K - n_last + j]), 0, 1, None))
        _lib.check(lib.lgmi_bfs_fetch(h, _lib.i32p(dist), _lib.i32p(par_one)))
        if not (np.array_equal(rows_d[j], dist) and np.array_equal(rows_p[j], par_one)):
            raise RuntimeError("lgmi_bfs_run_many row differs from the per-call search")
    del rows_d, rows_p
    dbuf.free()
    pbuf.free()

    # the same searches one call at a time (lgmi_bfs_run per root)
    for i in range(args.warmup):
        _lib.check(lib.lgmi_bfs_run(h, int(roots[i % len(roots)]), 0, 1, None))
    _lib.check(lib.lgmi_synchronize())
    t_pc = time.perf_counter()
    for i in range(K):
        _lib.check(lib.lgmi_bfs_run(h, int(roots[i % len(roots)]), 0, 1, None))
    _lib.check(lib.lgmi_synchronize())
    elapsed_pc = time.perf_counter() - t_pc
    gteps_per_call = edges / elapsed_pc / 1e9
    log(f"timed {K} BFS one call each: {elapsed_pc * 1e3:.2f} ms total, {gteps_per_call:.2f} GTEPS")

    # the same timed runs with the learned direction switch (outputs are the
    # same by construction; the GPU tests check it)
    gteps_learned = None
    from lgmi import policy
    params, _ = policy.load_weights()
    if params is not None:
        try:
            policy.set_learned(params)
            for i in range(args.warmup):
                _lib.check(lib.lgmi_bfs_run(h, int(roots[i % len(roots)]), 0, 1, None))
            _lib.check(lib.lgmi_synchronize())
            t_l = time.perf_counter()
            for i in range(K):
                _lib.check(lib.lgmi_bfs_run(h, int(roots[i % len(roots)]), 0, 1, None))
            _lib.check(lib.lgmi_synchronize())
            gteps_learned = edges / (time.perf_counter() - t_l) / 1e9
            log(f"timed {K} BFS with the learned switch, one call each: {gteps_learned:.2f} GTEPS")
        finally:
            policy.set_heuristic()

    # per-root device times (HIP events) for the harmonic mean and median
    per_root = []
    for i, r in enumerate(roots):
        _lib.check(lib.lgmi_bfs_run(h, int(r), 0, 1, _lib.ctypes.byref(st)))
        per_root.append(mcomp[i] / (st.device_ms * 1e-3) / 1e9)
    per_root = np.array(per_root)

    # roofline of the step kernel: HIP events around every step launch
    _lib.check(lib.lgmi_bfs_time_steps(h, 1))
    tot_bytes = tot_ms = 0.0
    bu_bytes = bu_ms = 0.0
    launches = 0
    steps_buf = (_lib.BfsStep * 4096)()
    n_td = n_bu = 0
    for r in roots:
        _lib.check(lib.lgmi_bfs_run(h, int(r), 0, 1, _lib.ctypes.byref(st)))
        n = lib.lgmi_bfs_steps(h, 4096, steps_buf)
        if n < 0:
            _lib.check(n)
        launches += st.launches
        tot_ms += st.steps_event_ms
        for k in range(n):
            s = steps_buf[k]
            b = step_bytes(s, nv)
            tot_bytes += b
            if s.dir == 1:
                bu_bytes += b
                bu_ms += s.event_ms
                n_bu += 1
            else:
                n_td += 1
    _lib.check(lib.lgmi_bfs_time_steps(h, 0))
    achieved = tot_bytes / (tot_ms * 1e-3) / 1e9 if tot_ms > 0 else 0.0
    bu_achieved = bu_bytes / (bu_ms * 1e-3) / 1e9 if bu_ms > 0 else 0.0
    log(f"level kernel: {launches} launches, avg {tot_ms / max(launches, 1) * 1e3:.1f} us, "
        f"{tot_bytes / max(launches, 1) / 1e6:.2f} MB/launch, {achieved:.0f} GB/s; "
        f"bottom-up steps {n_bu}: {bu_achieved:.0f} GB/s; top-down steps {n_td}")

    # CPU baselines on a sample of the roots, each 1 warm-up pass then the
    # median of three: (a) the reference's threaded path for distances and
    # parents, Parallel.gdistances + Parallel.bfs_tree, on the host's thread
    # share (the reported baseline); (b) the serial one-pass BFS with
    # distances and parents on one core
    sample_roots = roots[: args.cpu_roots]
    e_cpu = float(mcomp[: len(sample_roots)].sum())
    nthreads = cpu_threads()
    t_gd = []

    def cpu_threaded():
        t_gd.clear()
        for r in sample_roots:
            t0 = time.perf_counter()
            orc.par_gdistances_threaded(rp, col, [int(r)], nthreads)
            t_gd.append(time.perf_counter() - t0)
            orc.par_bfs_tree_threaded(rp, col, int(r), nthreads)

    t_cpu, t_runs = cpu_median(cpu_threaded)
    cpu_gteps = e_cpu / t_cpu / 1e9
    gd_gteps = e_cpu / sum(t_gd) / 1e9
    log(f"cpu baseline (threaded reference path, {nthreads} threads): {len(sample_roots)} roots, median of 3 passes "
        f"{t_cpu:.2f}s ({', '.join(f'{t:.2f}' for t in t_runs)}) -> {cpu_gteps:.4f} GTEPS; "
        f"Parallel.gdistances alone {gd_gteps:.4f} GTEPS")

    def cpu_serial():
        for r in sample_roots:
            orc.bfs_dist_parent(rp, col, int(r))

    t_ser, t_ser_runs = cpu_median(cpu_serial)
    ser_gteps = e_cpu / t_ser / 1e9
    log(f"cpu serial one-pass BFS: median of 3 passes {t_ser:.2f}s "
        f"({', '.join(f'{t:.2f}' for t in t_ser_runs)}) -> {ser_gteps:.4f} GTEPS")

    traffic = read_traffic()
    from lgmi import tuning
    tuning_values = tuning.get()
    result = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / K, 5),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (project RMAT generator, seed 1)",
        "config": {
            "workload": "single-source BFS (distances + parents), symmetrised RMAT scale 22, edge factor 16, "
                        "a/b/c = 0.57/0.19/0.19, self-loops and duplicates removed, Int32 CSR; "
                        "64 seeded roots of degree >= 1",
            "nv": nv, "nnz": nnz, "roots": len(roots), "policy": "alpha=15/beta=18 heuristic",
            "entry_point": "lgmi_bfs_run_many (8 searches at a time; rows checked against lgmi_bfs_run)",
            "gteps_per_call": round(gteps_per_call, 4),
            "gteps_harmonic_mean_device": round(float(len(per_root) / np.sum(1.0 / per_root)), 3),
            "gteps_median_device": round(float(np.median(per_root)), 3),
            "gteps_learned_switch_per_call": round(gteps_learned, 4) if gteps_learned is not None else None,
            "tuning": tuning_values,
        },
        "roofline": {
            "kernel": "bfs_level",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": round(tot_bytes / max(launches, 1), 1),
            "avg_launch_us": round(tot_ms / max(launches, 1) * 1e3, 3),
            "bottom_up_achieved": round(bu_achieved, 1),
            "bottom_up_frac": round(bu_achieved / HBM_PEAK_GBS, 4),
        },
        "cpu_baseline": {
            "value": round(cpu_gteps, 5),
            "unit": "GTEPS",
            "cores": nthreads,
            "kind": "port",
            "sample": f"{len(sample_roots)} of the 64 roots; the reference's threaded path for distances and parents "
                      f"(Parallel.gdistances, queue_segment_size 20, then Parallel.bfs_tree) on {nthreads} threads; "
                      f"1 warm-up pass, median of 3 passes ({t_cpu:.2f} s per pass)",
            "parallel_gdistances_only": round(gd_gteps, 5),
            "serial_one_pass": {"value": round(ser_gteps, 5), "cores": 1,
                                "sample": "the same roots, serial BFS writing distances and parents in one pass; "
                                          f"median of 3 passes ({t_ser:.2f} s per pass)"},
        },
    }
    return result


def read_traffic(name="bfs_level_rmat22_pmc.json"):
    """HBM bytes per launch of a kernel from its committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or
    None when no summary is present."""
    p = os.path.join(REPO, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def bc_bench(args, lgmi, _lib, orc):
    t0 = time.time()
    g = lgmi.generators.rmat(20, 16, seed=1)
    rp, col = g.csr()
    nv, nnz = g.nv(), len(col)
    deg = np.diff(rp)
    log(f"RMAT-20 built: nv={nv} nnz={nnz} ({time.time() - t0:.1f}s)")
    rng = np.random.default_rng(4242)
    srcs = rng.choice(np.nonzero(deg > 0)[0], size=1024, replace=False).astype(np.int64)
    g.handle()
    from lgmi.centrality import last_stats
    lgmi.betweenness_centrality(g, srcs + 1)               # warmup: allocations, graph capture
    reps = 3
    walls, stats = [], []
    for _ in range(reps):
        t1 = time.perf_counter()
        bc = lgmi.betweenness_centrality(g, srcs + 1)
        walls.append(time.perf_counter() - t1)
        stats.append(last_stats(g))
    wall = float(np.mean(walls))
    stt = {k: float(np.mean([s[k] for s in stats])) for k in ("device_ms", "forward_ms", "backward_ms")}
    launch = {0: "one launch per level", 1: "one persistent launch",
              2: "one persistent launch replayed from a captured hipGraph"}[int(stats[-1]["launch"])]
    m = nnz / 2
    teps = 1024 * m / wall
    log(f"BC 1024 sources ({launch}): {wall * 1e3:.1f} ms wall (mean of {reps}), device {stt['device_ms']:.1f} ms "
        f"(fwd {stt['forward_ms']:.1f}, bwd {stt['backward_ms']:.1f}), {teps / 1e9:.2f} GTEPS_BC")
    # algorithmic bytes per source: 8 B sigma read (forward) + 8 B coefficient
    # read (backward) per shortest-path DAG edge, 24 B per reached vertex
    # (sigma write, sigma read + coefficient write); DAG edges counted exactly
    # on a sample of the sources with the oracle's BFS distances
    src_i = np.repeat(np.arange(nv, dtype=np.int64), np.diff(rp))
    per_src = []
    for s0 in srcs[:16]:
        d = orc.gdistances(rp, col, [int(s0)]).astype(np.int64)
        du, dv = d[src_i], d[col]
        e_dag = int(np.count_nonzero((du >= 0) & (dv == du + 1)))
        per_src.append(16 * e_dag + 24 * int(np.count_nonzero(d >= 0)))
    bytes_alg = 1024 * float(np.mean(per_src))
    bc_gbs = bytes_alg / (stt["device_ms"] * 1e-3) / 1e9
    log(f"BC algorithmic bytes {bytes_alg / 1e9:.1f} GB (DAG-edge model, 16-source sample) -> {bc_gbs:.0f} GB/s "
        f"= {bc_gbs / HBM_PEAK_GBS:.3f} of HBM peak")
    # CPU baseline: threaded oracle on a sample of the sources, scaled; one
    # warm-up pass, then the median of three
    nthreads = cpu_threads()
    ns = args.cpu_bc_sources
    box = {}

    def cpu_pass():
        box["want"] = orc.betweenness_threaded(rp, col, rp, False, srcs[:ns], nthreads=nthreads,
                                               index_quirk=False)

    t_cpu, t_runs = cpu_median(cpu_pass)
    want = box["want"]
    cpu_teps = ns * m / t_cpu
    # parity spot check on the sample: rescale to the 1024-source k
    got = lgmi.betweenness_centrality(g, srcs[:ns] + 1)
    floor = 1e-15 * max(1.0, float(np.abs(want).max()))
    err = float(np.max(np.abs(got - want) / np.maximum(np.abs(want), floor)))
    log(f"BC cpu baseline: {ns} sources, {nthreads} threads, median of 3 passes {t_cpu:.2f}s "
        f"({', '.join(f'{t:.2f}' for t in t_runs)}) -> {cpu_teps / 1e9:.4f} GTEPS_BC; sample max rel err {err:.2e}")
    g._release()
    return {
        "workload": "exact Brandes betweenness, symmetrised RMAT scale 20 edge factor 16, "
                    "1024 explicit seeded sources in 64-lane groups, FP64",
        "value": round(teps / 1e9, 4),
        "unit": "GTEPS_BC (sources x undirected edges / s)",
        "wall_ms": round(wall * 1e3, 2),
        "device_ms": round(stt["device_ms"], 2),
        "forward_ms": round(stt["forward_ms"], 2),
        "backward_ms": round(stt["backward_ms"], 2),
        "launch": launch,
        "persistent_work_groups": {"launched": int(stats[-1]["blocks"]), "participants": int(stats[-1]["participants"])},
        "roofline": {"bound": "hbm", "kernel": "bc_persist" if stats[-1]["launch"] else "bc_forward/bc_backward",
                     "achieved": round(bc_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bc_gbs / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("bc_persist_rmat20_pmc.json") if stats[-1]["launch"] else None,
                     "bytes_model": "per source 16 B per shortest-path DAG edge + 24 B per reached vertex, "
                                    "DAG edges from a 16-source sample"},
        "sample_max_rel_err": err,
        "cpu_baseline": {"value": round(cpu_teps / 1e9, 5), "unit": "GTEPS_BC", "cores": nthreads,
                         "kind": "port",
                         "sample": f"{ns} of the 1024 sources, threaded oracle Brandes; 1 warm-up pass, "
                                   f"median of 3 passes ({t_cpu:.2f} s per pass)"},
        "speedup_vs_cpu": round(teps / cpu_teps, 1),
    }


LDS_BYTES_PER_CLK_CU = 128      # LDS bytes per clock per CU (CDNA3's rate, a floor for CDNA4)
ENGINE_CLOCK_GHZ = 2.4          # MI355X peak engine clock
CU_COUNT = 256


def config4_bench(args, lgmi, orc, tune=False):
    """4096 graphs, n log-uniform in [200, 20000], half barabasi_albert(n, 4),
    half watts_strogatz(n, 8, 0.1), seeded per graph: all-source exact
    betweenness (normalize=true, endpoints=false) for every graph in one
    persistent launch, replayed from a captured hipGraph, with the
    autotuner's degree-bin and lane-group thresholds (batch_tuning.json, or
    a fresh autotune with --config4); the CPU oracle runs a seeded
    subsample and its time is scaled by sum n*m."""
    from lgmi import batch
    t0 = time.time()
    graphs = lgmi.generators.uneven_batch(args.config4_graphs, 200, 20_000, seed=1)
    nm = [g.nv() * (len(g.csr()[1]) // 2) for g in graphs]
    nm_tot = float(sum(nm))
    sum_n = sum(g.nv() for g in graphs)
    sum_n2 = float(sum(float(g.nv()) ** 2 for g in graphs))
    sum_n_nnz = float(sum(float(g.nv()) * len(g.csr()[1]) for g in graphs))
    log(f"batch built: {len(graphs)} graphs, sum n {sum_n}, sum n*m {nm_tot:.3e} ({time.time() - t0:.1f}s)")
    if tune:
        tuned = batch.autotune(graphs, log=log)
        if args.save_tuning:
            batch.save_tuning(tuned, args.save_tuning)
        best = tuned["best"]
        cands = tuned["candidates"]
    else:
        batch.load_tuning()
        with open(batch.TUNING_PATH) as f:
            best = json.load(f)["best"]
        cands = None
    runs = []
    for _ in range(2):
        t1 = time.perf_counter()
        res, ms = lgmi.betweenness_centrality_batch(graphs, return_ms=True)
        runs.append((ms, time.perf_counter() - t1))
    ms = runs[-1][0]
    wall = runs[-1][1]
    log(f"batch: device {ms:.1f} ms, wall {wall * 1e3:.1f} ms (first run {runs[0][0]:.1f} ms), "
        f"{nm_tot / (ms * 1e-3) / 1e9:.2f} G n*m/s")
    # LDS traffic model of the Brandes passes (per source): forward, 8 B
    # state read per adjacency entry plus a 16 B atomic add for about half
    # of them (the entries towards the next level); backward, 8 B state read
    # per adjacency entry; per vertex 32 B (level list, sigma, coefficient,
    # partial).  Adjacency entries come from L2 (a graph's CSR stays
    # cached while its sources run): 4 B per entry and pass.
    lds_bytes = 24.0 * sum_n_nnz + 32.0 * sum_n2
    l2_bytes = 8.0 * sum_n_nnz
    lds_peak = LDS_BYTES_PER_CLK_CU * ENGINE_CLOCK_GHZ * CU_COUNT          # GB/s
    lds_gbs = lds_bytes / (ms * 1e-3) / 1e9
    # CPU sample: systematic sampling with probability proportional to the
    # work n*m (graphs in random order, one at each of k equally spaced
    # points of the cumulative work), so the sample follows where the time
    # goes; estimator: the mean per-work time of the sample times the total
    # work.  Each sampled graph: one warm-up run, then the median of three.
    k = min(args.cpu_batch_graphs, len(graphs))
    order = np.random.default_rng(7).permutation(len(graphs))
    cum = np.cumsum(np.array([nm[i] for i in order], dtype=np.float64))
    pts = (np.arange(k) + 0.5) / k * cum[-1]
    idx = [int(order[j]) for j in np.searchsorted(cum, pts)]
    nthreads = cpu_threads()
    per_work, t_cpu, maxerr = [], 0.0, 0.0
    for i in idx:
        rp, col = graphs[i].csr()
        box = {}

        def run_one():
            box["want"] = orc.betweenness_threaded(rp, col, rp, False, np.arange(graphs[i].nv()), nthreads=nthreads,
                                                   index_quirk=False)

        t_i, _ = cpu_median(run_one)
        t_cpu += t_i
        per_work.append(t_i / nm[i])
        want = box["want"]
        floor = 1e-15 * max(1.0, float(np.abs(want).max()))
        maxerr = max(maxerr, float((np.abs(res[i] - want) / np.maximum(np.abs(want), floor)).max()))
    cpu_s = float(np.mean(per_work)) * nm_tot
    log(f"cpu oracle: {len(idx)} graphs (work-proportional sample, n {sorted(graphs[i].nv() for i in idx)}), "
        f"{nthreads} threads, {t_cpu:.2f}s -> {cpu_s:.0f}s estimated; max rel err {maxerr:.2e}")
    out = {
        "workload": "4096 graphs, n log-uniform in [200, 20000], half barabasi_albert(n, 4), half "
                    "watts_strogatz(n, 8, 0.1); all-source exact betweenness, FP64, one persistent launch "
                    "replayed from a captured hipGraph",
        "value": round(nm_tot / (ms * 1e-3) / 1e9, 4), "unit": "G (sum n*m)/s",
        "device_ms": round(ms, 2), "wall_ms": round(wall * 1e3, 2), "sum_nm": nm_tot,
        "tuning": best,
        "roofline": {"bound": "lds", "kernel": "bc_batch", "achieved": round(lds_gbs, 1), "peak": round(lds_peak, 1),
                     "unit": "GB/s", "frac": round(lds_gbs / lds_peak, 4), "traffic": None,
                     "l2_adjacency_gbs": round(l2_bytes / (ms * 1e-3) / 1e9, 1),
                     "bytes_model": "LDS: per source 24 B per adjacency entry + 32 B per vertex; peak "
                                    f"{LDS_BYTES_PER_CLK_CU} B/clk/CU x {CU_COUNT} CUs x {ENGINE_CLOCK_GHZ} GHz"},
        "sample_max_rel_err": maxerr,
        "cpu_baseline": {"value": round(nm_tot / cpu_s / 1e9, 5), "unit": "G (sum n*m)/s", "cores": nthreads,
                         "kind": "port",
                         "sample": f"{len(idx)} graphs of the batch drawn with probability proportional to n*m "
                                   f"(systematic, seeded), threaded oracle Brandes, each 1 warm-up + median of 3 "
                                   f"({t_cpu:.2f} s in all); total = mean time per n*m x sum n*m"},
        "speedup_vs_cpu": round(cpu_s / (ms * 1e-3), 1),
    }
    if cands is not None:
        out["autotune_candidates"] = cands
    return out


def config5_bench(args, lgmi, _lib):
    """The learned direction switch against the alpha=15 / beta=18 rule on
    the three graphs it is specified on (directed RMAT-23 with edge factor
    16, barabasi_albert(8_000_000, 8), the 4096 x 4096 grid), with the
    committed network: per graph `config5_roots` of the held-out roots the
    training run kept aside, each BFS run under both policies, outputs
    compared bit for bit, device times summed."""
    import zlib
    from lgmi import policy
    params, meta = policy.load_weights()
    if params is None:
        return None
    gen = lgmi.generators
    specs = [("rmat23_directed", lambda: gen.rmat(23, 16, seed=23, directed=True)),
             ("ba_8M_8", lambda: gen.barabasi_albert(8_000_000, 8, seed=8)),
             ("grid_4096x4096", lambda: gen.grid([4096, 4096]))]
    n_train = {"rmat23_directed": 2000, "ba_8M_8": 2000, "grid_4096x4096": 40}
    out = {"graphs": {}}
    tot_h = tot_l = 0.0
    same_all = True
    for name, make in specs:
        t0 = time.time()
        g = make()
        rp = g.csr()[0]
        rng = np.random.default_rng(zlib.crc32(name.encode()))   # the training run's root draw
        cand = np.nonzero(np.diff(rp) > 0)[0]
        ntr = n_train[name]
        pick = rng.choice(cand, size=min(len(cand), ntr + 200), replace=False)
        roots = pick[ntr:ntr + args.config5_roots]
        g.handle()
        policy.evaluate(g, roots[:2], params)                    # warm-up
        th, tl, same = policy.evaluate(g, roots, params)
        g._release()
        out["graphs"][name] = {"nv": g.nv(), "roots": int(len(roots)), "sum_ms_heuristic": round(float(th.sum()), 3),
                               "sum_ms_learned": round(float(tl.sum()), 3),
                               "ratio": round(float(th.sum() / tl.sum()), 4), "outputs_identical": bool(same)}
        tot_h += float(th.sum())
        tot_l += float(tl.sum())
        same_all = same_all and same
        log(f"config 5 {name}: heuristic {th.sum():.1f} ms, learned {tl.sum():.1f} ms, ratio "
            f"{th.sum() / tl.sum():.3f}, identical={same} ({time.time() - t0:.1f}s)")
    out["ratio_all"] = round(tot_h / tot_l, 4)
    out["outputs_identical"] = same_all
    out["workload"] = ("learned MLP vs alpha=15/beta=18 direction switch; held-out roots of the training draw; "
                       "device time of full BFS runs (distances + parents)")
    return out


def dropin_bench(lgmi, _lib):
    """The per-call drop-in path on RMAT-22: a caller that holds only a
    host CSR (what the Julia shim builds from fadjlist) creates a handle
    (upload and derived structures), runs one BFS with host outputs
    (lgmi_bfs: device search + the dist / parent copy back) and destroys
    the handle; beside it the same BFS on a resident handle."""
    import ctypes
    lib = _lib.load()
    g = lgmi.generators.rmat(22, 16, seed=1)
    rp, col = g.csr()
    nv = g.nv()
    root = int(np.argmax(np.diff(rp)))
    d = np.empty(nv, np.int32)
    p = np.empty(nv, np.int32)
    rows = []
    for rep in range(3):
        h = ctypes.c_void_p()
        t0 = time.perf_counter()
        _lib.check(lib.lgmi_graph_create(nv, len(col), _lib.i32p(rp), _lib.i32p(col), 0, None, None, 0,
                                         ctypes.byref(h)))
        _lib.check(lib.lgmi_synchronize())
        t1 = time.perf_counter()
        _lib.check(lib.lgmi_bfs(h, root, 0, _lib.i32p(d), _lib.i32p(p)))
        t2 = time.perf_counter()
        _lib.check(lib.lgmi_bfs(h, root, 0, _lib.i32p(d), _lib.i32p(p)))
        t3 = time.perf_counter()
        lib.lgmi_graph_destroy(h)
        t4 = time.perf_counter()
        rows.append((t1 - t0, t2 - t1, t3 - t2, t4 - t3))
    med = np.median(np.array(rows), axis=0) * 1e3
    out = {"graph": "RMAT-22 (host CSR, 128.3M entries)", "create_ms": round(float(med[0]), 2),
           "first_bfs_host_outputs_ms": round(float(med[1]), 2), "resident_bfs_host_outputs_ms": round(float(med[2]), 2),
           "destroy_ms": round(float(med[3]), 2),
           "per_call_ms": round(float(med[0] + med[1] + med[3]), 2),
           "note": "median of 3; per_call = create + first BFS with host outputs + destroy"}
    log(f"drop-in per call on RMAT-22: create {med[0]:.1f} ms, first BFS with host outputs {med[1]:.1f} ms, "
        f"resident BFS with host outputs {med[2]:.1f} ms, destroy {med[3]:.1f} ms")
    return out


def profile_pass(lgmi, _lib, nbfs, levels_out=""):
    """Workload for rocprofv3 runs: one 256 MiB device-to-device copy (the
    runtime's copy kernel, a known byte count for calibrating FETCH_SIZE /
    WRITE_SIZE) followed by `nbfs` BFS runs on RMAT-22.  With `levels_out`
    the level records of those runs (direction, mode, launch index,
    counters) go to that JSON file, in launch order, for
    --direction-summary."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    lib = _lib.load()
    g = lgmi.generators.rmat(22, 16, seed=1)
    deg = np.diff(g.csr()[0])
    h = g.handle()
    rng = np.random.default_rng(20260101)
    roots = rng.choice(np.nonzero(deg > 0)[0], size=64, replace=False).astype(np.int32)
    st = _lib.BfsStats()
    buf = (_lib.BfsStep * 4096)()
    runs = []
    for i in range(nbfs):
        _lib.check(lib.lgmi_bfs_run(h, int(roots[i % 64]), 0, 1, ctypes.byref(st)))
        if levels_out:
            n = lib.lgmi_bfs_steps(h, 4096, buf)
            if n < 0:
                _lib.check(n)
            runs.append({"launches": int(st.launches), "levels": int(st.levels),
                         "recs": [{"dir": buf[k].dir, "variant": buf[k].variant, "launch": buf[k].launch,
                                   "n_front": buf[k].n_front, "m_front": buf[k].m_front,
                                   "n_next": buf[k].n_next, "scanned": buf[k].scanned, "cand": buf[k].cand}
                                  for k in range(n)]})
    _lib.check(lib.lgmi_synchronize())
    if levels_out:
        with open(levels_out, "w") as f:
            json.dump({"nv": int(g.nv()), "runs": runs}, f)
    log(f"profile pass done: copy {nbytes} B x3, {nbfs} BFS")


def direction_summary(trace_csv, levels_json):
    """Per-direction table of the level-kernel dispatches (bfs_first for
    level 0 of a single source, then bfs_level) of a profiled
    --profile-pass run: each dispatch of rocprofv3's kernel trace (in
    start order) is labelled from the level records of the same run (the
    launches of one search are consecutive; a launch that ran no level is
    one issued after the search ended; a solo launch runs several), and its
    duration and the algorithmic bytes of its levels are summed per label."""
    import csv
    from types import SimpleNamespace
    rows = [r for r in csv.DictReader(open(trace_csv)) if "bfs_level" in r["Kernel_Name"] or
            "bfs_first" in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3 for r in rows]   # us
    lv = json.load(open(levels_json))
    nv = lv["nv"]
    labels, nbytes = [], []
    for run in lv["runs"]:
        by_launch = {}
        for rec in run["recs"]:
            by_launch.setdefault(rec["launch"], []).append(rec)
        for li in range(run["launches"]):
            recs = by_launch.get(li, [])
            if li == 0 and recs:      # single-source searches: launch 0 is bfs_first
                lab = "level 0 (bfs_first)"
            elif not recs:
                lab = "after the last level"
            elif len(recs) > 1 or recs[0]["variant"] == 3:
                lab = "top-down, solo"
            elif recs[0]["dir"] == 1:
                lab = "bottom-up"
            else:
                lab = "top-down, " + {0: "sweep", 1: "range", 2: "queue"}[recs[0]["variant"]]
            labels.append(lab)
            nbytes.append(sum(step_bytes(SimpleNamespace(**r), nv) for r in recs))
    if len(labels) != len(dur):
        raise SystemExit(f"{len(dur)} bfs_level dispatches in the trace, {len(labels)} launches in the level log")
    agg = {}
    for lab, t, b in zip(labels, dur, nbytes):
        a = agg.setdefault(lab, [0, 0.0, 0.0])
        a[0] += 1
        a[1] += t
        a[2] += b
    tot = sum(a[1] for a in agg.values())
    lines = ["| dispatches | count | total ms | avg us | share | algorithmic GB/s | fraction of 8 TB/s |",
             "|---|---|---|---|---|---|---|"]
    for lab, (n, t, b) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        gbs = b / (t * 1e-6) / 1e9 if t > 0 else 0.0
        lines.append(f"| {lab} | {n} | {t / 1e3:.2f} | {t / n:.2f} | {100 * t / tot:.1f} % | {gbs:.0f} | "
                     f"{gbs / HBM_PEAK_GBS:.4f} |")
    print("\n".join(lines))


def lg_load_bench(lgmi, g):
    """loadgraph of an LG text file holding symmetrised RMAT-20 (15.7M edge
    lines): header and edge-line parsing on the host (lgmi_lg_parse_edges),
    then the device ingest, beside the same parse followed by the host CSR
    builder."""
    import tempfile

    from lgmi import persistence as P
    from lgmi.graphs import csr_from_edges
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "rmat20.lg")
        with open(path, "w") as f:
            P.savegraph(f, g, "rmat20")
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        g2 = P.loadgraph(path, "rmat20")
        t_load = time.perf_counter() - t0
        # the same file: parse only, then the host builder
        t0 = time.perf_counter()
        r = P._Reader(P._auto_decompress(P._read_all(path)))
        h = P._parse_header(r.readline().strip())
        s, dd = P._parse_edges(r, h)
        t_parse = time.perf_counter() - t0
        t0 = time.perf_counter()
        rp_h, col_h = csr_from_edges(h.nv, s - 1, dd - 1, symmetrize=True)
        t_host = time.perf_counter() - t0
    rp, col = g.csr()
    same = bool(np.array_equal(g2.csr()[0], rp) and np.array_equal(g2.csr()[1], col) and
                np.array_equal(rp_h, rp) and np.array_equal(col_h, col))
    res = {"lines": int(h.ne), "mb": round(size / 1e6, 1), "load_ms": round(t_load * 1e3, 1),
           "parse_ms": round(t_parse * 1e3, 1), "device_ingest_ms": round((t_load - t_parse) * 1e3, 1),
           "host_builder_ms": round(t_host * 1e3, 1), "same_csr": same,
           "cpu": "the same parse, then lgmi_csr_build (host C++, 1 thread)"}
    log(f"extras: loadgraph RMAT-20 LG ({size / 1e6:.0f} MB, {h.ne} lines) {t_load * 1e3:.0f} ms "
        f"(parse {t_parse * 1e3:.0f} ms); parse + host builder {(t_parse + t_host) * 1e3:.0f} ms; same={same}")
    return res


def extras_bench(args, lgmi, _lib, orc):
    """The widened rows at scale, each beside a CPU run of the same work:
    stress centrality and the distance statistics (closeness / radiality /
    eccentricity) over the 1024 RMAT-20 sources of the betweenness workload,
    connected components and device ingest of RMAT-22, and a depth-2
    neighbourhood of an RMAT-22 hub."""
    import ctypes
    from lgmi import distances as D
    from lgmi.graphs import csr_from_edges
    out = {}
    nthreads = min(16, os.cpu_count() or 1)
    g = lgmi.generators.rmat(20, 16, seed=1)
    rp, col = g.csr()
    nv, nnz = g.nv(), len(col)
    rng = np.random.default_rng(4242)
    srcs = rng.choice(np.nonzero(np.diff(rp) > 0)[0], size=1024, replace=False).astype(np.int64)
    g.handle()
    from lgmi.centrality import last_stats
    # stress: 1024 sources, one persistent launch
    lgmi.stress_centrality(g, srcs + 1)
    t1 = time.perf_counter()
    st = lgmi.stress_centrality(g, srcs + 1)
    wall = time.perf_counter() - t1
    st_stats = last_stats(g)
    ns = 16
    c0 = time.perf_counter()
    want = orc.stress(rp, col, rp, False, srcs[:ns], nthreads=nthreads)
    t_cpu = time.perf_counter() - c0
    got = lgmi.stress_centrality(g, srcs[:ns] + 1)
    out["stress_rmat20_1024"] = {"ms": round(wall * 1e3, 2), "device_ms": round(st_stats["device_ms"], 2),
                                 "gteps": round(1024 * nnz / 2 / wall / 1e9, 3),
                                 "cpu_ms_per_source": round(t_cpu / ns * 1e3, 2), "cpu_threads": nthreads,
                                 "speedup_vs_cpu": round(t_cpu / ns * 1024 / wall, 1),
                                 "sample_exact": bool((got == want).all())}
    log(f"extras: stress 1024 sources {wall * 1e3:.1f} ms, cpu {t_cpu / ns * 1e3:.1f} ms/source "
        f"({nthreads} threads), exact={out['stress_rmat20_1024']['sample_exact']}")
    # distance statistics: 1024 sources
    D.distance_stats(g, srcs + 1)
    t1 = time.perf_counter()
    ds = D.distance_stats(g, srcs + 1)
    wall = time.perf_counter() - t1
    c0 = time.perf_counter()
    ok = True
    for i in range(8):
        d = orc.gdistances(rp, col, [int(srcs[i])]).astype(np.int64)
        r = d[d >= 0]
        ok = ok and [int(r.sum()), len(r), int(r.max())] == ds[i].tolist()
    t_cpu = (time.perf_counter() - c0) / 8
    out["distance_stats_rmat20_1024"] = {"ms": round(wall * 1e3, 2), "gteps": round(1024 * nnz / 2 / wall / 1e9, 3),
                                         "cpu_ms_per_source": round(t_cpu * 1e3, 2), "cpu_threads": 1,
                                         "speedup_vs_cpu": round(t_cpu * 1024 / wall, 1), "sample_exact": ok}
    log(f"extras: distance stats 1024 sources {wall * 1e3:.1f} ms, cpu BFS {t_cpu * 1e3:.1f} ms/source, exact={ok}")
    # RMAT-22: ingest, components, neighbourhood
    g22 = lgmi.generators.rmat(22, 16, seed=1)
    rp22, col22 = g22.csr()
    src22 = np.repeat(np.arange(g22.nv()), np.diff(rp22))
    keep = src22 <= col22
    es, ed = src22[keep], col22[keep]
    perm = np.random.default_rng(5).permutation(len(es))
    es, ed = np.ascontiguousarray(es[perm], np.int32), np.ascontiguousarray(ed[perm], np.int32)
    lib = _lib.load()
    for rep in range(2):
        h = ctypes.c_void_p()
        ne = np.zeros(1, dtype=np.int64)
        t1 = time.perf_counter()
        _lib.check(lib.lgmi_graph_from_edges(_lib.i32p(es), _lib.i32p(ed), len(es), 0, g22.nv(), ctypes.byref(h),
                                             _lib.i64p(ne)))
        wall = time.perf_counter() - t1
        lib.lgmi_graph_destroy(h)
    c0 = time.perf_counter()
    rp_h, col_h = csr_from_edges(g22.nv(), es, ed, symmetrize=True)
    t_cpu = time.perf_counter() - c0
    out["ingest_rmat22"] = {"edges": int(len(es)), "ms": round(wall * 1e3, 1), "ne": int(ne[0]),
                            "cpu_ms": round(t_cpu * 1e3, 1), "cpu": "lgmi_csr_build (host C++, 1 thread)",
                            "same_csr": bool((rp_h == rp22).all() and (col_h == col22).all())}
    log(f"extras: ingest RMAT-22 ({len(es)} edges) {wall * 1e3:.1f} ms incl. host copies; host builder "
        f"{t_cpu * 1e3:.1f} ms")
    out["loadgraph_rmat20"] = lg_load_bench(lgmi, g)
    lab = np.zeros(g22.nv(), dtype=np.int64)
    g22.handle()
    lgmi.connected_components_(lab, g22)
    t1 = time.perf_counter()
    lgmi.connected_components_(lab, g22)
    wall = time.perf_counter() - t1
    try:
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import connected_components as sp_cc
        m = csr_matrix((np.ones(len(col22), np.int8), col22, rp22), shape=(g22.nv(), g22.nv()))
        c0 = time.perf_counter()
        ncomp, _ = sp_cc(m, directed=False)
        t_cpu = time.perf_counter() - c0
    except ImportError:
        ncomp, t_cpu = -1, float("nan")
    out["components_rmat22"] = {"ms": round(wall * 1e3, 2), "components": int(len(np.unique(lab))),
                                "cpu_ms": round(t_cpu * 1e3, 1), "cpu": "scipy.sparse.csgraph (1 thread)",
                                "same_count": int(len(np.unique(lab))) == int(ncomp)}
    log(f"extras: components RMAT-22 {wall * 1e3:.1f} ms ({len(np.unique(lab))} components), scipy "
        f"{t_cpu * 1e3:.1f} ms")
    hub = int(np.argmax(np.diff(rp22)))
    verts = np.empty(g22.nv(), dtype=np.int32)
    dists = np.empty(g22.nv(), dtype=np.int32)
    cnt = np.zeros(1, dtype=np.int32)
    for rep in range(2):
        t1 = time.perf_counter()
        _lib.check(lib.lgmi_neighborhood(g22.handle(), hub, 2, 0, _lib.i32p(verts), _lib.i32p(dists), _lib.i32p(cnt)))
        wall = time.perf_counter() - t1
    out["neighborhood_rmat22_hub_d2"] = {"ms": round(wall * 1e3, 2), "vertices": int(cnt[0]),
                                         "note": "C-ABI call incl. the host copies of the queue and distances"}
    log(f"extras: neighborhood(hub, 2) on RMAT-22: {int(cnt[0])} vertices in {wall * 1e3:.1f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile-pass", type=int, default=0,
                    help="run only the counter-profiling workload with this many BFS runs")
    ap.add_argument("--levels-out", default="", help="with --profile-pass: write the runs' level records here")
    ap.add_argument("--direction-summary", nargs=2, metavar=("TRACE_CSV", "LEVELS_JSON"),
                    help="print the per-direction table of a profiled --profile-pass run and exit")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-roots", type=int, default=4)
    ap.add_argument("--cpu-bc-sources", type=int, default=32)
    ap.add_argument("--no-bc", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="only the RMAT-22 BFS and RMAT-20 betweenness legs (no batch, learned-switch or drop-in legs)")
    ap.add_argument("--bc-only", action="store_true", help="run only the RMAT-20 betweenness workload")
    ap.add_argument("--extras", action="store_true",
                    help="run only the widened rows (stress, distance measures, ingest, components, neighborhood)")
    ap.add_argument("--config4", action="store_true",
                    help="run only the uneven small-graph batch (betweenness) workload, autotuning first")
    ap.add_argument("--config4-graphs", type=int, default=4096)
    ap.add_argument("--cpu-batch-graphs", type=int, default=8)
    ap.add_argument("--config5-roots", type=int, default=20)
    ap.add_argument("--save-tuning", default="", help="write the batch autotuner's choice to this JSON file")
    args = ap.parse_args()
    if args.direction_summary:
        direction_summary(*args.direction_summary)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import lgmi
    from lgmi import _lib
    from oracle import oracle as orc
    if _lib.load().lgmi_device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.profile_pass:
        profile_pass(lgmi, _lib, args.profile_pass, args.levels_out)
        return
    if args.config4:
        print(json.dumps(config4_bench(args, lgmi, orc, tune=True)), flush=True)
        return
    if args.extras:
        print(json.dumps(extras_bench(args, lgmi, _lib, orc)), flush=True)
        return
    if args.bc_only:
        print(json.dumps(bc_bench(args, lgmi, _lib, orc)), flush=True)
        return
    result = bfs_bench(args, lgmi, _lib, orc)
    if not args.no_bc:
        result["bc"] = bc_bench(args, lgmi, _lib, orc)
    if not args.quick:
        result["batch"] = config4_bench(args, lgmi, orc)
        result["learned_switch"] = config5_bench(args, lgmi, _lib)
        result["dropin"] = dropin_bench(lgmi, _lib)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
