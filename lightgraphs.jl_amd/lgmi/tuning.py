"""Launch thresholds of the BFS steps and of the betweenness heavy-row
segments: the autotuner, and the JSON file it writes (``tuning.json`` next
to this module), which the package applies before it creates its first
graph handle.

The reference's threaded BFS has one such knob, ``queue_segment_size``
(src/Parallel/traversals/gdistances.jl:135-139), the chunk of the frontier
queue a thread claims at a time.  The GPU path has several, none of which
changes a result (the tests run every step variant against the oracle):

* ``queue_min_edges`` / ``queue_max_edges`` / ``queue_front`` -- a top-down
  step runs in queue mode (every work-group reads the frontier list and
  expands an equal share of its rows' entries) when the frontier has at most
  ``queue_front`` vertices and between the two edge bounds;
* ``range_front`` / ``range_light`` -- otherwise range mode (each
  work-group owns a vertex range and takes the hub rows' entries in it)
  when the frontier holds rows longer than 256 entries, has at most
  ``range_front`` vertices and at most ``range_light`` edges in shorter rows;
* ``wave_row`` -- in the bitmap sweep, rows of at least this many entries
  are expanded by a whole wavefront, shorter ones packed into the lanes;
* ``lane_rounds`` -- a bottom-up candidate scans this many rounds of four
  entries in its own lane before a whole wavefront takes its row;
* ``solo_front`` / ``solo_edges`` -- a level whose frontier has at most
  this many vertices and edges runs in solo mode: one work-group expands it
  and goes on to the next level inside the same launch while the frontier
  stays that small (0 turns solo mode off);
* ``bc_seg_len`` -- betweenness cuts rows longer than this into segments of
  this many entries (fixed for a graph at its first betweenness call).

``autotune()`` runs a coordinate search over candidate values on the BFS
workload of the throughput benchmark (symmetrised RMAT-22, its 64 seeded
roots), a high-diameter grid (2048 x 2048) and a power-law graph
(barabasi_albert(2M, 8)), each scored by its summed device time relative to
the defaults, and the betweenness segment length on the betweenness
workload of the benchmark (RMAT-20, 1024 seeded sources).  ``python -m lgmi.tuning`` runs it and writes the file.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _lib

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

FIELDS = ["wave_row", "lane_rounds", "range_front", "queue_front", "range_light", "queue_min_edges",
          "queue_max_edges", "bc_seg_len", "solo_front", "solo_edges"]

# Candidate values of the coordinate search, in the order they are searched.
CANDIDATES = {
    "solo_edges": [0, 2048, 4096, 16384, 32768],
    "solo_front": [512, 1024, 2048],
    "queue_min_edges": [0, 256, 1024, 4096, 16384, 65536, 1 << 40],
    "queue_max_edges": [16384, 65536, 262144, 1048576],
    "queue_front": [256, 1024, 2048, 4096],
    "range_front": [0, 256, 1024, 2048],
    "range_light": [4096, 32768, 262144],
    "wave_row": [8, 16, 32, 64, 128, 257],
    "lane_rounds": [1, 2, 4, 8, 16],
}
BC_CANDIDATES = [256, 512, 1024, 2048, 4096]

_applied = False


def get() -> dict:
    t = _lib.Tuning()
    _lib.check(_lib.load().lgmi_get_tuning(_lib.ctypes.byref(t)))
    return {f: int(getattr(t, f)) for f in FIELDS}


def set_values(values: dict | None) -> None:
    """Set the thresholds (missing keys keep their current value; None
    restores the built-in defaults)."""
    lib = _lib.load()
    if values is None:
        _lib.check(lib.lgmi_set_tuning(None))
        return
    cur = get()
    cur.update({k: int(v) for k, v in values.items() if k in FIELDS})
    t = _lib.Tuning(**cur)
    _lib.check(lib.lgmi_set_tuning(_lib.ctypes.byref(t)))


def load(path: str = TUNING_PATH) -> bool:
    """Apply the thresholds stored in `path` (False when it does not exist)."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        set_values(json.load(f)["best"])
    return True


def apply_once() -> None:
    """Apply tuning.json once per process (the first graph handle calls it)."""
    global _applied
    if not _applied:
        _applied = True
        load()


def _bfs_ms(g, roots, reps: int = 3) -> float:
    """Summed device time (ms) of full BFS runs (distances + parents) from
    `roots`, each root the median of `reps` runs."""
    lib = _lib.load()
    h = g.handle()
    st = _lib.BfsStats()
    tot = 0.0
    for r in roots:
        t = []
        for _ in range(reps):
            _lib.check(lib.lgmi_bfs_run(h, int(r), 0, 1, _lib.ctypes.byref(st)))
            t.append(st.device_ms)
        tot += float(np.median(t))
    return tot


def _workloads(log):
    from . import generators as gen
    out = []
    t0 = time.time()
    g = gen.rmat(22, 16, seed=1)
    deg = np.diff(g.csr()[0])
    rng = np.random.default_rng(20260101)           # the throughput benchmark's roots
    out.append(("rmat22", g, rng.choice(np.nonzero(deg > 0)[0], size=64, replace=False)))
    g = gen.grid([2048, 2048])
    out.append(("grid2048", g, np.array([0, 2048 * 700 + 1300])))
    g = gen.barabasi_albert(2_000_000, 8, seed=5)
    rng = np.random.default_rng(11)
    out.append(("ba2M", g, rng.choice(g.nv(), size=16, replace=False)))
    log(f"tuning workloads built ({time.time() - t0:.1f}s)")
    return out


def autotune(log=print) -> dict:
    """Coordinate search of the BFS thresholds, then the betweenness segment
    length; applies and returns the choice with every candidate's score."""
    set_values(None)
    work = _workloads(log)
    for _, g, roots in work:
        _bfs_ms(g, roots[:2], reps=1)                # allocations, first launches
    base = {name: _bfs_ms(g, roots) for name, g, roots in work}
    log(f"defaults: " + ", ".join(f"{k} {v:.2f} ms" for k, v in base.items()))

    def score(values):
        set_values(values)
        per = {name: _bfs_ms(g, roots) for name, g, roots in work}
        return sum(per[n] / base[n] for n in per) / len(per), per

    best = get()
    best_score, best_per = score(best)
    history = []
    for field, cands in CANDIDATES.items():
        for v in cands:
            if v == best[field]:
                continue
            trial = dict(best, **{field: v})
            sc, per = score(trial)
            history.append({"field": field, "value": v, "score": round(sc, 4),
                            "ms": {k: round(x, 3) for k, x in per.items()}})
            log(f"{field}={v}: score {sc:.4f} ({', '.join(f'{k} {x:.2f}' for k, x in per.items())})")
            if sc < best_score * 0.995:
                best, best_score, best_per = trial, sc, per
        log(f"{field} -> {best[field]} (score {best_score:.4f})")
    set_values(best)
    bc_best, bc_hist = tune_bc_segments(best["bc_seg_len"], log)
    best["bc_seg_len"] = bc_best
    set_values(best)
    for name, g, roots in work:
        g._release()
    return {"best": best, "score": round(best_score, 4), "defaults_ms": {k: round(v, 3) for k, v in base.items()},
            "best_ms": {k: round(v, 3) for k, v in best_per.items()}, "bfs_candidates": history,
            "bc_candidates": bc_hist,
            "workloads": {"rmat22": "symmetrised RMAT-22, 64 seeded roots", "grid2048": "2048 x 2048 grid, 2 roots",
                          "ba2M": "barabasi_albert(2_000_000, 8), 16 roots",
                          "bc": "RMAT-20 betweenness, 1024 seeded sources"}}


def tune_bc_segments(current: int, log=print):
    """The betweenness segment length: each candidate on a fresh graph
    handle (so that it builds its own segments) of the betweenness workload
    (RMAT-20, 1024 seeded sources), three timed runs after a warm-up, in two
    passes over the candidates in opposite orders (a single pass in one
    order let the GPU's warming over the run decide between close
    candidates); a candidate's time is the median of its six runs.  Returns
    the choice (a candidate must beat the best so far by 1 %) and the
    candidates' times."""
    from . import generators as gen
    from .centrality import betweenness_centrality, last_stats
    from .graphs import SimpleGraph
    g20 = gen.rmat(20, 16, seed=1)
    rng = np.random.default_rng(4242)
    srcs = rng.choice(np.nonzero(np.diff(g20.csr()[0]) > 0)[0], size=1024, replace=False) + 1
    runs = {L: [] for L in BC_CANDIDATES}
    for order in (list(BC_CANDIDATES), list(reversed(BC_CANDIDATES))):
        for L in order:
            set_values({"bc_seg_len": L})
            g = SimpleGraph.from_csr(*g20.csr())
            betweenness_centrality(g, srcs)
            for _ in range(3):
                betweenness_centrality(g, srcs)
                runs[L].append(last_stats(g)["device_ms"])
            g._release()
    bc_hist = []
    bc_best, bc_best_ms = current, None
    for L in BC_CANDIDATES:
        m = float(np.median(runs[L]))
        bc_hist.append({"bc_seg_len": L, "device_ms": round(m, 3)})
        log(f"bc_seg_len={L}: {m:.2f} ms")
        if bc_best_ms is None or m < bc_best_ms * 0.99:
            bc_best, bc_best_ms = L, m
    set_values({"bc_seg_len": current})
    return bc_best, bc_hist


def save(result: dict, path: str = TUNING_PATH) -> None:
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


if __name__ == "__main__":
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="autotune the BFS and betweenness launch thresholds")
    ap.add_argument("--out", default=TUNING_PATH)
    ap.add_argument("--bc-only", action="store_true",
                    help="time the betweenness segment lengths with the other thresholds from tuning.json; "
                         "print the candidates, write nothing")
    a = ap.parse_args()
    say = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.bc_only:
        load()
        choice, hist = tune_bc_segments(get()["bc_seg_len"], say)
        print(json.dumps({"bc_seg_len": choice, "bc_candidates": hist}))
    else:
        res = autotune(log=say)
        save(res, a.out)
        print(json.dumps({"best": res["best"], "score": res["score"]}))
