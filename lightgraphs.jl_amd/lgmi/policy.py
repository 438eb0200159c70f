"""Learned BFS direction policy: label collection, GPU training, evaluation.

No counterpart in the reference (its BFS is top-down only,
src/traversals/bfs.jl:108-122).  The device BFS chooses top-down or
bottom-up per level; this module replaces the fixed Beamer alpha=15 /
beta=18 rule with a small FP32 MLP (16 features -> 64 ReLU -> 2 logits)
trained on the GPU by liblgmi's own kernels:

1. collect(): for sampled levels of BFS runs from seeded roots, run the
   level both ways on the identical state (lgmi_bfs_probe) and record the
   two device times beside the 16 features the BFS computes on-device.
2. train(): label = the faster direction, weight = |t_td - t_bu|
   (cost-sensitive: a level where the choice does not matter carries no
   weight); full-batch Adam on the GPU (lgmi_mlp_train).
3. evaluate(): held-out roots under the heuristic and the learned policy,
   device time per BFS, outputs compared bit for bit.

The policy never changes an output: distances and parents are the same in
either direction (the smallest-id parent rule), which evaluate() checks.
"""
from __future__ import annotations

import json
import os
import time
import zlib

import numpy as np

from . import _lib

FEATURES = [
    "log2(1+|F|)", "log2(1+m_F)", "log2(1+m_U)", "log2(1+|U|)", "log2(1+level)", "|F|/nv",
    "log2 m_F - log2 m_U", "log2 |F| - log2 |F_prev|", "previous direction",
    "hist deg 0", "hist deg 1-3", "hist deg 4-15", "hist deg 16-63", "hist deg 64-255", "hist deg 256-1023",
    "hist deg >=1024",
]
NW = _lib.LGMI_POLICY_NWEIGHTS
NTRAIN = NW - 32
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "policy_weights.npz")


def init_params(x: np.ndarray, seed: int = 0) -> np.ndarray:
    """He-initialised weights; mu/sd standardise the training features."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NW, dtype=np.float32)
    p[:64 * 16] = rng.normal(0, np.sqrt(2.0 / 16), 64 * 16)
    p[64 * 16 + 64:64 * 16 + 64 + 128] = rng.normal(0, np.sqrt(1.0 / 64), 128)
    mu = x.mean(axis=0)
    sd = x.std(axis=0)
    sd = np.where(sd > 1e-3, sd, 1.0)
    p[NTRAIN:NTRAIN + 16] = mu
    p[NTRAIN + 16:] = sd
    return p


def probe(g, root: int, levels) -> np.ndarray:
    """Rows (feat[16], td_us, bu_us, level, n_front) for the listed levels."""
    lib = _lib.load()
    lv = np.ascontiguousarray(sorted(set(int(x) for x in levels)), dtype=np.int32)
    buf = (_lib.ProbeRec * max(1, len(lv)))()
    n = lib.lgmi_bfs_probe(g.handle(), int(root), 0, _lib.i32p(lv if len(lv) else np.zeros(1, np.int32)),
                           len(lv), buf)
    if n < 0:
        _lib.check(n)
    out = np.zeros((n, 20), dtype=np.float64)
    for i in range(n):
        r = buf[i]
        out[i, :16] = list(r.feat)
        out[i, 16] = r.td_us
        out[i, 17] = r.bu_us
        out[i, 18] = r.level
        out[i, 19] = r.n_front
    return out


def bfs_levels(g, root: int) -> int:
    lib = _lib.load()
    st = _lib.BfsStats()
    _lib.check(lib.lgmi_bfs_run(g.handle(), int(root), 0, 1, _lib.ctypes.byref(st)))
    return int(st.levels)


def collect(g, roots, max_probe_levels: int = 64, seed: int = 0, log=None) -> np.ndarray:
    """Probe records for every root.  Graphs with at most max_probe_levels
    levels are probed at every level; deeper ones at a seeded sample."""
    rng = np.random.default_rng(seed)
    rows = []
    for i, r in enumerate(roots):
        depth = bfs_levels(g, r)
        if depth <= max_probe_levels:
            levels = np.arange(depth)
        else:
            levels = np.sort(rng.choice(depth, size=max_probe_levels, replace=False))
        p = probe(g, r, levels)
        rows.append(np.concatenate([p, np.full((len(p), 1), i, dtype=np.float64)], axis=1))
        if log and (i + 1) % max(1, len(roots) // 4) == 0:
            log(f"collect: {i + 1}/{len(roots)} roots")
    return np.concatenate(rows) if rows else np.zeros((0, 21))


MAX_STEPS = 16384


def _forced_run(lib, h, root, forced, buf):
    """One BFS with every step forced to `forced`, HIP events around every
    step launch; returns the step records (features, per-launch ms)."""
    _lib.check(lib.lgmi_bfs_force_direction(h, forced))
    _lib.check(lib.lgmi_bfs_run(h, int(root), 0, 1, None))
    n = lib.lgmi_bfs_steps(h, MAX_STEPS, buf)
    if n < 0:
        _lib.check(n)
    feat = np.array([list(buf[i].feat) for i in range(n)], dtype=np.float64).reshape(n, 16)
    ms = np.array([buf[i].event_ms for i in range(n)], dtype=np.float64)
    nf = np.array([buf[i].n_front for i in range(n)], dtype=np.float64)
    return feat, ms, nf


def collect_pipelined(g, roots, log=None) -> np.ndarray:
    """Labels from whole BFS runs: each root is searched once with every
    step forced top-down and once forced bottom-up, inside the normal
    back-to-back launch stream, with HIP events around every step.  The
    frontier at level L is the same in both runs, so every level yields a
    (t_td, t_bu) pair on identical state.  The previous-direction feature is
    set to the faster direction of the previous level (what a good policy
    would have chosen).  Rows as collect()."""
    lib = _lib.load()
    h = g.handle()
    buf = (_lib.BfsStep * MAX_STEPS)()
    _lib.check(lib.lgmi_bfs_time_steps(h, 1))
    rows = []
    try:
        for i, r in enumerate(roots):
            f_td, t_td, nf = _forced_run(lib, h, r, 0, buf)
            _, t_bu, _ = _forced_run(lib, h, r, 1, buf)
            n = min(len(t_td), len(t_bu))
            rec = np.zeros((n, 21))
            rec[:, :16] = f_td[:n]
            rec[:, 16] = t_td[:n] * 1e3
            rec[:, 17] = t_bu[:n] * 1e3
            best = (rec[:, 17] < rec[:, 16]).astype(np.float64)
            rec[1:, 8] = best[:-1]
            rec[0, 8] = 0.0
            rec[:, 18] = np.arange(n)
            rec[:, 19] = nf[:n]
            rec[:, 20] = i
            rows.append(rec)
            if log and (i + 1) % max(1, len(roots) // 4) == 0:
                log(f"collect: {i + 1}/{len(roots)} roots")
    finally:
        lib.lgmi_bfs_force_direction(h, -1)
        lib.lgmi_bfs_time_steps(h, 0)
    return np.concatenate(rows) if rows else np.zeros((0, 21))


def labels(recs):
    """Features, labels (1 = bottom-up faster) and weights for one record
    array or a list of them (one per graph).  A level's weight is the time a
    wrong choice would cost, |t_td - t_bu|, relative to the best total of
    its own BFS run (column 20 holds the root index); every graph then gets
    the same total weight, so a high-diameter graph's many small decisions
    count as much as a few large ones elsewhere."""
    if isinstance(recs, np.ndarray):
        recs = [recs]
    xs, ys, ws = [], [], []
    for rec in recs:
        td, bu = rec[:, 16], rec[:, 17]
        cost = np.abs(td - bu)
        root = rec[:, 20].astype(np.int64)
        tot = np.bincount(root, weights=np.minimum(td, bu))
        wrel = cost / np.maximum(tot[root], 1e-9)
        wrel = wrel / max(wrel.sum(), 1e-30)
        xs.append(rec[:, :16])
        ys.append((bu < td).astype(np.int32))
        ws.append(wrel)
    x = np.ascontiguousarray(np.concatenate(xs), dtype=np.float32)
    y = np.ascontiguousarray(np.concatenate(ys), dtype=np.int32)
    w = np.concatenate(ws)
    w = np.ascontiguousarray(w / w.mean(), dtype=np.float32)
    return x, y, w


def loss_grad(x, y, w, params):
    lib = _lib.load()
    loss = np.zeros(1, np.float32)
    grad = np.zeros(NW, np.float32)
    _lib.check(lib.lgmi_mlp_loss_grad(_lib.f32p(x), _lib.i32p(y), _lib.f32p(w) if w is not None else None,
                                      len(y), _lib.f32p(np.ascontiguousarray(params, np.float32)),
                                      _lib.f32p(loss), _lib.f32p(grad)))
    return float(loss[0]), grad


def train(x, y, w, steps: int = 3000, lr: float = 3e-3, seed: int = 0, params=None):
    lib = _lib.load()
    p = init_params(x, seed) if params is None else np.array(params, dtype=np.float32)
    loss = np.zeros(1, np.float32)
    _lib.check(lib.lgmi_mlp_train(_lib.f32p(x), _lib.i32p(y), _lib.f32p(w) if w is not None else None, len(y),
                                  _lib.f32p(p), int(steps), float(lr), _lib.f32p(loss)))
    return p, float(loss[0])


def logits(x, params):
    lib = _lib.load()
    out = np.zeros((len(x), 2), np.float32)
    if len(x):
        _lib.check(lib.lgmi_mlp_forward(_lib.f32p(np.ascontiguousarray(x, np.float32)), len(x),
                                        _lib.f32p(np.ascontiguousarray(params, np.float32)), _lib.f32p(out)))
    return out


def heuristic_choice(rec: np.ndarray, alpha=15.0, beta=18.0) -> np.ndarray:
    """The alpha/beta rule the device applies, restated on the features:
    after a top-down step go bottom-up when m_F > m_U / alpha; after a
    bottom-up step go top-down when |F| shrank and |F| <= nv / beta.
    (Level 0 is always top-down.)"""
    mf = np.exp2(rec[:, 1]) - 1
    mu = np.exp2(rec[:, 2]) - 1
    prev_bu = rec[:, 8] > 0.5
    frac = rec[:, 5]
    shrink = rec[:, 7] < 0
    td_rule = np.where(mf > mu / alpha, 1, 0)
    bu_rule = np.where(shrink & (frac <= 1.0 / beta), 0, 1)
    out = np.where(prev_bu, bu_rule, td_rule)
    out[rec[:, 18] == 0] = 0
    return out


def decision_cost(rec: np.ndarray, choice: np.ndarray) -> float:
    """Sum over probed levels of the time of the chosen direction (us)."""
    return float(np.where(choice == 1, rec[:, 17], rec[:, 16]).sum())


def set_learned(params) -> None:
    lib = _lib.load()
    p = np.ascontiguousarray(params, np.float32)
    _lib.check(lib.lgmi_bfs_set_policy(_lib.LGMI_POLICY_LEARNED, 15.0, 18.0, _lib.f32p(p), NW))


def set_heuristic(alpha=15.0, beta=18.0) -> None:
    _lib.check(_lib.load().lgmi_bfs_set_policy(_lib.LGMI_POLICY_HEURISTIC, alpha, beta, None, 0))


def evaluate(g, roots, params, reps: int = 1, check_outputs: bool = True):
    """Per-root device ms under the heuristic and the learned policy, and
    whether the outputs were identical."""
    lib = _lib.load()
    h = g.handle()
    st = _lib.BfsStats()
    nv = g.nv()
    t_h, t_l = [], []
    same = True
    d0 = np.empty(nv, np.int32)
    p0 = np.empty(nv, np.int32)
    d1 = np.empty(nv, np.int32)
    p1 = np.empty(nv, np.int32)
    for r in roots:
        th, tl = [], []
        for k in range(reps):
            set_heuristic()
            _lib.check(lib.lgmi_bfs_run(h, int(r), 0, 1, _lib.ctypes.byref(st)))
            th.append(st.device_ms)
            if check_outputs and k == 0:
                _lib.check(lib.lgmi_bfs_fetch(h, _lib.i32p(d0), _lib.i32p(p0)))
            set_learned(params)
            _lib.check(lib.lgmi_bfs_run(h, int(r), 0, 1, _lib.ctypes.byref(st)))
            tl.append(st.device_ms)
            if check_outputs and k == 0:
                _lib.check(lib.lgmi_bfs_fetch(h, _lib.i32p(d1), _lib.i32p(p1)))
                same = same and bool((d0 == d1).all() and (p0 == p1).all())
        t_h.append(float(np.median(th)))
        t_l.append(float(np.median(tl)))
    set_heuristic()
    return np.array(t_h), np.array(t_l), same


def save(params, meta: dict, path: str = WEIGHTS_PATH) -> None:
    np.savez(path, params=np.asarray(params, np.float32), meta=np.array(json.dumps(meta)))


def load_weights(path: str = WEIGHTS_PATH):
    if not os.path.exists(path):
        return None, None
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), json.loads(str(z["meta"]))


def bench_roots(g, n: int = 64, seed: int = 20260101) -> np.ndarray:
    """The throughput benchmark's seeded roots (degree >= 1)."""
    deg = np.diff(g.csr()[0])
    return np.random.default_rng(seed).choice(np.nonzero(deg > 0)[0], size=n, replace=False)


def config5(out_dir: str, n_train: int = 2000, n_eval: int = 200, grid_train: int = 40, rmat22_train: int = 500,
            log=print) -> dict:
    """The learned-switch workload: directed RMAT-23 (edge factor 16),
    barabasi_albert(8_000_000, 8) and a 4096 x 4096 grid, plus the
    undirected RMAT-22 of the throughput benchmark; seeded training and
    held-out roots per graph (on RMAT-22 the benchmark's 64 roots are the
    held-out set and never trained on); training on the GPU; evaluation
    against the alpha/beta heuristic with bit-identical outputs."""
    from . import generators as gen
    t0 = time.time()
    graphs = [
        ("rmat23_directed", lambda: gen.rmat(23, 16, seed=23, directed=True), n_train),
        ("ba_8M_8", lambda: gen.barabasi_albert(8_000_000, 8, seed=8), n_train),
        ("grid_4096x4096", lambda: gen.grid([4096, 4096]), grid_train),
        ("rmat22_undirected", lambda: gen.rmat(22, 16, seed=1), rmat22_train),
    ]
    recs, built = {}, {}
    for name, make, ntr in graphs:
        g = make()
        rp = g.csr()[0]
        deg = np.diff(rp)
        if g.is_directed():
            deg = deg + np.diff(g.in_csr()[0])
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        cand = np.nonzero(np.diff(rp) > 0)[0]
        pick = rng.choice(cand, size=min(len(cand), ntr + n_eval), replace=False)
        train_roots, eval_roots = pick[:ntr], pick[ntr:ntr + n_eval]
        if name == "rmat22_undirected":
            eval_roots = bench_roots(g)
            train_roots = np.setdiff1d(pick, eval_roots)[:ntr]
        log(f"{name}: nv={g.nv()} nnz={len(g.csr()[1])} built in {time.time() - t0:.1f}s")
        rec = collect_pipelined(g, train_roots, log=lambda m, n=name: log(f"{n}: {m}"))
        recs[name] = rec
        built[name] = (g, eval_roots)
        log(f"{name}: {len(rec)} probed levels ({time.time() - t0:.1f}s)")
    x, y, w = labels(list(recs.values()))
    params, loss = train(x, y, w, steps=3000, lr=3e-3, seed=0)
    log(f"trained on {len(y)} levels, loss {loss:.4f} ({time.time() - t0:.1f}s)")
    out = {"train_levels": int(len(y)), "final_loss": loss, "graphs": {}}
    tot_h = tot_l = 0.0
    all_same = True
    for name, (g, eval_roots) in built.items():
        th, tl, same = evaluate(g, eval_roots, params, reps=1)
        rec = recs[name]
        choice_l = (logits(rec[:, :16], params).argmax(axis=1))
        choice_h = heuristic_choice(rec)
        best = decision_cost(rec, (rec[:, 17] < rec[:, 16]).astype(int))
        out["graphs"][name] = {
            "nv": g.nv(), "held_out_roots": int(len(eval_roots)),
            "sum_ms_heuristic": float(th.sum()), "sum_ms_learned": float(tl.sum()),
            "ratio": float(th.sum() / tl.sum()), "outputs_identical": same,
            "train_probe_cost_us": {"heuristic": decision_cost(rec, choice_h), "learned": decision_cost(rec, choice_l),
                                    "oracle_best": best},
        }
        tot_h += th.sum()
        tot_l += tl.sum()
        all_same = all_same and same
        log(f"{name}: heuristic {th.sum():.1f} ms, learned {tl.sum():.1f} ms, ratio {th.sum() / tl.sum():.3f}, "
            f"identical={same}")
    out["ratio_all"] = float(tot_h / tot_l)
    out["outputs_identical"] = all_same
    out["seconds"] = time.time() - t0
    save(params, {"train_levels": int(len(y)), "loss": loss, "features": FEATURES})
    os.makedirs(out_dir, exist_ok=True)
    save(params, {"train_levels": int(len(y)), "loss": loss, "features": FEATURES},
         os.path.join(out_dir, "policy_weights.npz"))
    with open(os.path.join(out_dir, "direction_policy.json"), "w") as f:
        json.dump(out, f, indent=1)
    return out


if __name__ == "__main__":
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="train and evaluate the learned BFS direction policy")
    ap.add_argument("--out", default="policy_out")
    ap.add_argument("--train-roots", type=int, default=2000)
    ap.add_argument("--eval-roots", type=int, default=200)
    ap.add_argument("--grid-train-roots", type=int, default=40)
    a = ap.parse_args()
    res = config5(a.out, a.train_roots, a.eval_roots, a.grid_train_roots,
                  log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))
