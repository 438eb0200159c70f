"""Uneven small-graph betweenness batch: launch-threshold autotuner.

The batch kernel (lgmi_betweenness_batch) runs a graph either one source
per wavefront (n <= wave_max_n) or one source per work-group, expands
adjacency rows with lane groups of 1 to 8 lanes taking 2 to 16 entries
each per round (a whole wavefront for longer rows), and splits graphs with
n * (nnz + n) above a work target into several source ranges.  autotune()
times candidate (wave_max_n, work target, lane group, slots) settings on a
seeded sample of the
batch and keeps the fastest; the choice is written to a JSON
file next to this module and applied with lgmi_betweenness_batch_tuning.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .centrality import batch_tuning, betweenness_centrality_batch

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_tuning.json")


def work(graphs) -> float:
    return float(sum(g.nv() * (len(g.csr()[1]) + g.nv()) for g in graphs))


def autotune(graphs, sample: int = 512, seed: int = 0, candidates=None, log=None) -> dict:
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(graphs), size=min(sample, len(graphs)), replace=False)
    sub = [graphs[i] for i in idx]
    candidates = candidates or [(wmax, 0.0, kg, sl) for kg, sl in ((1, 8), (1, 16), (2, 4), (2, 8), (4, 4))
                                for wmax in (0, -1)]
    results = []
    for wmax, target, kg, sl in candidates:
        batch_tuning(wmax, target, -1, kg, sl)
        betweenness_centrality_batch(sub[:16])                 # warm up
        _, ms = betweenness_centrality_batch(sub, return_ms=True)
        results.append({"wave_max_n": wmax, "task_work_target": target, "lane_group": kg, "slots": sl,
                        "device_ms": ms})
        if log:
            log(f"autotune: wave_max_n={wmax} target={target:g} lane_group={kg} slots={sl}: {ms:.1f} ms")
    best = min(results, key=lambda r: r["device_ms"])
    batch_tuning(best["wave_max_n"], best["task_work_target"], -1, best["lane_group"], best["slots"])
    out = {"best": best, "candidates": results, "sample_graphs": len(sub)}
    return out


def save_tuning(result: dict, path: str = TUNING_PATH) -> None:
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load_tuning(path: str = TUNING_PATH) -> bool:
    if not os.path.exists(path):
        return False
    with open(path) as f:
        best = json.load(f)["best"]
    batch_tuning(best["wave_max_n"], best["task_work_target"], -1, best.get("lane_group", 0), best.get("slots", 0))
    return True
