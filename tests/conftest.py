import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "lightgraphs.jl_amd"))
sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblgmi.so on the device)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc
    orc.build()
    return orc


def csr_of(n, src, dst, directed, drop_self_loops=False):
    """Sorted CSR (+ in-CSR) of an edge list through the package builder."""
    from lgmi.graphs import csr_from_edges, csr_transpose
    rp, col = csr_from_edges(n, np.asarray(src), np.asarray(dst), symmetrize=not directed,
                             drop_self_loops=drop_self_loops)
    if directed:
        irp, icol = csr_transpose(rp, col)
    else:
        irp, icol = rp, col
    return rp, col, irp, icol


GOLDEN_NX = ["nx_er_300_1200.npz", "nx_ba_400_4.npz", "nx_ws_400_8.npz", "nx_rmat9_sym.npz",
             "nx_rmat8_dir.npz", "nx_rand_dir_200.npz"]


# Layer sizes of a graph whose shortest-path counts outgrow FP32's 24-bit
# significand: consecutive layers are completely joined, so from the single
# vertex of layer 0 the count to layer k is the product of the sizes of
# layers 1..k-1 (15^7 = 170859375 > 2^24 by layer 8, odd, so FP32 rounds it;
# the largest, about 1.1e14, stays below 2^53).  A 40-vertex path hangs off
# the last layer.
LAYERS = [1, 15, 15, 13, 15, 15, 14, 15, 15, 15, 15, 15, 3]


def layered_edges(tail=40):
    """0-based undirected edge list (src, dst) and vertex count of the
    layered graph above."""
    src, dst, start = [], [], []
    n = 0
    for s in LAYERS:
        start.append(n)
        n += s
    for i in range(len(LAYERS) - 1):
        for a in range(start[i], start[i] + LAYERS[i]):
            for b in range(start[i + 1], start[i + 1] + LAYERS[i + 1]):
                src.append(a)
                dst.append(b)
    prev = start[-1]
    for t in range(tail):
        src.append(prev)
        dst.append(n)
        prev = n
        n += 1
    return n, np.array(src, np.int64), np.array(dst, np.int64)
