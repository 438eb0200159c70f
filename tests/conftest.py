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
