"""Shared fixtures.  Tests marked ``gpu`` need a HIP device and the built
library; everything else runs on the CPU (oracle, host code, ABI surface)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libsgpar_mi.so")


@pytest.fixture(scope="session")
def ctx():
    import ipdps_2021_graph_coarsening_amd as mi
    c = mi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def grid256():
    from ipdps_2021_graph_coarsening_amd import graphs
    rp, col, _ = graphs.grid2d(256)
    return rp, col


def random_graph(rng, n, p, weighted=False, lcc=False):
    """Erdos-Renyi-like symmetric test graph."""
    from ipdps_2021_graph_coarsening_amd import graphs
    m = max(1, int(p * n * (n - 1) / 2))
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    w = rng.integers(1, 9, m) if weighted else None
    return graphs.from_edges(n, u, v, w, lcc=lcc)
