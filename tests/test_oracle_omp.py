"""The OpenMP restatement (oracle/sgpar_omp.c, the multi-core CPU baseline)
against the Serial-semantics oracle: the same hierarchy (level sizes), the
same continuing PCG state, iteration counts within 10 % and the same cut
up to 1 % (its dot products are OpenMP reductions, another summation order)."""
import numpy as np
import pytest

import oracle
from ipdps_2021_graph_coarsening_amd import graphs


@pytest.mark.parametrize("name", ["grid256", "rmat14w", "rgg60k", "path3000"])
def test_omp_port_matches_oracle(name):
    if not oracle.omp_available():
        pytest.skip("libsgpar_omp.so not built")
    if name == "grid256":
        (rp, col, _), w, seed = graphs.grid2d(256), None, 12345
    elif name == "rmat14w":
        rp, col, w = graphs.rmat(14, 16, seed=1)
        seed = 1
    elif name == "rgg60k":
        (rp, col, _), w, seed = graphs.rgg(60000, seed=5), None, 3
    else:
        (rp, col, _), w, seed = graphs.path(3000), None, 2
    o = oracle.partition_graph(rp, col, w, seed, seed)
    p = oracle.omp_partition_graph(rp, col, w, seed, seed)
    assert list(p["levels_n"]) == list(o["levels_n"])
    assert p["pcg_state"] == o["pcg_state"]
    for a, b in zip(p["iters"], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(p["edge_cut"] - o["edge_cut"]) <= max(1, 0.01 * o["edge_cut"])
    n = len(rp) - 1
    assert np.count_nonzero(p["part"] == 0) == n // 2
    src = np.repeat(np.arange(n), np.diff(rp))
    assert p["edge_cut"] == int(np.count_nonzero(p["part"][src] != p["part"][col])) // 2
