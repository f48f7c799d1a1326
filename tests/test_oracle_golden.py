"""Pins the oracle against the recorded reference Serial run (fixture in
tests/golden) -- the check that fixes the XorShift pool seeding, the
permutation, both heavy-neighbour rules, the map construction, the coarse
build, the FP64 power iteration and the split all at once."""
import json
import os

import numpy as np

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_grid256_serial_run(grid256):
    g = json.load(open(os.path.join(GOLD, "grid256_serial_seed12345.json")))
    rp, col = grid256
    assert len(rp) - 1 == g["n"] and int(rp[-1]) == g["m_stored"]
    r = oracle.partition_graph(rp, col, None, g["hec_seed"], g["pcg_state"], g["pcg_inc"],
                               g["refine_alg"], g["tol"])
    assert list(r["levels_n"]) == g["levels_n_fine_to_coarse"]
    assert list(r["iters"]) == g["power_iterations_fine_to_coarse"]
    assert r["edge_cut"] == g["edge_cut"]
    part = r["part"]
    assert set(np.unique(part)) <= {0, 1}
    assert abs(int(part.sum()) - g["n"] // 2) <= 1


def test_golden_sensitivity_to_seed(grid256):
    """A different seed must change the hierarchy: the golden match above is
    not an accident of the level sizes being seed-independent."""
    rp, col = grid256
    a = oracle.partition_graph(rp, col, None, 12345, 12345)
    b = oracle.partition_graph(rp, col, None, 12346, 12345)
    assert list(a["levels_n"]) != list(b["levels_n"])
