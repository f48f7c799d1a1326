"""The gosh2_* binaries' coarsening (GOSH v2: sgp_coarsen_GOSH_v2,
heuristics.h:444-601, with parallel_map_construct_prefilled, :378-442;
-DGOSH_V2) on the GPU against the oracle: single levels and hierarchies
bit-exact on graphs with and without seeds, bisections (power iteration:
tolerance parity; FM: bit-exact) and the -DGOSH_V2 drop-in."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import build as mibuild
from ipdps_2021_graph_coarsening_amd import graphs

pytestmark = pytest.mark.gpu


@pytest.fixture
def g2(ctx):
    ctx.set_coarsening(7)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(43)
    yield "grid64", *graphs.grid2d(64)
    yield "rgg", *graphs.rgg(9000, seed=3)
    yield "rmat_w", *graphs.rmat(13, 16, seed=5)
    yield "rmat_dense", *graphs.rmat(11, 64, seed=7)
    rp, col, w = random_graph(rng, 4000, 0.002, weighted=True, lcc=True)
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_gosh2_level_bitexact(g2, case):
    name, rp, col, w = case
    g = g2.graph(rp, col, w)
    for seed in (2, 1234):
        vc, nc, _, _ = g.hec_level(1, seed)
        ovc, onc = oracle.coarsen_gosh2(rp, col, w, seed)
        assert nc == onc and np.array_equal(vc, ovc)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_gosh2_hierarchy_bitexact(g2, case):
    name, rp, col, w = case
    H = g2.graph(rp, col, w).hierarchy(41)
    O = oracle.Hierarchy(rp, col, w, 41, variant=7)
    assert H.nlevels == O.nlevels
    for l in range(H.nlevels):
        a, b = H.level(l), O.levels[l]
        assert a["n"] == b["n"] and a["m"] == b["m"]
        assert np.array_equal(a["rowptr"], b["rowptr"]) and np.array_equal(a["col"], b["col"])
        assert np.array_equal(a["vtx_w"], b["vtx_w"])
        if l:
            assert np.array_equal(a["w"], b["w"])
        if l + 1 < H.nlevels:
            assert np.array_equal(a["vcmap"], b["vcmap"])


def test_gosh2_bisection_vs_oracle(g2):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.rmat(14, 16, seed=9)
    r = mi.partition_graph(g2, g2.graph(rp, col, w), hec_seed=8, pcg_state=8)
    o = oracle.partition_graph(rp, col, w, 8, 8, variant=7)
    L = r.stats["nlevels"]
    assert list(r.stats["level_n"][:L]) == [int(x) for x in o["levels_n"]]
    for a, b in zip(r.stats["level_iters"][:L], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(r.edge_cut - o["edge_cut"]) <= max(2, 0.01 * o["edge_cut"])


def test_gosh2_fm_bisection_bitexact(g2):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, _ = graphs.rmat(14, 8, seed=13)
    g2.set_refinement(True)
    try:
        r = mi.partition_graph(g2, g2.graph(rp, col), hec_seed=14, pcg_state=1)
    finally:
        g2.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 14, variant=7)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


def test_dropin_gosh2(tmp_path):
    if not os.path.exists(mibuild.DROPIN_GOSH2_EXE):
        pytest.skip("the drop-in was not built (it needs the reference's cmg headers at build time)")
    rp, col, _ = graphs.rmat(13, 8, seed=19)
    csr, cfg = tmp_path / "g.csr", tmp_path / "config.txt"
    graphs.write_csr(str(csr), rp, col)
    cfg.write_text("0 0 0 2 1 1e-10\n")
    metrics, part = tmp_path / "m.json", tmp_path / "p.txt"
    r = subprocess.run([mibuild.DROPIN_GOSH2_EXE, str(csr), str(metrics), str(cfg), "-", str(part)],
                       env=dict(os.environ, SGPAR_FIXED_TIME="33"), capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    (t,) = json.loads(metrics.read_text())
    o = oracle.partition_graph(rp, col, None, 33, 33, variant=7)
    assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
