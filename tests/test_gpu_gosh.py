"""The gosh_* binaries' coarsening (GOSH centres: GOSH_clusters,
heuristics.h:102-217, and sgp_coarsen_GOSH, :312-377; -DGOSH) on the GPU
against the oracle: single levels and hierarchies bit-exact, bisections
(power iteration: tolerance parity; FM: bit-exact) and the -DGOSH drop-in."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import build as mibuild
from ipdps_2021_graph_coarsening_amd import graphs
from test_oracle_gosh import _csr_from_rows

pytestmark = pytest.mark.gpu


@pytest.fixture
def gosh(ctx):
    ctx.set_coarsening(6)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(29)
    yield "grid64", *graphs.grid2d(64)
    yield "rgg", *graphs.rgg(9000, seed=3)
    yield "rmat_w", *graphs.rmat(13, 16, seed=5)
    yield "rmat_hubs", *graphs.rmat(11, 48, seed=7)
    rp, col, w = random_graph(rng, 4000, 0.002, weighted=True)  # isolated vertices
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_gosh_level_bitexact(gosh, case):
    name, rp, col, w = case
    g = gosh.graph(rp, col, w)
    vc, nc, _, _ = g.hec_level(1, 2)
    ovc, onc = oracle.coarsen_gosh(rp, col)
    assert nc == onc and np.array_equal(vc, ovc)


def test_gosh_level_multigraph_bitexact(gosh):
    """Self-loops and repeated neighbours count in a vertex's entries; rows
    longer than a wavefront take several steps."""
    rng = np.random.default_rng(6)
    n = 3000
    rows = [[] for _ in range(n)]
    for a, b in rng.integers(0, n, (9000, 2)):
        rows[a].append(int(b))
        if a != b:
            rows[b].append(int(a))
    for h in (5, 77):
        for b in rng.integers(0, n, 400):
            rows[h].append(int(b))
            rows[int(b)].append(h)
    for i in range(0, n, 9):
        rows[i].append(i)
    rp, col = _csr_from_rows(rows)
    vc, nc, _, _ = gosh.graph(rp, col).hec_level(1, 1)
    ovc, onc = oracle.coarsen_gosh(rp, col)
    assert nc == onc and np.array_equal(vc, ovc)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_gosh_hierarchy_bitexact(gosh, case):
    name, rp, col, w = case
    H = gosh.graph(rp, col, w).hierarchy(41)
    O = oracle.Hierarchy(rp, col, w, 41, variant=6)
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


def test_gosh_bisection_vs_oracle(gosh):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.rmat(14, 16, seed=9)
    r = mi.partition_graph(gosh, gosh.graph(rp, col, w), hec_seed=8, pcg_state=8)
    o = oracle.partition_graph(rp, col, w, 8, 8, variant=6)
    L = r.stats["nlevels"]
    assert list(r.stats["level_n"][:L]) == [int(x) for x in o["levels_n"]]
    for a, b in zip(r.stats["level_iters"][:L], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(r.edge_cut - o["edge_cut"]) <= max(2, 0.01 * o["edge_cut"])


def test_gosh_fm_bisection_bitexact(gosh):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, _ = graphs.rgg(15000, seed=13)
    gosh.set_refinement(True)
    try:
        r = mi.partition_graph(gosh, gosh.graph(rp, col), hec_seed=14, pcg_state=1)
    finally:
        gosh.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 14, variant=6)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


def test_dropin_gosh(tmp_path):
    if not os.path.exists(mibuild.DROPIN_GOSH_EXE):
        pytest.skip("the drop-in was not built (it needs the reference's cmg headers at build time)")
    rp, col, _ = graphs.rgg(12000, seed=19)
    csr, cfg = tmp_path / "g.csr", tmp_path / "config.txt"
    graphs.write_csr(str(csr), rp, col)
    cfg.write_text("0 0 0 2 1 1e-10\n")
    metrics, part = tmp_path / "m.json", tmp_path / "p.txt"
    r = subprocess.run([mibuild.DROPIN_GOSH_EXE, str(csr), str(metrics), str(cfg), "-", str(part)],
                       env=dict(os.environ, SGPAR_FIXED_TIME="33"), capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    (t,) = json.loads(metrics.read_text())
    o = oracle.partition_graph(rp, col, None, 33, 33, variant=6)
    assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
