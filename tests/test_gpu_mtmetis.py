"""The mtmetis_* binaries' coarsening (heavy-edge matching, then the leaf,
twin and relative passes of heuristics.h:1737-1888; -DMTMETIS) on the GPU
against the oracle: single levels (1 and 2) and hierarchies bit-exact on
graphs where every pass pairs vertices (R-MATs, stars joined in a path,
K(2, m) blocks whose middles are twins), an FM bisection bit-exact, and the
-DMTMETIS drop-in."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import build as mibuild
from ipdps_2021_graph_coarsening_amd import graphs
from test_oracle_mtmetis import _star_forest, twin_blocks

pytestmark = pytest.mark.gpu


@pytest.fixture
def mm(ctx):
    ctx.set_coarsening(2)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(37)
    yield "rmat", *graphs.rmat(13, 8, seed=3)
    yield "rmat_w", *graphs.rmat(12, 4, seed=9)
    yield "stars", *_star_forest(rng)
    yield "twins", *twin_blocks(rng, 30)
    rp, col, w = random_graph(rng, 3000, 0.002, weighted=True, lcc=True)
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_mtmetis_level_bitexact(mm, case):
    name, rp, col, w = case
    g = mm.graph(rp, col, w)
    for level, seed in ((1, 2), (2, 77)):
        vc, nc, _, _ = g.hec_level(level, seed)
        ovc, onc, _ = oracle.coarsen_mtmetis(level, rp, col, w, seed)
        assert nc == onc and np.array_equal(vc, ovc)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_mtmetis_hierarchy_bitexact(mm, case):
    name, rp, col, w = case
    H = mm.graph(rp, col, w).hierarchy(41)
    O = oracle.Hierarchy(rp, col, w, 41, variant=2)
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


def test_mtmetis_fm_bisection_bitexact(mm):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, _ = graphs.rmat(14, 8, seed=13)
    mm.set_refinement(True)
    try:
        r = mi.partition_graph(mm, mm.graph(rp, col), hec_seed=14, pcg_state=1)
    finally:
        mm.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 14, variant=2)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


def test_dropin_mtmetis(tmp_path):
    if not os.path.exists(mibuild.DROPIN_MTMETIS_EXE):
        pytest.skip("the drop-in was not built (it needs the reference's cmg headers at build time)")
    rp, col, _ = graphs.rmat(13, 8, seed=19)
    csr, cfg = tmp_path / "g.csr", tmp_path / "config.txt"
    graphs.write_csr(str(csr), rp, col)
    cfg.write_text("0 0 0 2 1 1e-10\n")
    metrics, part = tmp_path / "m.json", tmp_path / "p.txt"
    r = subprocess.run([mibuild.DROPIN_MTMETIS_EXE, str(csr), str(metrics), str(cfg), "-", str(part)],
                       env=dict(os.environ, SGPAR_FIXED_TIME="33"), capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    (t,) = json.loads(metrics.read_text())
    o = oracle.partition_graph(rp, col, None, 33, 33, variant=2)
    assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
