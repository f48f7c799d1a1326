"""The mis_* binaries' coarsening (MIS-2 aggregates: mis_2,
heuristics.h:8-100, and sgp_coarsen_mis_2, :218-289; -DMIS) on the GPU
against the oracle: single levels and hierarchies bit-exact, bisections
(power iteration: tolerance parity; FM: bit-exact) and the -DMIS drop-in."""
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
def mis(ctx):
    ctx.set_coarsening(5)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(23)
    yield "grid64", *graphs.grid2d(64)
    yield "rgg", *graphs.rgg(9000, seed=3)
    yield "rmat_w", *graphs.rmat(13, 16, seed=5)
    rp, col, w = random_graph(rng, 4000, 0.002, weighted=True)  # isolated vertices
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_mis2_level_bitexact(mis, case):
    name, rp, col, w = case
    g = mis.graph(rp, col, w)
    for seed in (2, 1234):
        vc, nc, _, _ = g.hec_level(1, seed)
        ovc, onc = oracle.coarsen_mis2(rp, col, seed)
        assert nc == onc and np.array_equal(vc, ovc)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_mis2_hierarchy_bitexact(mis, case):
    name, rp, col, w = case
    H = mis.graph(rp, col, w).hierarchy(41)
    O = oracle.Hierarchy(rp, col, w, 41, variant=5)
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


def test_mis2_bisection_vs_oracle(mis):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.rgg(20000, seed=11)
    r = mi.partition_graph(mis, mis.graph(rp, col, w), hec_seed=8, pcg_state=8)
    o = oracle.partition_graph(rp, col, w, 8, 8, variant=5)
    L = r.stats["nlevels"]
    assert list(r.stats["level_n"][:L]) == [int(x) for x in o["levels_n"]]
    for a, b in zip(r.stats["level_iters"][:L], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(r.edge_cut - o["edge_cut"]) <= max(2, 0.01 * o["edge_cut"])


def test_mis2_fm_bisection_bitexact(mis):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, _ = graphs.rgg(15000, seed=13)
    mis.set_refinement(True)
    try:
        r = mi.partition_graph(mis, mis.graph(rp, col), hec_seed=14, pcg_state=1)
    finally:
        mis.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 14, variant=5)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


def test_dropin_mis(tmp_path):
    if not os.path.exists(mibuild.DROPIN_MIS_EXE):
        pytest.skip("the drop-in was not built (it needs the reference's cmg headers at build time)")
    rp, col, _ = graphs.rgg(12000, seed=19)
    csr, cfg = tmp_path / "g.csr", tmp_path / "config.txt"
    graphs.write_csr(str(csr), rp, col)
    cfg.write_text("0 0 0 2 1 1e-10\n")
    metrics, part = tmp_path / "m.json", tmp_path / "p.txt"
    r = subprocess.run([mibuild.DROPIN_MIS_EXE, str(csr), str(metrics), str(cfg), "-", str(part)],
                       env=dict(os.environ, SGPAR_FIXED_TIME="33"), capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    (t,) = json.loads(metrics.read_text())
    o = oracle.partition_graph(rp, col, None, 33, 33, variant=5)
    assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
