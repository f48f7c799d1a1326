"""The match_* binaries' coarsening (heavy-edge matching: sgp_coarsen_match,
heuristics.h:1584-1735, 1890-1895; -DPUREMATCH) on the GPU against the
oracle: single levels (levels 1 and 2: random and first-heaviest
neighbours) and hierarchies bit-exact, bisections (power iteration:
tolerance parity; FM: bit-exact), the refusals, and the -DPUREMATCH
drop-ins."""
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
def mt(ctx):
    ctx.set_coarsening(1)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(31)
    yield "grid64", *graphs.grid2d(64)
    yield "rgg", *graphs.rgg(9000, seed=3)
    yield "rmat_w", *graphs.rmat(13, 16, seed=5)
    yield "rmat_hubs", *graphs.rmat(11, 48, seed=7)
    rp, col, w = random_graph(rng, 4000, 0.002, weighted=True, lcc=True)
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_match_level_bitexact(mt, case):
    name, rp, col, w = case
    g = mt.graph(rp, col, w)
    for level, seed in ((1, 2), (2, 1234)):
        vc, nc, vperm, hn = g.hec_level(level, seed)
        ovc, onc, _ = oracle.coarsen_match(level, rp, col, w, seed)
        assert nc == onc and np.array_equal(vc, ovc)
        ovp, _, _ = oracle.generate_permutation(len(rp) - 1, seed)
        assert np.array_equal(vperm, ovp)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_match_hierarchy_bitexact(mt, case):
    name, rp, col, w = case
    H = mt.graph(rp, col, w).hierarchy(41)
    O = oracle.Hierarchy(rp, col, w, 41, variant=1)
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


def test_match_refusals(mt):
    import ipdps_2021_graph_coarsening_amd as mi
    rp = np.array([0, 1, 2, 2], np.uint32)  # vertex 2 has no entries
    col = np.array([1, 0], np.uint32)
    with pytest.raises(mi.MiError):
        mt.graph(rp, col).hec_level(1, 1)
    rp = np.array([0, 1, 2, 3], np.uint32)  # vertex 2: a self-loop only
    col = np.array([1, 0, 2], np.uint32)
    with pytest.raises(mi.MiError):
        mt.graph(rp, col).hec_level(2, 1)


def test_match_bisection_vs_oracle(mt):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.rmat(14, 16, seed=9)
    r = mi.partition_graph(mt, mt.graph(rp, col, w), hec_seed=8, pcg_state=8)
    o = oracle.partition_graph(rp, col, w, 8, 8, variant=1)
    L = r.stats["nlevels"]
    assert list(r.stats["level_n"][:L]) == [int(x) for x in o["levels_n"]]
    for a, b in zip(r.stats["level_iters"][:L], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(r.edge_cut - o["edge_cut"]) <= max(2, 0.01 * o["edge_cut"])


def test_match_fm_bisection_bitexact(mt):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, _ = graphs.rgg(15000, seed=13)
    mt.set_refinement(True)
    try:
        r = mi.partition_graph(mt, mt.graph(rp, col), hec_seed=14, pcg_state=1)
    finally:
        mt.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 14, variant=1)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


@pytest.mark.parametrize("fm", [False, True])
def test_dropin_match(tmp_path, fm):
    exe = mibuild.DROPIN_MATCH_FM_EXE if fm else mibuild.DROPIN_MATCH_EXE
    if not os.path.exists(exe):
        pytest.skip("the drop-in was not built (it needs the reference's cmg headers at build time)")
    rp, col, _ = graphs.rgg(12000, seed=19)
    csr, cfg = tmp_path / "g.csr", tmp_path / "config.txt"
    graphs.write_csr(str(csr), rp, col)
    cfg.write_text("0 0 0 2 1 1e-10\n")
    metrics, part = tmp_path / "m.json", tmp_path / "p.txt"
    r = subprocess.run([exe, str(csr), str(metrics), str(cfg), "-", str(part)],
                       env=dict(os.environ, SGPAR_FIXED_TIME="33"), capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    (t,) = json.loads(metrics.read_text())
    if fm:
        o = oracle.partition_graph_fm(rp, col, None, 33, variant=1)
        assert t["edge-cut"] == o["edge_cut"]
        got = np.loadtxt(part, dtype=np.int64)
        assert np.array_equal(got, o["part"].astype(np.int64))
    else:
        o = oracle.partition_graph(rp, col, None, 33, 33, variant=1)
        assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
