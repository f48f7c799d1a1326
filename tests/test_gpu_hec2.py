"""The hec2_* binaries' coarsening (HEC heavy neighbours with
parallel_map_construct_v2, heuristics.h:695-764, -DHEC_V2) on the GPU
against the oracle: levels (1 and 2) and hierarchies bit-exact -- coarse
vertices left without members included -- a bisection (power iteration:
tolerance parity), an FM bisection bit-exact, and the -DHEC_V2 drop-ins."""
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
def v2(ctx):
    ctx.set_coarsening(4)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(19)
    yield "grid64", *graphs.grid2d(64)
    yield "rgg", *graphs.rgg(9000, seed=2)
    yield "rmat_w", *graphs.rmat(13, 16, seed=4)
    rp, col, w = random_graph(rng, 4000, 0.002, weighted=True)  # isolated vertices: random heavy picks
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_hec2_level_bitexact(v2, case):
    name, rp, col, w = case
    g = v2.graph(rp, col, w)
    for level, seed in ((1, 5), (2, 9)):
        vc, nc, _, _ = g.hec_level(level, seed)
        ovc, onc, _ = oracle.coarsen_hec2(level, rp, col, w, seed)
        assert nc == onc and np.array_equal(vc, ovc)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_hec2_hierarchy_bitexact(v2, case):
    name, rp, col, w = case
    H = v2.graph(rp, col, w).hierarchy(31)
    O = oracle.Hierarchy(rp, col, w, 31, variant=4)
    assert H.nlevels == O.nlevels
    empty = 0
    for l in range(H.nlevels):
        a, b = H.level(l), O.levels[l]
        assert a["n"] == b["n"] and a["m"] == b["m"]
        assert np.array_equal(a["rowptr"], b["rowptr"]) and np.array_equal(a["col"], b["col"])
        assert np.array_equal(a["vtx_w"], b["vtx_w"])
        empty += int((a["vtx_w"] == 0).sum())
        if l:
            assert np.array_equal(a["w"], b["w"])
        if l + 1 < H.nlevels:
            assert np.array_equal(a["vcmap"], b["vcmap"])
    assert empty > 0  # ids left without members, as the reference leaves them


def test_hec2_bisection_vs_oracle(v2):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.rgg(20000, seed=11)
    r = mi.partition_graph(v2, v2.graph(rp, col, w), hec_seed=8, pcg_state=8)
    o = oracle.partition_graph(rp, col, w, 8, 8, variant=4)
    L = r.stats["nlevels"]
    assert list(r.stats["level_n"][:L]) == [int(x) for x in o["levels_n"]]
    for a, b in zip(r.stats["level_iters"][:L], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(r.edge_cut - o["edge_cut"]) <= max(2, 0.01 * o["edge_cut"])


def test_hec2_fm_bisection_bitexact(v2):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, _ = graphs.rgg(15000, seed=13)
    v2.set_refinement(True)
    try:
        r = mi.partition_graph(v2, v2.graph(rp, col), hec_seed=14, pcg_state=1)
    finally:
        v2.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 14, variant=4)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


@pytest.mark.parametrize("fm", [False, True])
def test_dropin_hec2(tmp_path, fm):
    exe = mibuild.DROPIN_HEC2_FM_EXE if fm else mibuild.DROPIN_HEC2_EXE
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
        o = oracle.partition_graph_fm(rp, col, None, 33, variant=4)
        assert t["edge-cut"] == o["edge_cut"]
        assert np.array_equal(np.loadtxt(part, dtype=np.int64), o["part"].astype(np.int64))
    else:
        o = oracle.partition_graph(rp, col, None, 33, 33, variant=4)
        assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
