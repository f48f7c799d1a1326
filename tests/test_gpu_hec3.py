"""The hec3_* binaries' coarsening (HEC heavy neighbours with
parallel_map_construct_v3, heuristics.h:766-826, -DHEC_V3) on the GPU
against the oracle: levels and hierarchies bit-exact, bisections (power
iteration: tolerance parity; FM: bit-exact), and the -DHEC_V3 drop-ins."""
import os

import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import build as mibuild
from ipdps_2021_graph_coarsening_amd import graphs

pytestmark = pytest.mark.gpu


@pytest.fixture
def v3(ctx):
    ctx.set_coarsening(3)
    yield ctx
    ctx.set_coarsening(0)


def _graphs():
    rng = np.random.default_rng(17)
    yield "grid64", *graphs.grid2d(64)
    yield "rgg", *graphs.rgg(9000, seed=2)
    yield "rmat_w", *graphs.rmat(13, 16, seed=4)
    rp, col, w = random_graph(rng, 4000, 0.002, weighted=True)  # isolated vertices: random heavy picks
    yield "er_w", rp, col, w


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_hec3_level_bitexact(v3, case):
    name, rp, col, w = case
    g = v3.graph(rp, col, w)
    for level, seed in ((1, 5), (2, 9)):
        vc, nc, _, hn = g.hec_level(level, seed)
        ovc, onc, ohn = oracle.coarsen_hec3(level, rp, col, w, seed)
        assert np.array_equal(hn, ohn)
        assert nc == onc and np.array_equal(vc, ovc)


@pytest.mark.parametrize("case", list(_graphs()), ids=lambda c: c[0])
def test_hec3_hierarchy_bitexact(v3, case):
    name, rp, col, w = case
    H = v3.graph(rp, col, w).hierarchy(31)
    O = oracle.Hierarchy(rp, col, w, 31, variant=3)
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


def test_hec3_bisection_vs_oracle(v3):
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.rgg(20000, seed=6)
    r = mi.partition_graph(v3, v3.graph(rp, col, w), hec_seed=8, pcg_state=8)
    o = oracle.partition_graph(rp, col, w, 8, 8, variant=3)
    L = r.stats["nlevels"]
    assert list(r.stats["level_n"][:L]) == [int(x) for x in o["levels_n"]]
    for a, b in zip(r.stats["level_iters"][:L], o["iters"]):
        assert abs(int(a) - int(b)) <= max(2, 0.1 * int(b))
    assert abs(r.edge_cut - o["edge_cut"]) <= max(2, 0.01 * o["edge_cut"])
    assert r.pcg_state == o["pcg_state"]


@pytest.mark.parametrize("name", ["grid96", "rgg"])
def test_hec3_fm_bisection_bitexact(v3, name):
    import ipdps_2021_graph_coarsening_amd as mi
    if name == "grid96":
        rp, col = graphs.grid2d(96)[:2]
    else:
        rp, col, _ = graphs.rgg(15000, seed=7)
    v3.set_refinement(True)
    try:
        r = mi.partition_graph(v3, v3.graph(rp, col), hec_seed=12, pcg_state=1)
    finally:
        v3.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 12, variant=3)
    assert r.edge_cut == o["edge_cut"] and np.array_equal(r.part, o["part"])


def test_dropin_hec3_variants(tmp_path):
    """The drop-ins compiled with -DHEC_V3 (and -DFM): the hec3_spec and
    hec3_fm flows."""
    import json
    import subprocess
    if not (os.path.exists(mibuild.DROPIN_HEC3_EXE) and os.path.exists(mibuild.DROPIN_HEC3_FM_EXE)):
        pytest.skip("the drop-ins were not built (they need the reference's cmg headers at build time)")
    rp, col, _ = graphs.rgg(12000, seed=9)
    csr, cfg = tmp_path / "g.csr", tmp_path / "config.txt"
    graphs.write_csr(str(csr), rp, col)
    cfg.write_text("0 0 0 2 1 1e-10\n")
    out = {}
    for tag, exe in (("spec", mibuild.DROPIN_HEC3_EXE), ("fm", mibuild.DROPIN_HEC3_FM_EXE)):
        metrics, part = tmp_path / f"m_{tag}.json", tmp_path / f"p_{tag}.txt"
        r = subprocess.run([exe, str(csr), str(metrics), str(cfg), "-", str(part)],
                           env=dict(os.environ, SGPAR_FIXED_TIME="21"), capture_output=True, text=True, timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        out[tag] = (json.loads(metrics.read_text()), np.loadtxt(part, dtype=np.uint32, ndmin=1))
    o = oracle.partition_graph(rp, col, None, 21, 21, variant=3)
    (t,) = out["spec"][0]
    assert [l["number-vertices"] for l in t["coarse-levels"]] == [int(x) for x in o["levels_n"]][::-1]
    of = oracle.partition_graph_fm(rp, col, None, 21, variant=3)
    (tf,) = out["fm"][0]
    assert tf["edge-cut"] == of["edge_cut"] and np.array_equal(out["fm"][1], of["part"])
