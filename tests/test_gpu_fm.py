"""FM refinement path on the GPU against the oracle (eigensolve_kokkos.h:208-1368,
the FM branch of sgp_eigensolve :1506-1587, the hec_fm bisection of
sgpar.h:2363-2431): one fm_refine call, the GGGP and spectral initial
partitions, the multilevel FM eigensolve and the whole FM bisection.  Every
comparison is bit-exact (integer work) except the spectral initial
partition, whose power iteration matches the oracle to its tolerance
(see test_fm_init_spectral)."""
import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import graphs
from test_oracle_fm import _cases, _cut2, _two_cliques

pytestmark = pytest.mark.gpu


def _check_refine(ctx, rp, col, w, vw, part):
    g = ctx.graph(rp, col, w)
    got, cut, info = g.fm_refine(part, vw)
    want, ocut, oinfo = oracle.fm_refine(rp, col, w, vw, part)
    assert (info["swaps"], info["argmin"], info["min_imb"], info["maxE"]) == \
        (oinfo["swaps"], oinfo["argmin"], oinfo["min_imb"], oinfo["maxE"])
    assert cut == ocut
    assert np.array_equal(got, want.astype(np.uint32))
    return got, cut


@pytest.mark.parametrize("case", range(24))
def test_fm_refine_random_vs_oracle(ctx, case):
    rp, col, w, vw, part = _cases()[case]
    _check_refine(ctx, rp, col, w, vw, part)


@pytest.mark.parametrize("k", [40, 200])
def test_fm_refine_grid_vs_oracle(ctx, k):
    rp, col = graphs.grid2d(k)[:2]
    rng = np.random.default_rng(k)
    n = len(rp) - 1
    for part in ((rng.random(n) < 0.5), np.arange(n) % k < k // 2, rng.random(n) < 0.1):
        _check_refine(ctx, rp, col, None, None, part.astype(np.uint8))


def test_fm_refine_hubs_and_heavy_weights_vs_oracle(ctx):
    """Rows longer than a wavefront (neighbour updates in several steps) and
    weights large enough that the buckets span millions (the scans skip
    empty stretches by their counts)."""
    rng = np.random.default_rng(8)
    n = 5000
    u = np.concatenate([rng.integers(0, n, 20000), np.full(3000, 11), np.full(900, 4000)])
    v = np.concatenate([rng.integers(0, n, 20000), rng.integers(0, n, 3000), rng.integers(0, n, 900)])
    for wscale in (1, 5000):
        w = rng.integers(1, 9, len(u)) * wscale
        rp, col, ww = graphs.from_edges(n, u, v, w, lcc=True)
        nn = len(rp) - 1
        vw = rng.integers(1, 7, nn).astype(np.uint32)
        for part in (rng.random(nn) < 0.5, rng.random(nn) < 0.2):
            _check_refine(ctx, rp, col, ww, vw, part.astype(np.uint8))


def _multigraph(rng, n, ne, loops):
    """Symmetric CSR that keeps repeated edges and self-loops (both
    directions of every edge stored, rows in edge order)."""
    E = [tuple(x) for x in rng.integers(0, n, (ne, 2))]
    E += E[: ne // 10]  # repeated edges
    E += [(i, i) for i in range(0, n, loops)]
    wts = rng.integers(1, 5, len(E))
    rows = [[] for _ in range(n)]
    for (a, b), x in zip(E, wts):
        rows[a].append((b, x))
        if a != b:
            rows[b].append((a, x))
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c, _ in r], np.uint32)
    w = np.array([x for r in rows for _, x in r], np.uint32)
    return rp, col, w


def test_fm_refine_repeated_columns_and_self_loops_vs_oracle(ctx):
    """A row that names a neighbour twice (or the vertex itself) is taken
    entry by entry, as the reference's loop does."""
    rng = np.random.default_rng(2)
    for n, ne in ((300, 6000), (60, 3000)):
        rp, col, w = _multigraph(rng, n, ne, 7)
        for s in range(3):
            part = (np.random.default_rng(s).random(n) < 0.5).astype(np.uint8)
            _check_refine(ctx, rp, col, w, None, part)


def test_fm_refine_refuses_an_asymmetric_graph(ctx):
    """Gains of a graph whose rows disagree leave the bucket range; the
    search stops with an error instead of indexing outside it."""
    import ipdps_2021_graph_coarsening_amd as mi
    rp = np.array([0, 0, 1, 2, 3, 4], np.uint32)  # 1, 2, 3, 4 -> 0; row 0 empty
    col = np.array([0, 0, 0, 0], np.uint32)
    # maxE = 1; moving vertex 1 to side 0 takes vertex 0's gain to -2
    with pytest.raises(mi.MiError):
        ctx.graph(rp, col).fm_refine(np.array([0, 1, 1, 1, 1], np.uint32))


@pytest.mark.parametrize("seed", range(6))
def test_fm_init_gggp_vs_oracle(ctx, seed):
    rng = np.random.default_rng(50 + seed)
    n = [12, 60, 120, 180, 200, 7][seed]
    rp, col, w = random_graph(rng, n, min(1.0, 6.0 / n), weighted=bool(seed % 2))
    nn = len(rp) - 1
    vw = rng.integers(1, 9, nn).astype(np.uint32) if seed % 3 else None
    got = ctx.graph(rp, col, w).fm_init(vw, seed=1)
    want = oracle.init_gggp(rp, col, w, vw)
    assert np.array_equal(got, want.astype(np.uint32))


def test_fm_init_gggp_two_cliques(ctx):
    rp, col = _two_cliques(7)
    got = ctx.graph(rp, col).fm_init()
    assert _cut2(rp, col, None, got) == 2 and int(got.sum()) == 7


def test_fm_init_spectral(ctx):
    """More than 200 vertices: init_spectral.  Two clusters joined by a few
    edges: the Fiedler vector separates them, so the GPU's power iteration
    (equal to the oracle's to its tolerance) gives the oracle's partition;
    on a plain random graph the sides agree but for vertices whose
    eigenvector entries the tolerance leaves undecided."""
    rng = np.random.default_rng(3)
    k = 160
    u = np.concatenate([rng.integers(0, k, 1500), rng.integers(k, 2 * k, 1500), rng.integers(0, k, 3)])
    v = np.concatenate([rng.integers(0, k, 1500), rng.integers(k, 2 * k, 1500), rng.integers(k, 2 * k, 3)])
    rp, col, w = graphs.from_edges(2 * k, u, v, None, lcc=True)
    nn = len(rp) - 1
    vw = rng.integers(1, 4, nn).astype(np.uint32)
    got = ctx.graph(rp, col, w).fm_init(vw, seed=9)
    want, _ = oracle.init_spectral(rp, col, w, vw, 9)
    assert np.array_equal(got, want.astype(np.uint32))
    rp, col, w = random_graph(rng, 1500, 0.004, weighted=True, lcc=True)
    got = ctx.graph(rp, col, w).fm_init(None, seed=4)
    want, _ = oracle.init_spectral(rp, col, w, None, 4)
    assert np.count_nonzero(got != want) <= len(got) // 50


@pytest.mark.parametrize("name", ["grid64", "random", "weighted"])
def test_fm_eigensolve_vs_oracle(ctx, name):
    rng = np.random.default_rng(12)
    if name == "grid64":
        rp, col = graphs.grid2d(64)[:2]
        w = None
    else:
        rp, col, w = random_graph(rng, 4000, 0.0015, weighted=name == "weighted", lcc=True)
    g = ctx.graph(rp, col, w)
    H = g.hierarchy(21)
    ev, cut, st = H.fm_eigensolve(21)
    O = oracle.Hierarchy(rp, col, w, 21)
    opart, ocut, opasses = O.fm_eigensolve(21)
    assert H.nlevels == O.nlevels
    assert cut == ocut
    assert np.array_equal(ev, opart.astype(np.float64))
    assert list(st["level_iters"][:H.nlevels]) == [int(x) for x in opasses]


@pytest.mark.parametrize("name", ["grid128", "random", "weighted"])
def test_partition_graph_fm_vs_oracle(ctx, name):
    import ipdps_2021_graph_coarsening_amd as mi
    rng = np.random.default_rng(31)
    if name == "grid128":
        rp, col = graphs.grid2d(128)[:2]
        w = None
    else:
        rp, col, w = random_graph(rng, 6000, 0.001, weighted=name == "weighted", lcc=True)
    ctx.set_refinement(True)
    try:
        r = mi.partition_graph(ctx, ctx.graph(rp, col, w), hec_seed=5, pcg_state=77)
    finally:
        ctx.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, w, 5)
    assert r.edge_cut == o["edge_cut"]
    assert np.array_equal(r.part, o["part"])
    assert r.pcg_state == 77  # the FM path draws nothing from the pcg stream
    assert list(r.stats["level_iters"][:o["nlevels"]]) == [int(x) for x in o["passes"]]


def test_bisect_batch_fm_equals_single(ctx):
    import ipdps_2021_graph_coarsening_amd as mi
    rng = np.random.default_rng(41)
    gs = [random_graph(rng, n, 6.0 / n, lcc=True) for n in (800, 1500, 300)]
    ctx.set_refinement(True)
    try:
        G = [ctx.graph(rp, col, w) for rp, col, w in gs]
        res, _ = mi.bisect_batch(ctx, G, hec_seeds=[3, 4, 5], pcg_states=[1, 2, 3])
        for k, (rp, col, w) in enumerate(gs):
            one = mi.partition_graph(ctx, G[k], hec_seed=3 + k, pcg_state=1 + k)
            assert res[k].edge_cut == one.edge_cut and np.array_equal(res[k].part, one.part)
            o = oracle.partition_graph_fm(rp, col, w, 3 + k)
            assert res[k].edge_cut == o["edge_cut"] and np.array_equal(res[k].part, o["part"])
    finally:
        ctx.set_refinement(False)


@pytest.mark.parametrize("name", ["grid96", "rgg"])
def test_partition_4way_fm_vs_oracle(ctx, name):
    """The 4-way driver (sgpar.cpp:53-89) in FM mode against the oracle's FM
    bisections composed the same way (bisect, sgp_use_partition, bisect each
    side with the same seed): bit-exact labels and cuts.  The two sides run
    side by side on the worker contexts."""
    import ipdps_2021_graph_coarsening_amd as mi
    if name == "grid96":
        rp, col = graphs.grid2d(96)[:2]
    else:
        rp, col, _ = graphs.rgg(12000, seed=3)
    ctx.set_refinement(True)
    try:
        r4 = mi.partition_4way(ctx, ctx.graph(rp, col), hec_seed=6, pcg_state=2)
    finally:
        ctx.set_refinement(False)
    o = oracle.partition_graph_fm(rp, col, None, 6)
    (r0, c0), (r1, c1) = graphs.use_partition(rp, col, o["part"])
    o0 = oracle.partition_graph_fm(r0, c0, None, 6)
    o1 = oracle.partition_graph_fm(r1, c1, None, 6)
    assert r4.extra["cuts"] == [o["edge_cut"], o0["edge_cut"], o1["edge_cut"]]
    side = o["part"].astype(bool)
    exp = np.empty(len(rp) - 1, np.uint32)
    exp[~side] = o0["part"]
    exp[side] = 2 + o1["part"]
    assert np.array_equal(r4.part, exp)
