"""The bucketed coarse build (P^T A P with duplicate merge: rows grouped by
coarse vertex, small rows sorted in LDS straight from their members, large
rows split into 4096-column windows, small windows sorted in LDS, big
windows summed in dense LDS accumulators) against the oracle's build, bit
for bit (row offsets, columns, merged weights, cluster weights), with the
bucketed form forced on every level (mi_ctx_set_bucketed_build(ctx, 0)).

Cases pick out each branch and each edge of it:
- HEC maps of weighted R-MAT graphs (hub rows, many windows);
- random maps with skewed cluster sizes onto 5,000 and 20,000 coarse
  vertices (1 and 5 windows per large row);
- stars whose centre row holds 2047, 2048 and 2049 entries (the small /
  large boundary), one whose centre row's first window holds 70,000 entries
  (a big window summed by two workgroups) and one at 4096 + 1 coarse
  vertices (a second window of one column);
- rows whose every entry is a self-loop (nothing left), vertices without
  edges inside clusters, coarse vertices without members, weights up to
  2^20 merged many times over.
"""
import numpy as np
import pytest

import oracle
from ipdps_2021_graph_coarsening_amd import graphs
from conftest import random_graph

pytestmark = pytest.mark.gpu


def _check(ctx, rp, col, w, vcmap, nc):
    vcmap = np.ascontiguousarray(vcmap, dtype=np.uint32)
    orr, oc, ow, ocvw = oracle.build_coarse(rp, col, w, vcmap, nc)
    ctx.set_bucketed_build(0)
    try:
        g = ctx.graph(rp, col, w)
        gc, cvw = g.build_coarse(vcmap, nc)
        r, c, ww = gc.to_host()
        gc.close()
        g.close()
    finally:
        ctx.set_bucketed_build()
    assert np.array_equal(r, orr)
    assert np.array_equal(c, oc)
    assert np.array_equal(ww, ow)
    assert np.array_equal(cvw, ocvw)
    return orr


@pytest.mark.parametrize("scale,seed", [(13, 1), (15, 2)])
def test_bucketed_build_hec_maps_rmat(ctx, scale, seed):
    rp, col, w = graphs.rmat(scale, 16, seed=seed)
    for level in (1, 2):
        vc, nc, _, _ = oracle.coarsen_hec(level, rp, col, w, seed)
        orr = _check(ctx, rp, col, w, vc, nc)
        assert np.diff(orr).max() > 0


@pytest.mark.parametrize("nc", [5000, 20000])
def test_bucketed_build_skewed_random_maps(ctx, nc):
    rng = np.random.default_rng(nc)
    rp, col, w = graphs.rmat(15, 24, seed=3)
    n = len(rp) - 1
    # a few huge clusters, many small ones, some coarse vertices left empty
    big = rng.integers(0, 40, n)
    vc = np.where(rng.random(n) < 0.3, big, rng.integers(0, nc - 10, n)).astype(np.uint32)
    # weights up to 2^20, symmetric: a hash of the unordered pair
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(rp))
    lo, hi = np.minimum(src, col), np.maximum(src, col)
    w = ((lo * np.uint64(2654435761) + hi * np.uint64(40503)) % np.uint64(1 << 20) + np.uint64(1)).astype(np.uint32)
    _check(ctx, rp, col, w, vc, nc)


def _star(leaves, seed, leaf_clusters, nc, center_cluster=0):
    """A star (centre 0) plus a ring over the leaves; leaves mapped onto
    leaf_clusters (values in [1, nc)), the centre alone in its cluster;
    vertices without edges after the leaves, spread over [1, nc), so that
    there are at least nc vertices."""
    rng = np.random.default_rng(seed)
    k = leaves + 1
    n = max(k, nc)
    u = np.concatenate([np.zeros(leaves, np.int64), np.arange(1, k)])
    v = np.concatenate([np.arange(1, k), np.roll(np.arange(1, k), 1)])
    wt = rng.integers(1, 1000, len(u))
    rp, col, w = graphs.from_edges(n, u, v, wt, lcc=False)
    vc = np.empty(n, np.uint32)
    vc[0] = center_cluster
    vc[1:k] = leaf_clusters
    vc[k:] = rng.integers(1, nc, n - k)
    return rp, col, w, vc


@pytest.mark.parametrize("leaves", [2047, 2048, 2049])
def test_bucketed_build_small_large_boundary(ctx, leaves):
    # centre row: `leaves` entries plus none from its own cluster; its
    # coarse row holds exactly `leaves` fine entries (2047..2049 around the
    # 2048 boundary; 2049 leaves -> 2049 entries, a large row)
    nc = 9000
    rng = np.random.default_rng(leaves)
    lc = rng.integers(1, nc, leaves)
    rp, col, w, vc = _star(leaves, leaves, lc, nc)
    _check(ctx, rp, col, w, vc, nc)


def test_bucketed_build_big_window_two_workgroups(ctx):
    # 70,000 leaves all mapped into the first 4096 coarse vertices: the
    # centre's coarse row has one window of 70,000 entries (> 65,536: two
    # accumulating workgroups), many merged duplicates
    leaves, nc = 70000, 30000
    rng = np.random.default_rng(5)
    lc = rng.integers(1, 4096, leaves)
    rp, col, w, vc = _star(leaves, 5, lc, nc)
    _check(ctx, rp, col, w, vc, nc)


def test_bucketed_build_window_of_one_column(ctx):
    # nc = 4097: the second window of a large row holds the single column 4096
    leaves, nc = 6000, 4097
    lc = np.random.default_rng(6).integers(1, nc, leaves)
    lc[:50] = 4096
    rp, col, w, vc = _star(leaves, 6, lc, nc)
    _check(ctx, rp, col, w, vc, nc)


def test_bucketed_build_self_loops_isolated_and_empty(ctx):
    rng = np.random.default_rng(11)
    n = 30000
    # a clique of 70 vertices (4830 entries, every one a self-loop once the
    # clique is one cluster), a random graph on the rest, 500 isolated
    # vertices
    cl = np.arange(70)
    cu, cv = np.meshgrid(cl, cl)
    keep = cu < cv
    ru = rng.integers(100, n - 500, 60000)
    rv = rng.integers(100, n - 500, 60000)
    u = np.concatenate([cu[keep], ru])
    v = np.concatenate([cv[keep], rv])
    rp, col, w = graphs.from_edges(n, u, v, rng.integers(1, 50, len(u)), lcc=False)
    nc = 12000
    vc = rng.integers(1, nc - 100, n).astype(np.uint32)  # coarse vertices >= nc - 100 stay empty
    vc[:70] = 0                                           # the clique: one cluster of self-loops
    vc[n - 500:] = rng.integers(0, 3, 500)                # isolated vertices joining clusters
    _check(ctx, rp, col, w, vc, nc)


@pytest.mark.parametrize("seed", [0, 1])
def test_bucketed_build_random_graphs(ctx, seed):
    rng = np.random.default_rng(seed)
    rp, col, w = random_graph(rng, 40000, 0.0015, weighted=True)
    vc, nc, _, _ = oracle.coarsen_hec(2, rp, col, w, seed)
    _check(ctx, rp, col, w, vc, nc)
