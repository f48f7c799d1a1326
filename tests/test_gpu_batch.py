"""mi_bisect_batch: the bisections of many graphs, with the power iterations
of their small levels run together (one workgroup per graph and level),
equal graph by graph and bit for bit the one-graph calls
(mi_partition_graph): partitions, cuts, iteration counts, the continuing
PCG states and the final eigenvalue logs."""
import numpy as np
import pytest

import ipdps_2021_graph_coarsening_amd as mi
from ipdps_2021_graph_coarsening_amd import graphs
from conftest import random_graph

pytestmark = pytest.mark.gpu


def _graphs():
    rng = np.random.default_rng(8)
    out = [graphs.rgg(n, seed=30 + i) for i, n in enumerate((1000, 2500, 4000, 9000))]
    out += [graphs.grid2d(64), graphs.grid2d(128), graphs.rmat(11, 16, seed=4), graphs.rmat(13, 16, seed=1),
            random_graph(rng, 3000, 0.002, weighted=True, lcc=True), graphs.path(40), graphs.grid3d(12)]
    return out


@pytest.mark.parametrize("normLap", [0, 1])
def test_bisect_batch_equals_single_calls(ctx, normLap):
    G = [ctx.graph(rp, col, w) for rp, col, w in _graphs()]
    cfg = mi.Config(4, normLap, 0, 1, 1e-10, 2)
    seeds = [3 + k for k in range(len(G))]
    states = [100 + 7 * k for k in range(len(G))]
    single = [mi.partition_graph(ctx, g, cfg, hec_seed=seeds[k], pcg_state=states[k]) for k, g in enumerate(G)]
    res, ph = mi.bisect_batch(ctx, G, cfg, hec_seeds=seeds, pcg_states=states)
    assert len(res) == len(G) and all(t >= 0 for t in ph)
    for k, (a, b) in enumerate(zip(res, single)):
        assert np.array_equal(a.part, b.part), k
        assert a.edge_cut == b.edge_cut and a.pcg_state == b.pcg_state, k
        sa, sb = a.stats, b.stats
        assert sa["level_n"] == sb["level_n"] and sa["level_iters"] == sb["level_iters"], k
        if normLap == 0:
            assert sa["eigen_log"] == sb["eigen_log"], k
    for g in G:
        g.close()


def test_bisect_batch_resumed_launches_and_empty(ctx):
    """Launches of 5 iterations give the same batch; an empty batch is fine."""
    G = [ctx.graph(rp, col, w) for rp, col, w in _graphs()[:6]]
    seeds, states = [1] * len(G), [2] * len(G)
    ref, _ = mi.bisect_batch(ctx, G, hec_seeds=seeds, pcg_states=states)
    try:
        ctx.set_small_levels(1 << 17, 5)
        res, _ = mi.bisect_batch(ctx, G, hec_seeds=seeds, pcg_states=states)
    finally:
        ctx.set_small_levels()
    for a, b in zip(res, ref):
        assert np.array_equal(a.part, b.part) and a.stats["level_iters"] == b.stats["level_iters"]
    empty, _ = mi.bisect_batch(ctx, [])
    assert empty == []
    for g in G:
        g.close()


def test_bisect_batch_many_graphs_repeated(ctx):
    """96 RGGs of 2,000 to 60,000 vertices, three batches in a row: every
    graph's vector is solved on one worker context and usually split on
    another, with the workers allocating meanwhile; each batch equals the
    one-graph calls in partition and cut."""
    rng = np.random.default_rng(12)
    ns = np.exp(rng.uniform(np.log(2e3), np.log(6e4), 96)).astype(int)
    G = [ctx.graph(*graphs.rgg(int(n), seed=500 + i)[:2]) for i, n in enumerate(ns)]
    seeds, states = [1] * len(G), [1] * len(G)
    single = [mi.partition_graph(ctx, g, hec_seed=1, pcg_state=1) for g in G]
    for rep in range(3):
        res, _ = mi.bisect_batch(ctx, G, hec_seeds=seeds, pcg_states=states)
        for k, (a, b) in enumerate(zip(res, single)):
            assert a.edge_cut == b.edge_cut and np.array_equal(a.part, b.part), (rep, k)
    for g in G:
        g.close()
