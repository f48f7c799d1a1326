"""The device input path (mi_graph_from_edges) against the host converter
graphs.from_edges, which restates the reference's mtx2csr.cpp:116-289
(symmetrise, drop self-loops, merge duplicates, sort rows, keep the largest
connected component in vertex order): the same rowptr, columns and merged
weights, bit for bit.  Cases: duplicates in both directions, self-loops,
isolated vertices, several components including a tie for the largest
(the reference keeps the one whose first vertex comes first), a graph with
no edges, weights summed over duplicates, unit weights without w, and a
chain long enough that the union-find needs many hooks."""
import numpy as np
import pytest

from ipdps_2021_graph_coarsening_amd import graphs

pytestmark = pytest.mark.gpu


def _check(ctx, n, u, v, w, lcc):
    rp, col, ww = graphs.from_edges(n, u, v, w, lcc=lcc)
    g = ctx.graph_from_edges(n, u, v, w, lcc=lcc)
    r2, c2, w2 = g.to_host()
    g.close()
    assert np.array_equal(r2, rp)
    assert np.array_equal(c2, col)
    if w is not None:
        assert np.array_equal(w2, ww)
    else:
        assert np.all(w2 == 1)  # unit weights (the host converter counts multiplicities)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("lcc", [True, False])
@pytest.mark.parametrize("weighted", [True, False])
def test_from_edges_random(ctx, seed, lcc, weighted):
    rng = np.random.default_rng(seed)
    n = 5000 + 1000 * seed
    ne = 9000
    u = rng.integers(0, n, ne)
    v = rng.integers(0, n, ne)
    u[:200] = v[:200]                       # self-loops
    u[200:400], v[200:400] = v[400:600], u[400:600]  # duplicates in the other direction
    w = rng.integers(1, 1000, ne) if weighted else None
    _check(ctx, n, u, v, w, lcc)


def test_from_edges_tie_for_largest_component(ctx):
    # three paths of 50 vertices (two of them tied as largest) and isolated
    # vertices; the first tied component (lowest first vertex) is kept
    n = 400
    u, v = [], []
    for start in (300, 10, 150):            # components at 10.., 150.., 300..
        for k in range(49):
            u.append(start + k)
            v.append(start + k + 1)
    u = np.array(u[:-5])                    # the component at 150 is shorter
    v = np.array(v[:-5])
    rp, col, _ = graphs.from_edges(n, u, v, None, lcc=True)
    assert len(rp) - 1 == 50
    _check(ctx, n, u, v, None, True)
    g = ctx.graph_from_edges(n, u, v, None, lcc=True)
    r2, c2, _ = g.to_host()
    g.close()
    assert c2.min() == 0 and r2[-1] == 98


def test_from_edges_no_edges_and_chain(ctx):
    _check(ctx, 7, np.zeros(0, np.int64), np.zeros(0, np.int64), None, True)
    # a long chain given in reverse order: the union-find hooks along it
    n = 200000
    u = np.arange(n - 1, 0, -1)
    v = u - 1
    _check(ctx, n, u, v, np.ones(n - 1, np.int64), True)


def test_from_edges_rmat(ctx):
    rng = np.random.default_rng(7)
    n = 1 << 14
    # R-MAT-like edges: hubs, duplicates, small components
    u = rng.integers(0, n, 120000) & rng.integers(0, n, 120000)
    v = rng.integers(0, n, 120000) & rng.integers(0, n, 120000)
    _check(ctx, n, u, v, rng.integers(1, 9, len(u)), True)
    _check(ctx, n, u, v, None, True)
