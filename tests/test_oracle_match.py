"""Pins of the oracle's heavy-edge matching (the match_* binaries,
-DPUREMATCH: sgp_coarsen_match, heuristics.h:1584-1735, 1890-1895).  The
reference ships no fixture for it, so:

  * a second model of each round that does not walk the list: the pairs of
    a round are the lexicographically first maximal matching of its active
    edges (u, hn(u)) in list order, found by rounds of local minima (an
    undecided edge that is the earliest undecided edge at both its ends is
    taken, edges touching a taken vertex are dropped); that is the order-free
    form the GPU computes, checked against the sequential walk;
  * the result is a matching: every coarse vertex holds one or two
    vertices, pairs are edges of the graph, pair ids come in round order
    before the singletons, which are numbered in index order;
  * it is maximal: no two singletons are neighbours;
  * the random and first-heaviest neighbours come from HEC's pinned pool,
    permutation and draws.
"""
import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import graphs


def _rows(rp, col):
    return [list(map(int, col[rp[i]:rp[i + 1]])) for i in range(len(rp) - 1)]


def lfmm(edges):
    """Lexicographically first maximal matching of edges (in order), by
    rounds of local minima; returns the indices taken."""
    und = [k for k, (u, v) in enumerate(edges) if u != v]
    taken, matched = [], set()
    while und:
        first = {}
        for k in und:
            for x in edges[k]:
                first.setdefault(x, k)
                first[x] = min(first[x], k)
        now = [k for k in und if all(first[x] == k for x in edges[k])]
        taken += now
        for k in now:
            matched.update(edges[k])
        und = [k for k in und if k not in now and not (set(edges[k]) & matched)]
    return sorted(taken)


def match_model(level, rp, col, w, seed):
    rows = _rows(rp, col)
    n = len(rows)
    wt = np.ones(int(rp[-1]), np.int64) if w is None else np.asarray(w, np.int64)
    vperm, _, st = oracle.generate_permutation(n, seed)
    pos = np.empty(n, np.int64)
    pos[vperm] = np.arange(n)
    if level == 1:
        dr, _ = oracle.urand64_stream(st, n)
        hn = [rows[i][int(dr[i]) % len(rows[i])] for i in range(n)]
    else:
        hn = []
        for i in range(n):
            ws = wt[rp[i]:rp[i + 1]]
            hn.append(rows[i][int(np.argmax(ws))])  # argmax: the first of the largest
    vc = [-1] * n
    nc, lst, swap = 0, [int(x) for x in vperm], 1
    while lst:
        act = [u for u in lst if (int(pos[u] < pos[hn[u]]) ^ swap) and vc[hn[u]] < 0]
        edges = [(u, hn[u]) for u in act]
        for k in lfmm(edges):
            u, v = edges[k]
            vc[u] = vc[v] = nc
            nc += 1
        nxt = []
        for u in lst:
            if vc[u] >= 0:
                continue
            cand = [(int(pos[v]) if level == 1 else int(wt[e]), e, v)
                    for e, v in zip(range(rp[u], rp[u + 1]), rows[u]) if vc[v] < 0]
            if cand:
                hn[u] = max(cand)[2]  # the largest key, the last of equals
                nxt.append(u)
        lst, swap = nxt, swap ^ 1
    for i in range(n):
        if vc[i] < 0:
            vc[i] = nc
            nc += 1
    return np.array(vc, np.uint32), nc


def _cases():
    rng = np.random.default_rng(12)
    yield "grid", graphs.grid2d(20)
    yield "rgg", graphs.rgg(1500, seed=2)
    yield "rmat_w", graphs.rmat(9, 8, seed=4)
    rp, col, w = random_graph(rng, 700, 0.008, weighted=True, lcc=True)
    yield "er_w", (rp, col, w)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
@pytest.mark.parametrize("level", [1, 2])
def test_match_against_order_free_model(case, level):
    name, (rp, col, w) = case
    for seed in (3, 41):
        vc, nc, rounds = oracle.coarsen_match(level, rp, col, w, seed)
        mv, mnc = match_model(level, rp, col, w, seed)
        assert nc == mnc and np.array_equal(vc, mv)
        assert rounds >= 1


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_match_is_a_maximal_matching(case):
    name, (rp, col, w) = case
    rows = _rows(rp, col)
    n = len(rows)
    vc, nc, _ = oracle.coarsen_match(2, rp, col, w, 7)
    members = [[] for _ in range(nc)]
    for i in range(n):
        members[vc[i]].append(i)
    sizes = np.array([len(x) for x in members])
    assert sizes.min() >= 1 and sizes.max() <= 2
    npairs = int((sizes == 2).sum())
    assert (sizes[:npairs] == 2).all() and (sizes[npairs:] == 1).all()
    for a, b in (x for x in members if len(x) == 2):
        assert b in rows[a]
    single = [x[0] for x in members[npairs:]]
    assert single == sorted(single)
    sset = set(single)
    assert not any(v in sset for u in single for v in rows[u] if v != u)
    assert npairs > n // 4


def test_match_refuses_what_the_reference_cannot_map():
    rp = np.array([0, 1, 2, 2], np.uint32)  # vertex 2 has no entries
    col = np.array([1, 0], np.uint32)
    assert oracle.coarsen_match(1, rp, col, None, 1)[1] is None
    # only a self-loop left: vertex 2's sole neighbour is itself
    rp = np.array([0, 1, 2, 3], np.uint32)
    col = np.array([1, 0, 2], np.uint32)
    assert oracle.coarsen_match(2, rp, col, None, 1)[1] is None


def test_match_hierarchy():
    rp, col, w = graphs.rgg(6000, seed=5)
    H = oracle.Hierarchy(rp, col, w, 9, variant=1)
    assert H.nlevels >= 4
    for l in range(H.nlevels - 1):
        a, b = H.levels[l], H.levels[l + 1]
        vc, nc, _ = oracle.coarsen_match(l + 1, a["rowptr"], a["col"], a["w"], 9)
        assert nc == b["n"] and np.array_equal(vc, a["vcmap"])
