"""Pins of the oracle's MIS-2 coarsening (the mis_* binaries, -DMIS:
mis_2, heuristics.h:8-100; sgp_coarsen_mis_2, :218-289).  The reference
ships no fixture for it, so:

  * the set is a maximal distance-2 independent set: no two members within
    two hops, every other vertex within two hops of one;
  * a second model of each round that takes the maximum (state, random,
    index) over each vertex's explicit two-hop ball (a breadth-first search)
    instead of two rounds of neighbour maxima over the tuple arrays;
  * the aggregation sweep against a second model that solves its fixed
    point (the step at which each vertex is reached: the smallest index of
    a neighbour reached before its own step) by repeated relaxation and
    follows each vertex's reacher to its set vertex;
  * known answers: one aggregate for a star and a clique.
"""
import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import graphs


def _adj(rp, col):
    return [list(col[rp[i]:rp[i + 1]]) for i in range(len(rp) - 1)]


def _ball2(adj, i):
    b = {i}
    for v in adj[i]:
        b.add(v)
        b.update(adj[v])
    return b


def mis2_model(rp, col, seed):
    adj = _adj(rp, col)
    n = len(adj)
    st = oracle.pool_state(seed)
    rnd, _ = oracle.urand64_stream(st, n)
    rnd = [int(x) for x in rnd]
    state = [0] * n
    balls = [_ball2(adj, i) for i in range(n)]
    while any(s == 0 for s in state):
        best = []
        for i in range(n):
            best.append(max(balls[i], key=lambda v: (state[v], rnd[v], v)))
        new = list(state)
        for i in range(n):
            if state[i] == 0:
                b = best[i]
                if b == i:
                    new[i] = 1
                elif state[b] == 1 or best[b] == b:
                    new[i] = -1
        state = new
    return np.array(state, np.int32)


def agg_model(rp, col, state):
    adj = _adj(rp, col)
    n = len(adj)
    INF = 1 << 62
    T = [INF] * n
    ids = {}
    for i in range(n):
        if state[i] == 1:
            ids[i] = len(ids)
    own = {}
    for i in range(n):
        if state[i] == 1:
            T[i] = -1
            own[i] = ids[i]
            for v in adj[i]:
                T[v] = -1
                own[v] = ids[i]
    changed = True
    while changed:
        changed = False
        for v in range(n):
            if v in own:
                continue
            c = [i for i in adj[v] if T[i] < i]
            t = min(c) if c else INF
            if t < T[v]:
                T[v] = t
                changed = True
    src = {v: min(i for i in adj[v] if T[i] < i) for v in range(n) if v not in own}

    def agg(v):
        while v not in own:
            v = src[v]
        return own[v]

    return np.array([agg(v) for v in range(n)], np.uint32), len(ids)


def _cases():
    rng = np.random.default_rng(3)
    yield "grid", graphs.grid2d(24)[:2]
    yield "rgg", graphs.rgg(1500, seed=1)[:2]
    yield "rmat", graphs.rmat(9, 8, seed=2)[:2]
    rp, col, _ = random_graph(rng, 600, 0.01, lcc=False)
    yield "er", (rp, col)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_mis2_rounds_and_set(case):
    name, (rp, col) = case
    for seed in (1, 77):
        st, rounds = oracle.mis2(rp, col, seed)
        assert np.array_equal(st, mis2_model(rp, col, seed))
        adj = _adj(rp, col)
        members = set(np.flatnonzero(st == 1).tolist())
        for i in members:
            assert not (_ball2(adj, i) - {i}) & members
        for i in range(len(adj)):
            assert st[i] in (1, -1)
            if st[i] == -1:
                assert _ball2(adj, i) & members


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_mis2_aggregation_fixed_point(case):
    name, (rp, col) = case
    vc, nc = oracle.coarsen_mis2(rp, col, 5)
    st, _ = oracle.mis2(rp, col, 5)
    mv, mnc = agg_model(rp, col, st)
    assert nc == mnc and np.array_equal(vc, mv)


def test_mis2_star_and_clique_one_aggregate():
    n = 30
    E = [(0, i) for i in range(1, n)]
    adj = [[] for _ in range(n)]
    for a, b in E:
        adj[a].append(b)
        adj[b].append(a)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(x) for x in adj])
    col = np.array([v for x in adj for v in sorted(x)], np.uint32)
    vc, nc = oracle.coarsen_mis2(rp, col, 9)
    assert nc == 1 and not vc.any()
    k = 12
    adj = [[j for j in range(k) if j != i] for i in range(k)]
    rp = np.zeros(k + 1, np.uint32)
    rp[1:] = np.cumsum([len(x) for x in adj])
    col = np.array([v for x in adj for v in x], np.uint32)
    vc, nc = oracle.coarsen_mis2(rp, col, 3)
    assert nc == 1 and not vc.any()


def test_mis2_hierarchy():
    rp, col, w = graphs.rgg(5000, seed=2)
    H = oracle.Hierarchy(rp, col, w, 4, variant=5)
    assert H.nlevels >= 2
    for l in range(H.nlevels - 1):
        assert len(np.unique(H.levels[l]["vcmap"])) == H.levels[l + 1]["n"]
