"""Pins of the oracle's GOSH v2 coarsening (the gosh2_* binaries,
-DGOSH_V2: sgp_coarsen_GOSH_v2, heuristics.h:444-601, with
parallel_map_construct_prefilled, :378-442).  The reference ships no
fixture for it, so a second model of the whole level in the order-free form
the GPU computes is checked against the oracle's one-thread walk:

  * seeds (more entries than max(nnz/n, 50)) numbered in index order;
  * each of the two joining sweeps as a fixed point: a vertex is reached when
    a neighbour held an id before the sweep or is a smaller vertex reached
    itself; it takes the id of its candidate of the largest (key, entry)
    -- weight, then number of entries -- following candidates down to one
    that held an id before the sweep;
  * the hand-out sweep as MIS-2's (reach steps relaxed, reachers followed);
  * the heavy neighbours of the remaining vertices from the XorShift64 draws
    at their entry offsets (the pool itself is pinned with HEC);
  * the prefilled map by the closed form of HEC's map with positions equal
    to vertex ids (`mapspec.closed_form_prefilled`), itself checked against
    the oracle's walk on thousands of random instances.
"""
import numpy as np
import pytest

import mapspec
import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import graphs

INF = 0xFFFFFFFF


def _rows(rp, col):
    return [list(map(int, col[rp[i]:rp[i + 1]])) for i in range(len(rp) - 1)]


def _join_sweep(rp, col, vc, key):
    """One joining sweep as a fixed point (see the module docstring)."""
    n = len(rp) - 1
    rows = _rows(rp, col)
    a0 = vc != INF
    reach = np.zeros(n, bool)
    changed = True
    while changed:
        changed = False
        for i in range(n):
            if a0[i] or reach[i]:
                continue
            if any(a0[v] or (v < i and reach[v]) for v in rows[i]):
                reach[i] = True
                changed = True
    p = np.full(n, -1, np.int64)
    for i in np.flatnonzero(reach):
        cands = [(key(e, v), e, v) for e, v in zip(range(rp[i], rp[i + 1]), rows[i]) if a0[v] or (v < i and reach[v])]
        p[i] = max(cands)[2]
    out = vc.copy()
    for i in np.flatnonzero(reach):
        x = i
        while not a0[x]:
            x = p[x]
        out[i] = vc[x]
    return out


def _handout_sweep(rp, col, vc):
    """Mapped vertices hand their id to unmapped neighbours in an index-order
    walk that sees its own writes: v is reached at the smallest neighbour i
    that was mapped before its own turn."""
    n = len(rp) - 1
    rows = _rows(rp, col)
    big = 1 << 62
    T = np.where(vc != INF, -1, big)
    changed = True
    while changed:
        changed = False
        for v in range(n):
            if T[v] == -1:
                continue
            c = [i for i in rows[v] if T[i] < i]
            t = min(c) if c else big
            if t < T[v]:
                T[v] = t
                changed = True
    out = vc.copy()
    for v in range(n):
        if T[v] in (-1, big):
            continue
        x = v
        while T[x] != -1:
            x = T[x]
        out[v] = vc[x]
    return out


def gosh2_model(rp, col, w, seed):
    rp = np.asarray(rp, np.int64)
    n = len(rp) - 1
    deg = np.diff(rp)
    wt = np.ones(int(rp[-1]), np.int64) if w is None else np.asarray(w, np.int64)
    thr = max(int(rp[-1]) // n, 50)
    seedv = deg > thr
    vc = np.full(n, INF, np.int64)
    vc[seedv] = np.arange(int(seedv.sum()))
    nc0 = int(seedv.sum())
    vc = _join_sweep(rp, col, vc, lambda e, v: int(wt[e]))
    vc = _join_sweep(rp, col, vc, lambda e, v: int(deg[v]))
    vc = _handout_sweep(rp, col, vc)
    rem = np.flatnonzero(vc == INF)
    draws, _ = oracle.urand64_stream(oracle.pool_state(seed), int(deg[rem].sum()))
    hn = np.arange(n)
    off = 0
    for i in rem:
        best = None
        for j, e in enumerate(range(rp[i], rp[i + 1])):
            k = (int(wt[e]), int(draws[off + j]), j)
            if best is None or k > best:
                best = k
        off += deg[i]
        if best is None:
            return None, None
        hn[i] = col[rp[i] + best[2]]
    out, nc = mapspec.closed_form_prefilled(vc, hn, nc0)
    return out, nc


def test_prefilled_closed_form_random():
    rng = np.random.default_rng(1)
    for t in range(4000):
        n = int(rng.integers(1, 40))
        hn = rng.integers(0, n, n) if t % 3 else rng.integers(0, max(1, n // 4), n)
        pre = rng.random(n) < rng.random() * 0.6
        nc0 = int(pre.sum())
        vc0 = np.full(n, INF, np.uint32)
        vc0[pre] = rng.integers(0, max(1, nc0), nc0)
        hn = hn.astype(np.uint32)
        hn[pre] = np.flatnonzero(pre)
        ovc, onc = oracle.map_construct_prefilled(vc0, np.flatnonzero(~pre), hn, nc0)
        svc, snc = mapspec.closed_form_prefilled(vc0, hn, nc0)
        assert onc == snc and np.array_equal(ovc, svc), t


def _cases():
    rng = np.random.default_rng(21)
    yield "rmat", graphs.rmat(12, 8, seed=3)
    yield "rmat_dense", graphs.rmat(11, 64, seed=5)
    yield "rgg", graphs.rgg(3000, seed=2)
    rp, col, w = random_graph(rng, 1200, 0.006, weighted=True, lcc=True)
    yield "er_w", (rp, col, w)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_gosh2_against_order_free_model(case):
    name, (rp, col, w) = case
    for seed in (4, 31):
        vc, nc = oracle.coarsen_gosh2(rp, col, w, seed)
        mv, mnc = gosh2_model(rp, col, w, seed)
        assert nc == mnc and np.array_equal(vc, mv)
        assert (vc < nc).all()


def test_gosh2_hierarchy():
    rp, col, w = graphs.rmat(12, 16, seed=7)
    H = oracle.Hierarchy(rp, col, w, 9, variant=7)
    assert H.nlevels >= 2
    for l in range(H.nlevels - 1):
        a, b = H.levels[l], H.levels[l + 1]
        vc, nc = oracle.coarsen_gosh2(a["rowptr"], a["col"], a["w"], 9)
        assert nc == b["n"] and np.array_equal(vc, a["vcmap"])
