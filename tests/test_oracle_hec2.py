"""Pins of the oracle's HEC v2 map (the hec2_* binaries, -DHEC_V2:
parallel_map_construct_v2, heuristics.h:695-764, on one thread).  The
reference ships no fixture for it, so:

  * a second model of each round in the order-free form the GPU computes,
    checked against the oracle's walk on thousands of random (permutation,
    heavy-neighbour) functions and on HEC's own heavy neighbours of graphs:
    - the slot pos(hn(u)) goes to the list vertex of the smallest list
      index that names it;
    - the slots' pseudo-locks: a lock is touched by at most two slots (the
      slot of the same number, and the slot whose sample is the vertex of
      that number), so a slot takes its first lock when no earlier toucher
      took it -- an earlier slot touches a lock that is its first always,
      one that is its second only when it took its first -- and it pairs
      when it takes both and they differ;
    - a vertex written by several pairs keeps the last (largest) id;
    - an unmapped list vertex joins hn(u)'s id after the pairs, or, when
      hn(u) is an unmapped list vertex earlier in the list, whatever that
      one joins (a chase along hn through earlier list positions);
  * rounds map at least one vertex on these inputs, and every vertex ends
    mapped to an id below nc (ids may be left without members: a later pair
    can re-map both members of an earlier one, as the reference does).
"""
import numpy as np
import pytest

import oracle
from conftest import random_graph
from ipdps_2021_graph_coarsening_amd import graphs

INF = 0xFFFFFFFF


def v2_model(vperm, hn):
    n = len(vperm)
    pos = np.empty(n, np.int64)
    pos[vperm] = np.arange(n)
    vc = [INF] * n
    lst = [int(x) for x in vperm]
    nc = 0
    while lst:
        # slots: the first claim in list order
        smp = {}
        for k, u in enumerate(lst):
            smp.setdefault(int(pos[hn[u]]), u)
        slot_of = {u: v for v, u in smp.items()}
        first = {v: min(u, v) for v, u in smp.items()}
        second = {v: max(u, v) for v, u in smp.items()}

        def touchers(L):
            return [t for t in {L, slot_of.get(L)} if t is not None and t in smp]

        A = {}
        for v in sorted(smp):  # an earlier toucher of first(v) decides A(v)
            L = first[v]
            A[v] = not any(t < v and (first[t] == L or (second[t] == L and A[t])) for t in touchers(L))
        ok = []
        for v in sorted(smp):
            L2 = second[v]
            if A[v] and first[v] != L2 and not any(
                    t < v and (first[t] == L2 or (second[t] == L2 and A[t])) for t in touchers(L2)):
                ok.append(v)
        write = {}
        for r, v in enumerate(ok):
            for x in (smp[v], int(vperm[v])):
                write[x] = max(write.get(x, -1), nc + r)
        for x, c in write.items():
            vc[x] = c
        nc += len(ok)
        # joins, chasing earlier unmapped list vertices
        lpos = {u: k for k, u in enumerate(lst)}
        join = {}
        for u in lst:
            if vc[u] != INF:
                continue
            x = u
            while True:
                t = int(hn[x])
                if vc[t] != INF:
                    join[u] = vc[t]
                    break
                if t in lpos and lpos[t] < lpos[x]:
                    x = t
                    continue
                break
        for u, c in join.items():
            vc[u] = c
        nxt = [u for u in lst if vc[u] == INF]
        if len(nxt) == len(lst):
            return np.array(vc, np.uint32), None
        lst = nxt
    return np.array(vc, np.uint32), nc


def test_v2_random_functions():
    rng = np.random.default_rng(7)
    for trial in range(3000):
        n = int(rng.integers(1, 40))
        vperm = rng.permutation(n).astype(np.uint32)
        if trial % 3 == 0:
            hn = rng.integers(0, n, n).astype(np.uint32)
        elif trial % 3 == 1:  # mutual pairs and hubs
            hn = rng.integers(0, max(1, n // 5), n).astype(np.uint32)
        else:
            hn = np.roll(np.arange(n), 1).astype(np.uint32)
        vc, nc, _ = oracle.map_construct_v2(vperm, hn)
        mv, mnc = v2_model(vperm, hn)
        assert nc == mnc, trial
        if nc is not None:
            assert np.array_equal(vc, mv), trial
            assert (vc < nc).all()


def _cases():
    rng = np.random.default_rng(11)
    yield "grid", graphs.grid2d(24)
    yield "rgg", graphs.rgg(2000, seed=4)
    yield "rmat_w", graphs.rmat(10, 8, seed=6)
    rp, col, w = random_graph(rng, 800, 0.006, weighted=True)  # isolated vertices
    yield "er_w", (rp, col, w)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
@pytest.mark.parametrize("level", [1, 2])
def test_v2_levels(case, level):
    name, (rp, col, w) = case
    for seed in (3, 99):
        vc, nc, rounds = oracle.coarsen_hec2(level, rp, col, w, seed)
        _, _, vperm, hn = oracle.coarsen_hec(level, rp, col, w, seed)
        mv, mnc = v2_model(vperm, hn)
        assert nc == mnc and np.array_equal(vc, mv)
        assert nc is not None and (vc < nc).all() and rounds >= 1


def test_v2_hierarchy():
    rp, col, w = graphs.rgg(6000, seed=5)
    H = oracle.Hierarchy(rp, col, w, 9, variant=4)
    assert H.nlevels >= 3
    for l in range(H.nlevels - 1):
        a, b = H.levels[l], H.levels[l + 1]
        vc, nc, _ = oracle.coarsen_hec2(l + 1, a["rowptr"], a["col"], a["w"], 9)
        assert nc == b["n"] and np.array_equal(vc, a["vcmap"])
