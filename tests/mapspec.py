"""Executable specification of the data-parallel map construction that the
GPU kernels (mi_hec.hip) evaluate: the closed form of the reference's
one-thread parallel_map_construct (heuristics.h:633-693).

Pass 1 acts on E = {u : pos(u) >= pos(hn(u))}.  Creators are the self-loops
of E and, for every v outside E, the lowest-position member of E pointing at
v; ids go to creators in position order.  Other members of E join the
cluster of hn(u) (resolved by walking hn inside E, where positions strictly
decrease).  Pass 2 acts on the still-unassigned set R: u acts iff none of
its unassigned children (w with hn(w) = u) acts; an acting u joins hn(u)
when hn(u) was assigned in pass 1, otherwise the lowest-position acting
child of hn(u) creates the pair and the others join it.
"""
import numpy as np

INF = 0xFFFFFFFF


def closed_form_map(vperm, hn):
    vperm = np.asarray(vperm, dtype=np.int64)
    hn = np.asarray(hn, dtype=np.int64)
    n = len(vperm)
    ar = np.arange(n)
    pos = np.empty(n, np.int64)
    pos[vperm] = ar
    inE = pos >= pos[hn]
    minpos = np.full(n, INF, np.int64)
    w = np.where(inE & (hn != ar))[0]
    np.minimum.at(minpos, hn[w], pos[w])
    creator = inE & ((hn == ar) | ((~inE[hn]) & (minpos[hn] == pos)))
    cf = creator[vperm].astype(np.int64)
    cid = np.full(n, INF, np.int64)
    cid[vperm] = np.cumsum(cf) - cf
    nc1 = int(cf.sum())
    vc = np.full(n, INF, np.int64)
    cr = np.where(creator)[0]
    vc[cr] = cid[cr]
    vc[hn[cr]] = cid[cr]
    for u in np.where(inE & ~creator)[0]:
        x = u
        while (not creator[x]) and inE[hn[x]]:
            x = hn[x]
        vc[u] = cid[x] if creator[x] else cid[vperm[minpos[hn[x]]]]
    R = vc == INF
    state = np.zeros(n, np.int64)  # 0 unknown, 1 acts, 2 claimed
    rw = np.where(R)[0]
    rpar = rw[R[hn[rw]]]
    indeg = np.zeros(n, np.int64)
    np.add.at(indeg, hn[rpar], 1)
    rounds = 0
    while True:
        rounds += 1
        has_act = np.zeros(n, bool)
        nonact = np.zeros(n, np.int64)
        a = rpar[state[rpar] == 1]
        has_act[hn[a]] = True
        b = rpar[state[rpar] == 2]
        np.add.at(nonact, hn[b], 1)
        new = state.copy()
        und = R & (state == 0)
        new[und & has_act] = 2
        new[und & ~has_act & (nonact == indeg)] = 1
        if (new == state).all():
            break
        state = new
    act = R & (state == 1)
    minpos2 = np.full(n, INF, np.int64)
    a = np.where(act & R[hn])[0]
    np.minimum.at(minpos2, hn[a], pos[a])
    cr2 = act & R[hn] & (minpos2[hn] == pos)
    cf2 = cr2[vperm].astype(np.int64)
    cid2 = np.full(n, INF, np.int64)
    cid2[vperm] = np.cumsum(cf2) - cf2 + nc1
    out = vc.copy()
    j = np.where(act & ~R[hn])[0]
    out[j] = vc[hn[j]]
    j = np.where(act & R[hn])[0]
    out[j] = cid2[vperm[minpos2[hn[j]]]]
    j = np.where(R & ~act)[0]
    out[j] = cid2[vperm[minpos2[j]]]
    return out.astype(np.uint32), nc1 + int(cf2.sum()), rounds


def closed_form_prefilled(vcmap0, hn, nc0):
    """The same closed form for parallel_map_construct_prefilled
    (heuristics.h:378-442): positions are vertex ids, the walkers are the
    vertices without a prefilled id (in id order), a prefilled vertex never
    walks and a walker whose heavy neighbour holds a prefilled id joins it,
    in pass 1 through the hn-walk inside E and in pass 2 as any assigned
    target; new ids continue from nc0."""
    hn = np.asarray(hn, dtype=np.int64)
    n = len(hn)
    ar = np.arange(n)
    vc0 = np.asarray(vcmap0, dtype=np.int64)
    pre = vc0 != INF
    inE = (~pre) & (ar >= hn)
    minpos = np.full(n, INF, np.int64)
    w = np.where(inE & (hn != ar))[0]
    np.minimum.at(minpos, hn[w], w)
    creator = inE & ((hn == ar) | ((~inE[hn]) & (~pre[hn]) & (minpos[hn] == ar)))
    cf = creator.astype(np.int64)
    cid = np.cumsum(cf) - cf + nc0
    nc1 = int(cf.sum())
    vc = vc0.copy()
    cr = np.where(creator)[0]
    vc[cr] = cid[cr]
    vc[hn[cr]] = cid[cr]
    for u in np.where(inE & ~creator)[0]:
        x = u
        while (not creator[x]) and inE[hn[x]]:
            x = hn[x]
        if creator[x]:
            vc[u] = cid[x]
        elif pre[hn[x]]:
            vc[u] = vc0[hn[x]]
        else:
            vc[u] = cid[minpos[hn[x]]]
    R = vc == INF
    state = np.zeros(n, np.int64)
    rw = np.where(R)[0]
    rpar = rw[R[hn[rw]]]
    indeg = np.zeros(n, np.int64)
    np.add.at(indeg, hn[rpar], 1)
    while True:
        has_act = np.zeros(n, bool)
        nonact = np.zeros(n, np.int64)
        a = rpar[state[rpar] == 1]
        has_act[hn[a]] = True
        b = rpar[state[rpar] == 2]
        np.add.at(nonact, hn[b], 1)
        new = state.copy()
        und = R & (state == 0)
        new[und & has_act] = 2
        new[und & ~has_act & (nonact == indeg)] = 1
        if (new == state).all():
            break
        state = new
    act = R & (state == 1)
    minpos2 = np.full(n, INF, np.int64)
    a = np.where(act & R[hn])[0]
    np.minimum.at(minpos2, hn[a], a)
    cr2 = act & R[hn] & (minpos2[hn] == ar)
    cf2 = cr2.astype(np.int64)
    cid2 = np.cumsum(cf2) - cf2 + nc0 + nc1
    out = vc.copy()
    j = np.where(act & ~R[hn])[0]
    out[j] = vc[hn[j]]
    j = np.where(act & R[hn])[0]
    out[j] = cid2[minpos2[hn[j]]]
    j = np.where(R & ~act)[0]
    out[j] = cid2[minpos2[j]]
    return out.astype(np.uint32), nc0 + nc1 + int(cf2.sum())
