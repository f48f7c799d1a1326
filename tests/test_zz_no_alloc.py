"""mi_compute_partition_noalloc, the sort-and-split that allocates no device
memory, called through the C ABI against mi_compute_partition on the same
graph and vector: partitions and cuts equal bit for bit.

The calls come in a row on one context, each on a smaller graph than the
one before (the largest at the size limit of 2^21 vertices, then a mid-size
graph through the radix passes, then one of at most 4096 vertices through
the one-workgroup sort), so every call after the first finds the previous
call's larger contents in the library's scratch array.  A graph of
2^21 + 1 vertices is refused with MI_STATUS_ARG.  The vectors carry many
ties and both zeros, which the split orders by vertex id."""
import ctypes as C

import numpy as np
import pytest

import ipdps_2021_graph_coarsening_amd as mi
from ipdps_2021_graph_coarsening_amd import graphs

pytestmark = pytest.mark.gpu

MAX_N = 1 << 21


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        return p

    def h2d(self, d, a):
        assert self.L.hipMemcpy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, d):
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), d, a.nbytes, 2) == 0


def _vector(n, seed):
    rng = np.random.default_rng(seed)
    ev = rng.integers(-500, 500, n).astype(np.float64) / 7.0  # many ties
    ev[rng.integers(0, n, max(1, n // 50))] = 0.0
    ev[rng.integers(0, n, max(1, n // 50))] = -0.0
    return ev


def _noalloc(ctx, hip, g, ev, num_partitions, imb):
    d_ev, d_part = hip.alloc(8 * g.n), hip.alloc(4 * g.n)
    try:
        hip.h2d(d_ev, ev)
        cut = C.c_uint32(0xFFFFFFFF)
        rc = ctx._L.mi_compute_partition_noalloc(ctx.h, g.h, d_ev, num_partitions, imb, d_part, C.byref(cut))
        part = np.full(g.n, 7, dtype=np.uint32)
        if rc == 0:
            hip.d2h(part, d_part)
        return rc, part, int(cut.value)
    finally:
        hip.L.hipFree(d_ev)
        hip.L.hipFree(d_part)


def _reference_call(ctx, g, ev, num_partitions, imb):
    part = np.zeros(g.n, dtype=np.uint32)
    cut = C.c_uint32(0)
    ctx.check(ctx._L.mi_compute_partition(ctx.h, g.h, ev, num_partitions, imb, part, C.byref(cut)),
              "mi_compute_partition")
    return part, int(cut.value)


def test_noalloc_split_equals_compute_partition_in_a_row():
    hip = _Hip()
    cases = [
        (graphs.path(MAX_N), 2, 0),                 # at the size limit
        (graphs.rgg(150_000, seed=11), 4, 0),      # radix passes; num_partitions 4 splits in two
        (graphs.grid2d(60), 2, 5),                 # 3,600 vertices: one-workgroup sort; imbalance walk
    ]
    with mi.Context(0) as ctx:
        for k, ((rp, col, _), parts, imb) in enumerate(cases):
            print(f"noalloc case {k}: n={len(rp) - 1} m={int(rp[-1])} partitions={parts} imbalance={imb}",
                  flush=True)
            g = ctx.graph(rp, col)
            ev = _vector(g.n, 100 + k)
            ref_part, ref_cut = _reference_call(ctx, g, ev, parts, imb)
            rc, part, cut = _noalloc(ctx, hip, g, ev, parts, imb)
            assert rc == 0, ctx._L.mi_last_error(ctx.h)
            assert np.array_equal(part, ref_part), k
            assert cut == ref_cut, (k, cut, ref_cut)
            g.close()


def test_noalloc_split_refuses_more_than_its_scratch_holds():
    hip = _Hip()
    rp, col, _ = graphs.path(MAX_N + 1)
    with mi.Context(0) as ctx:
        g = ctx.graph(rp, col)
        rc, _, _ = _noalloc(ctx, hip, g, np.zeros(g.n), 2, 0)
        assert rc == 1  # MI_STATUS_ARG
        assert b"scratch" in ctx._L.mi_last_error(ctx.h)
        # the context still works afterwards
        small = graphs.grid2d(20)
        g2 = ctx.graph(small[0], small[1])
        ev = _vector(g2.n, 5)
        rc, part, cut = _noalloc(ctx, hip, g2, ev, 2, 0)
        assert rc == 0
        ref_part, ref_cut = _reference_call(ctx, g2, ev, 2, 0)
        assert np.array_equal(part, ref_part) and cut == ref_cut
        g2.close()
        g.close()
