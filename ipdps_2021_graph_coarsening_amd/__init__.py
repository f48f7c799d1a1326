"""MI355X-native multilevel spectral graph bisection.

Python binding of libsgpar_mi.so (C ABI in include/sgpar_mi.h).  The binding
mirrors the reference's sgpar API (cmg/sgpar.h:2230-2244): a graph goes in
as CSR, ``partition_graph`` returns the 0/1 part vector and the edge cut.
All compute runs in the HIP kernels of the library; when the library is
missing or no GPU is present the calls raise -- there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import graphs  # noqa: F401  (CSR file format and synthetic generators)

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsgpar_mi.so")
MAX_LEVELS = 202

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
vp = C.c_void_p


class MiError(RuntimeError):
    """A non-zero status from the library."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class Config(C.Structure):
    """Mirror of config_t (sgpar.h:186-193), same field order."""
    _fields_ = [("coarsening_alg", C.c_int), ("refine_alg", C.c_int), ("local_search_alg", C.c_int),
                ("num_iter", C.c_int), ("tol", C.c_double), ("num_partitions", C.c_int)]


class StepBench(C.Structure):
    """mi_step_bench"""
    _fields_ = [("step_ms", C.c_double), ("stream_ms", C.c_double), ("step_bytes", C.c_double),
                ("stream_bytes", C.c_double), ("impl_bytes", C.c_double), ("column_blocks", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("nlevels", C.c_int),
                ("level_n", C.c_uint32 * MAX_LEVELS),
                ("level_m", C.c_uint32 * MAX_LEVELS),
                ("level_iters", C.c_uint64 * MAX_LEVELS),
                ("level_max_reached", C.c_int * MAX_LEVELS),
                ("level_eigen_ms", C.c_double * MAX_LEVELS),
                ("coarsen_ms", C.c_double), ("eigen_ms", C.c_double),
                ("partition_ms", C.c_double), ("total_ms", C.c_double),
                ("permute_ms", C.c_double), ("heavy_ms", C.c_double),
                ("map_ms", C.c_double), ("build_ms", C.c_double),
                ("edge_cut", C.c_uint32),
                ("has_eigen_log", C.c_int),
                ("eigenvalue", C.c_double), ("eigen_min", C.c_double), ("eigen_max", C.c_double),
                ("mixed", C.c_int)]

    def as_dict(self) -> dict:
        L = self.nlevels
        return dict(nlevels=L, level_n=list(self.level_n[:L]), level_m=list(self.level_m[:L]),
                    level_iters=[int(x) for x in self.level_iters[:L]],
                    level_max_reached=list(self.level_max_reached[:L]),
                    level_eigen_ms=list(self.level_eigen_ms[:L]),
                    coarsen_ms=self.coarsen_ms, eigen_ms=self.eigen_ms,
                    partition_ms=self.partition_ms, total_ms=self.total_ms,
                    permute_ms=self.permute_ms, heavy_ms=self.heavy_ms, map_ms=self.map_ms,
                    build_ms=self.build_ms, edge_cut=self.edge_cut,
                    eigen_log=(dict(eigenvalue=self.eigenvalue, min=self.eigen_min, max=self.eigen_max)
                               if self.has_eigen_log else None),
                    mixed_levels=int(self.mixed))


_lib = None

# (name, restype, argtypes) of every C-ABI entry point
_SIGS = [
    ("mi_ctx_create", C.c_int, [C.c_int, C.POINTER(vp)]),
    ("mi_ctx_destroy", None, [vp]),
    ("mi_last_error", C.c_char_p, [vp]),
    ("mi_ctx_trim", C.c_int, [vp]),
    ("mi_ctx_set_graphs", C.c_int, [vp, C.c_int]),
    ("mi_ctx_set_precision", C.c_int, [vp, C.c_int, C.c_uint32]),
    ("mi_ctx_set_small_levels", C.c_int, [vp, C.c_uint32, C.c_uint32]),
    ("mi_ctx_set_locality", C.c_int, [vp, C.c_int, C.c_uint32]),
    ("mi_ctx_set_bucketed_build", C.c_int, [vp, C.c_uint32]),
    ("mi_ctx_set_refinement", C.c_int, [vp, C.c_int]),
    ("mi_ctx_set_coarsening", C.c_int, [vp, C.c_int]),
    ("mi_ctx_set_builder", C.c_int, [vp, C.c_int]),
    ("mi_fm_eigensolve", C.c_int, [vp, vp, C.c_uint64, C.c_double, vp, C.POINTER(C.c_uint32), C.POINTER(Stats)]),
    ("mi_fm_refine", C.c_int, [vp, vp, vp, u32p, C.POINTER(C.c_uint32), vp]),
    ("mi_fm_init", C.c_int, [vp, vp, vp, C.c_uint64, C.c_double, u32p]),
    ("mi_graph_create", C.c_int, [vp, C.c_uint32, u32p, u32p, vp, C.POINTER(vp)]),
    ("mi_graph_from_edges", C.c_int, [vp, C.c_uint32, C.c_uint64, u32p, u32p, vp, C.c_int, C.POINTER(vp)]),
    ("mi_graph_create_device", C.c_int, [vp, C.c_uint32, C.c_uint32, vp, vp, vp, C.POINTER(vp)]),
    ("mi_graph_destroy", None, [vp]),
    ("mi_graph_size", C.c_int, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("mi_graph_copy", C.c_int, [vp, vp, vp, vp, vp]),
    ("mi_hierarchy_build", C.c_int, [vp, vp, C.c_uint64, C.POINTER(vp), C.POINTER(Stats)]),
    ("mi_hier_create_from_host", C.c_int, [vp, C.c_int, u32p, C.POINTER(vp), C.POINTER(vp),
                                           C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    ("mi_hier_destroy", None, [vp]),
    ("mi_hier_nlevels", C.c_int, [vp]),
    ("mi_hier_level_size", C.c_int, [vp, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("mi_hier_level_copy", C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp]),
    ("mi_fiedler", C.c_int, [vp, vp, C.POINTER(C.c_uint64), C.c_uint64, C.c_int, C.c_double, vp,
                             C.POINTER(Stats)]),
    ("mi_power_iter", C.c_int, [vp, vp, f64p, C.c_int, C.c_double, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_int)]),
    ("mi_compute_partition", C.c_int, [vp, vp, f64p, C.c_int, C.c_int, u32p, C.POINTER(C.c_uint32)]),
    ("mi_compute_partition_noalloc", C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, C.POINTER(C.c_uint32)]),
    ("mi_partition_graph", C.c_int, [vp, vp, C.POINTER(Config), C.c_int, C.c_uint64,
                                     C.POINTER(C.c_uint64), C.c_uint64, vp, C.c_int,
                                     C.POINTER(C.c_uint32), C.POINTER(Stats)]),
    ("mi_bisect_batch", C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(Config), C.c_int, C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64), C.c_uint64, C.POINTER(vp), C.c_int,
                                  C.POINTER(C.c_uint32), C.POINTER(Stats), C.POINTER(C.c_double)]),
    ("mi_use_partition", C.c_int, [vp, vp, u32p, C.c_int, C.POINTER(vp), C.POINTER(vp)]),
    ("mi_partition_4way", C.c_int, [vp, vp, C.POINTER(Config), C.c_int, C.c_uint64, C.POINTER(C.c_uint64),
                                    C.c_uint64, vp, C.c_int, u32p, C.POINTER(C.c_uint32)]),
    ("mi_xs_draws", C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64, u64p]),
    ("mi_hec_level", C.c_int, [vp, vp, C.c_int, C.c_uint64, u32p, vp, vp, C.POINTER(C.c_uint32)]),
    ("mi_build_coarse", C.c_int, [vp, vp, u32p, C.c_uint32, C.POINTER(vp), vp]),
    ("mi_spmv_shift", C.c_int, [vp, vp, f64p, f64p]),
    ("mi_bench_step", C.c_int, [vp, vp, C.c_int, C.c_void_p]),
    ("mi_radix_sort_pairs", C.c_int, [vp, u64p, u32p, C.c_uint64, C.c_int, C.c_int]),
    ("mi_sort_random_keys", C.c_int, [vp, u64p, u32p, C.c_uint64]),
    ("mi_scan_u32", C.c_int, [vp, u32p, u32p, C.c_uint64, C.POINTER(C.c_uint32)]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib() -> C.CDLL:
    """Loads libsgpar_mi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m ipdps_2021_graph_coarsening_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(vp)


# Handles still open when the interpreter exits are closed by an exit hook,
# while the HIP runtime is certainly up and in a fixed order (hierarchies
# and graphs before the contexts they were made on), instead of by the
# garbage collector's finalisers in whatever order it reaches them.
_live: "weakref.WeakSet" = weakref.WeakSet()


def _close_live() -> None:
    for o in sorted(list(_live), key=lambda o: isinstance(o, Context)):
        try:
            o.close()
        except Exception:
            pass


atexit.register(_close_live)


class Context:
    """A library context on one HIP device (stream, pool, workspace)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = vp()
        rc = self._L.mi_ctx_create(device, C.byref(h))
        if rc:
            raise MiError(rc, f"mi_ctx_create(device={device}) failed (no usable HIP device?)")
        self.h = h
        _live.add(self)

    def check(self, rc: int, what: str) -> None:
        if rc:
            msg = self._L.mi_last_error(self.h)
            raise MiError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "h", None):
            self._L.mi_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_locality(self, enable: bool = True, min_vertices: int = 0) -> None:
        """Solve large gather-form levels in a locality order (default on)."""
        self.check(self._L.mi_ctx_set_locality(self.h, 1 if enable else 0, min_vertices), "mi_ctx_set_locality")

    def set_bucketed_build(self, min_entries: int = 0xFFFFFFFF) -> None:
        """Coarse builds of levels with >= min_entries entries take the
        bucketed form (0: every level; the default: none); results are
        identical."""
        self.check(self._L.mi_ctx_set_bucketed_build(self.h, min_entries), "mi_ctx_set_bucketed_build")

    def set_refinement(self, fm: bool) -> None:
        """Bisections refine by FM, as the reference's *_fm binaries (True), or
        by the multilevel power iteration (False, the default)."""
        self.check(self._L.mi_ctx_set_refinement(self.h, 1 if fm else 0), "mi_ctx_set_refinement")

    def set_coarsening(self, variant: int) -> None:
        """The coarsening of the hierarchies: 0 the hec_* binaries' (default),
        1 the match_* binaries' (heavy-edge matching), 2 the mtmetis_* binaries'
        (that matching plus the leaf, twin and relative passes), 4 the hec2_*
        binaries' (parallel_map_construct_v2),
        3 the hec3_* binaries' (parallel_map_construct_v3), 5 the mis_*
        binaries' (MIS-2 aggregates), 6 the gosh_* binaries' (GOSH centres), 7 the
        gosh2_* binaries' (GOSH v2), 8 hec_spec_serial's (the serial HEC:
        a Fisher-Yates shuffle and a greedy map in list order)."""
        self.check(self._L.mi_ctx_set_coarsening(self.h, variant), "mi_ctx_set_coarsening")

    def set_builder(self, builder: int) -> None:
        """The coarse-graph build: 0 the RADIX build (rows by column, default),
        1 the HASHMAP build of hec_spec_hashmap (rows in order of first
        occurrence)."""
        self.check(self._L.mi_ctx_set_builder(self.h, builder), "mi_ctx_set_builder")

    def set_small_levels(self, max_entries: int = 1 << 17, chunk: int = 0) -> None:
        """Levels of <= 16384 vertices and <= max_entries entries iterate in one
        workgroup (0: never), at most chunk iterations per launch (0: 8192)."""
        self.check(self._L.mi_ctx_set_small_levels(self.h, max_entries, chunk), "mi_ctx_set_small_levels")

    def set_precision(self, mixed: bool, min_entries: int = 0) -> None:
        """Mixed-precision eigensolve (FP32 on levels above the finest with at
        least min_entries entries, 0: 2^20; FP64 finest level) or plain FP64."""
        self.check(self._L.mi_ctx_set_precision(self.h, 1 if mixed else 0, min_entries), "mi_ctx_set_precision")

    def set_graphs(self, enable: bool) -> None:
        """Replay power iterations from a captured HIP graph, or launch every kernel (the default)."""
        self.check(self._L.mi_ctx_set_graphs(self.h, 1 if enable else 0), "mi_ctx_set_graphs")

    # ------------------------------------------------------------ graphs
    def graph(self, rowptr, col, w=None) -> "Graph":
        return Graph(self, rowptr, col, w)

    def graph_from_edges(self, n: int, u, v, w=None, lcc: bool = True) -> "Graph":
        """The reference converter's CSR graph of an edge list, built on the
        device (mi_graph_from_edges): symmetrised, self-loops dropped,
        duplicates merged (weights summed; no w: unit weights), rows sorted,
        with lcc only the largest connected component in vertex order."""
        u = _u32(u)
        v = _u32(v)
        if len(u) != len(v):
            raise ValueError("u and v differ in length")
        wa = None if w is None else _u32(w)
        h = vp()
        self.check(self._L.mi_graph_from_edges(self.h, n, len(u), u, v, _ptr(wa), 1 if lcc else 0, C.byref(h)),
                   "mi_graph_from_edges")
        return Graph(self, None, None, None, _handle=h)

    # ------------------------------------------------------------ building blocks
    def xs_draws(self, seed: int, first: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.uint64)
        self.check(self._L.mi_xs_draws(self.h, seed, first, count, out), "mi_xs_draws")
        return out

    def radix_sort_pairs(self, keys, vals, begin_bit=0, end_bit=64):
        k = np.array(keys, dtype=np.uint64, copy=True)
        v = np.array(vals, dtype=np.uint32, copy=True)
        self.check(self._L.mi_radix_sort_pairs(self.h, k, v, len(k), begin_bit, end_bit),
                   "mi_radix_sort_pairs")
        return k, v

    def sort_random_keys(self, keys, vals):
        """Stable 64-bit sort as used for the HEC permutation (mi_sort_random_keys)."""
        k = np.array(keys, dtype=np.uint64, copy=True)
        v = np.array(vals, dtype=np.uint32, copy=True)
        self.check(self._L.mi_sort_random_keys(self.h, k, v, len(k)), "mi_sort_random_keys")
        return k, v

    def scan_u32(self, a):
        a = _u32(a)
        out = np.empty_like(a)
        tot = C.c_uint32(0)
        self.check(self._L.mi_scan_u32(self.h, a, out, len(a), C.byref(tot)), "mi_scan_u32")
        return out, tot.value

    def hier_from_host(self, levels) -> "Hierarchy":
        """levels: list of dict(rowptr, col, w (or None), vcmap (None at coarsest))."""
        return Hierarchy.from_host(self, levels)


class Graph:
    """A CSR graph resident in HBM."""

    def __init__(self, ctx: Context, rowptr, col, w=None, _handle=None):
        self.ctx = ctx
        if _handle is not None:
            self.h = _handle
        else:
            rowptr = _u32(rowptr)
            col = _u32(col)
            wa = None if w is None else _u32(w)
            n = len(rowptr) - 1
            h = vp()
            ctx.check(ctx._L.mi_graph_create(ctx.h, n, rowptr, col, _ptr(wa), C.byref(h)),
                      "mi_graph_create")
            self.h = h
        _live.add(self)
        n, m = C.c_uint32(), C.c_uint32()
        ctx._L.mi_graph_size(self.h, C.byref(n), C.byref(m))
        self.n, self.m = n.value, m.value

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.mi_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ steps
    def hec_level(self, level: int, seed: int):
        """One HEC level: (vcmap, nc, vperm, hn)."""
        c = self.ctx
        vcmap = np.empty(self.n, dtype=np.uint32)
        vperm = np.empty(self.n, dtype=np.uint32)
        hn = np.empty(self.n, dtype=np.uint32)
        nc = C.c_uint32()
        c.check(c._L.mi_hec_level(c.h, self.h, level, seed, vcmap, _ptr(vperm), _ptr(hn), C.byref(nc)),
                "mi_hec_level")
        return vcmap, nc.value, vperm, hn

    def build_coarse(self, vcmap, nc: int):
        """Coarse graph under vcmap: (Graph, c_vtx_w)."""
        c = self.ctx
        h = vp()
        cvw = np.empty(max(nc, 1), dtype=np.uint32)
        c.check(c._L.mi_build_coarse(c.h, self.h, _u32(vcmap), nc, C.byref(h), _ptr(cvw)),
                "mi_build_coarse")
        return Graph(c, None, None, _handle=h), cvw[:nc]

    def spmv_shift(self, x) -> np.ndarray:
        c = self.ctx
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        c.check(c._L.mi_spmv_shift(c.h, self.h, x, y), "mi_spmv_shift")
        return y

    def bench_step(self, reps: int = 20) -> dict:
        """Times the power-iteration step on this graph (mi_bench_step)."""
        c = self.ctx
        b = StepBench()
        c.check(c._L.mi_bench_step(c.h, self.h, reps, C.byref(b)), "mi_bench_step")
        return {f: getattr(b, f) for f, _ in StepBench._fields_}

    def bench_spmv(self, reps: int = 20):
        """(ms, algorithmic bytes) of one whole power-iteration step."""
        b = self.bench_step(reps)
        return b["step_ms"], b["step_bytes"]

    def power_iter(self, u, normLap: int = 0, tol: float = 1e-10):
        c = self.ctx
        u = np.array(u, dtype=np.float64, copy=True)
        it = C.c_uint64()
        mr = C.c_int()
        c.check(c._L.mi_power_iter(c.h, self.h, u, normLap, tol, C.byref(it), C.byref(mr)),
                "mi_power_iter")
        return u, it.value, bool(mr.value)

    def fm_refine(self, part, vtx_w=None):
        """One fm_refine call (eigensolve_kokkos.h:973-1240): (sides, min_cut,
        info dict(swaps, argmin, min_imb, maxE))."""
        c = self.ctx
        p = np.ascontiguousarray(np.asarray(part) != 0, dtype=np.uint32)
        vw = None if vtx_w is None else _u32(vtx_w)
        mc = C.c_uint32()
        info = np.zeros(4, dtype=np.int64)
        c.check(c._L.mi_fm_refine(c.h, self.h, _ptr(vw), p, C.byref(mc), _ptr(info)), "mi_fm_refine")
        return p, mc.value, dict(swaps=int(info[0]), argmin=int(info[1]), min_imb=int(info[2]), maxE=int(info[3]))

    def fm_init(self, vtx_w=None, seed: int = 0, tol: float = 1e-10) -> np.ndarray:
        """The FM path's coarsest-level partition: init_gggp (n <= 200) or
        init_spectral (eigensolve_kokkos.h:1242-1368)."""
        c = self.ctx
        vw = None if vtx_w is None else _u32(vtx_w)
        p = np.empty(self.n, dtype=np.uint32)
        c.check(c._L.mi_fm_init(c.h, self.h, _ptr(vw), seed, tol, p), "mi_fm_init")
        return p

    def compute_partition(self, evec, num_partitions: int = 2, imbalance: int = 0):
        c = self.ctx
        part = np.empty(self.n, dtype=np.uint32)
        ec = C.c_uint32()
        c.check(c._L.mi_compute_partition(c.h, self.h, np.ascontiguousarray(evec, dtype=np.float64),
                                          num_partitions, imbalance, part, C.byref(ec)),
                "mi_compute_partition")
        return part, ec.value

    def use_partition(self, part):
        """Induced subgraphs of the sides of part (sgp_use_partition, sgpar.h:2666-2752),
        built on the device: (Graph, Graph), unit weights."""
        c = self.ctx
        h1, h2 = vp(), vp()
        c.check(c._L.mi_use_partition(c.h, self.h, _u32(part), 0, C.byref(h1), C.byref(h2)), "mi_use_partition")
        return Graph(c, None, None, _handle=h1), Graph(c, None, None, _handle=h2)

    def hierarchy(self, seed: int) -> "Hierarchy":
        return Hierarchy.build(self, seed)

    def to_host(self):
        """(rowptr, col, w) copied back from HBM (w is all ones for a unit graph)."""
        c = self.ctx
        rowptr = np.empty(self.n + 1, dtype=np.uint32)
        col = np.empty(max(self.m, 1), dtype=np.uint32)
        w = np.empty(max(self.m, 1), dtype=np.uint32)
        c.check(c._L.mi_graph_copy(c.h, self.h, _ptr(rowptr), _ptr(col), _ptr(w)), "mi_graph_copy")
        return rowptr, col[:self.m], w[:self.m]


class Hierarchy:
    """The multilevel hierarchy (level 0 = finest)."""

    def __init__(self):
        self.h = None

    @classmethod
    def build(cls, g: Graph, seed: int, stats: bool = True) -> "Hierarchy":
        self = cls()
        c = g.ctx
        self.ctx = c
        self.graph = g
        h = vp()
        st = Stats()
        c.check(c._L.mi_hierarchy_build(c.h, g.h, seed, C.byref(h), C.byref(st) if stats else None),
                "mi_hierarchy_build")
        self.h = h
        _live.add(self)
        self.stats = st.as_dict()
        return self

    @classmethod
    def from_host(cls, c: Context, levels) -> "Hierarchy":
        self = cls()
        self.ctx = c
        self.graph = None
        L = len(levels)
        keep = []
        ns = np.array([len(lv["rowptr"]) - 1 for lv in levels], dtype=np.uint32)

        def arr(a):
            if a is None:
                return None
            a = _u32(a)
            keep.append(a)
            return a.ctypes.data_as(vp)

        rp = (vp * L)(*[arr(lv["rowptr"]) for lv in levels])
        cp = (vp * L)(*[arr(lv["col"]) for lv in levels])
        wp = (vp * L)(*[arr(lv.get("w")) for lv in levels])
        vc = (vp * L)(*[arr(lv.get("vcmap")) if i + 1 < L else None for i, lv in enumerate(levels)])
        h = vp()
        c.check(c._L.mi_hier_create_from_host(c.h, L, ns, rp, cp, wp, vc, C.byref(h)),
                "mi_hier_create_from_host")
        self.h = h
        _live.add(self)
        self.stats = None
        return self

    def close(self):
        if getattr(self, "h", None):
            self.ctx._L.mi_hier_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nlevels(self) -> int:
        return self.ctx._L.mi_hier_nlevels(self.h)

    def level(self, l: int) -> dict:
        c = self.ctx
        n, m = C.c_uint32(), C.c_uint32()
        c.check(c._L.mi_hier_level_size(self.h, l, C.byref(n), C.byref(m)), "mi_hier_level_size")
        n, m = n.value, m.value
        rowptr = np.empty(n + 1, dtype=np.uint32)
        col = np.empty(max(m, 1), dtype=np.uint32)
        w = np.empty(max(m, 1), dtype=np.uint32)
        vtx_w = np.empty(n, dtype=np.uint32)
        has_map = l + 1 < self.nlevels
        vcmap = np.empty(n, dtype=np.uint32) if has_map else None
        c.check(c._L.mi_hier_level_copy(c.h, self.h, l, _ptr(rowptr), _ptr(col), _ptr(w), _ptr(vcmap),
                                        _ptr(vtx_w)), "mi_hier_level_copy")
        return dict(n=n, m=m, rowptr=rowptr, col=col[:m], w=w[:m], vcmap=vcmap, vtx_w=vtx_w)

    def fiedler(self, pcg_state: int, pcg_inc: int = 1, normLap: int = 0, tol: float = 1e-10):
        """(evec, stats dict, new pcg state)."""
        c = self.ctx
        n0 = C.c_uint32()
        c._L.mi_hier_level_size(self.h, 0, C.byref(n0), None)
        ev = np.empty(n0.value, dtype=np.float64)
        st = C.c_uint64(pcg_state)
        stats = Stats()
        c.check(c._L.mi_fiedler(c.h, self.h, C.byref(st), pcg_inc, normLap, tol, _ptr(ev), C.byref(stats)),
                "mi_fiedler")
        return ev, stats.as_dict(), st.value


    def fm_eigensolve(self, seed: int = 0, tol: float = 1e-10):
        """FM branch of the eigensolve: (0/1 evec, last cut, stats dict)."""
        c = self.ctx
        n0 = C.c_uint32()
        c._L.mi_hier_level_size(self.h, 0, C.byref(n0), None)
        ev = np.empty(n0.value, dtype=np.float64)
        cut = C.c_uint32()
        stats = Stats()
        c.check(c._L.mi_fm_eigensolve(c.h, self.h, seed, tol, _ptr(ev), C.byref(cut), C.byref(stats)),
                "mi_fm_eigensolve")
        return ev, cut.value, stats.as_dict()


@dataclass
class PartitionResult:
    part: np.ndarray
    edge_cut: int
    stats: dict
    pcg_state: int
    extra: dict = field(default_factory=dict)


def partition_graph(ctx: Context, g: Graph, config: Config | None = None, perc_imbalance: int = 0,
                    hec_seed: int = 0, pcg_state: int = 0, pcg_inc: int = 1) -> PartitionResult:
    """Whole bisection, the counterpart of sgp_partition_graph (sgpar.h:2363-2431)."""
    if config is None:
        config = Config(4, 0, 0, 1, 1e-10, 2)
    part = np.empty(g.n, dtype=np.uint32)
    st = C.c_uint64(pcg_state)
    ec = C.c_uint32()
    stats = Stats()
    ctx.check(ctx._L.mi_partition_graph(ctx.h, g.h, C.byref(config), perc_imbalance, hec_seed, C.byref(st),
                                        pcg_inc, _ptr(part), 0, C.byref(ec), C.byref(stats)),
              "mi_partition_graph")
    return PartitionResult(part=part, edge_cut=ec.value, stats=stats.as_dict(), pcg_state=st.value)


def bisect_batch(ctx: Context, graphs: list, config: Config | None = None, perc_imbalance: int = 0,
                 hec_seeds=None, pcg_states=None, pcg_inc: int = 1):
    """Bisections of many graphs with the small levels' power iterations run
    together (mi_bisect_batch); per graph the same result as partition_graph.
    Returns (list of PartitionResult, phase times in ms: coarsen, eigen, split)."""
    if config is None:
        config = Config(4, 0, 0, 1, 1e-10, 2)
    K = len(graphs)
    seeds = (C.c_uint64 * max(K, 1))(*(hec_seeds if hec_seeds is not None else [0] * K))
    states = (C.c_uint64 * max(K, 1))(*(pcg_states if pcg_states is not None else [0] * K))
    parts = [np.empty(g.n, dtype=np.uint32) for g in graphs]
    hs = (C.c_void_p * max(K, 1))(*[g.h for g in graphs])
    pp = (C.c_void_p * max(K, 1))(*[p.ctypes.data for p in parts])
    cuts = (C.c_uint32 * max(K, 1))()
    stats = (Stats * max(K, 1))()
    ph = (C.c_double * 3)()
    ctx.check(ctx._L.mi_bisect_batch(ctx.h, K, hs, C.byref(config), perc_imbalance, seeds, states, pcg_inc, pp, 0,
                                     cuts, stats, ph), "mi_bisect_batch")
    res = [PartitionResult(part=parts[k], edge_cut=cuts[k], stats=stats[k].as_dict(), pcg_state=states[k])
           for k in range(K)]
    return res, list(ph)


def partition_4way(ctx: Context, g: Graph, config: Config | None = None, perc_imbalance: int = 0,
                   hec_seed: int = 0, pcg_state: int = 0, pcg_inc: int = 1) -> PartitionResult:
    """Recursive 4-way partition (the driver of sgpar.cpp:53-89) on the device.
    part in {0..3}; edge_cut = sum of the three bisection cuts; extra['cuts']."""
    if config is None:
        config = Config(4, 0, 0, 1, 1e-10, 4)
    part = np.empty(g.n, dtype=np.uint32)
    st = C.c_uint64(pcg_state)
    cuts = np.zeros(3, dtype=np.uint32)
    ec = C.c_uint32()
    ctx.check(ctx._L.mi_partition_4way(ctx.h, g.h, C.byref(config), perc_imbalance, hec_seed, C.byref(st), pcg_inc,
                                       part.ctypes.data_as(vp), 0, cuts, C.byref(ec)), "mi_partition_4way")
    return PartitionResult(part=part, edge_cut=ec.value, stats={}, pcg_state=st.value,
                           extra={"cuts": [int(x) for x in cuts]})
