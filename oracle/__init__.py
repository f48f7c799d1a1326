"""CPU oracle for the multilevel spectral bisection path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by
the CPU-baseline leg of bench.py, always as the checker.  The product package
never imports it.

``oracle.core`` wraps liboracle.so (the C restatement in sgpar_oracle.c);
``oracle.ref`` wraps oracle/_ref/libsgpar_ref.so, the reference's own
non-Kokkos sgpar.h compiled from the reference tree, when it was built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsgpar_ref.so")

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and the reference wrapper when its sources exist)."""
    target = "all" if with_ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        build(with_ref=False)
    return C.CDLL(path)


class _Core:
    def __init__(self) -> None:
        L = _load(ORACLE_SO)
        self.L = L
        L.orc_pcg32.restype = C.c_uint32
        L.orc_pcg32.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
        L.orc_xs_step.restype = C.c_uint64
        L.orc_xs_step.argtypes = [C.c_uint64]
        L.orc_xs_urand64.restype = C.c_uint64
        L.orc_xs_urand64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_pool_state.restype = C.c_uint64
        L.orc_pool_state.argtypes = [C.c_uint64]
        L.orc_generate_permutation.restype = None
        L.orc_generate_permutation.argtypes = [C.c_uint32, C.POINTER(C.c_uint64), u32p, u64p]
        L.orc_heavy_neighbor.restype = None
        L.orc_heavy_neighbor.argtypes = [C.c_int, C.c_uint32, u32p, u32p, C.c_void_p,
                                         C.POINTER(C.c_uint64), u32p]
        L.orc_map_construct.restype = C.c_uint32
        L.orc_map_construct.argtypes = [C.c_uint32, u32p, u32p, u32p, u32p, C.POINTER(C.c_int)]
        L.orc_coarsen_hec.restype = C.c_uint32
        L.orc_coarsen_hec.argtypes = [C.c_int, C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64,
                                      u32p, u32p, u32p]
        L.orc_build_coarse.restype = C.c_uint32
        L.orc_build_coarse.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, u32p, C.c_uint32,
                                       C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p), u32p]
        L.orc_build_coarse_order.restype = C.c_uint32
        L.orc_build_coarse_order.argtypes = L.orc_build_coarse.argtypes + [C.c_int]
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_hierarchy_build.restype = C.c_void_p
        L.orc_hierarchy_build.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64]
        L.orc_hier_nlevels.restype = C.c_int
        L.orc_hier_nlevels.argtypes = [C.c_void_p]
        L.orc_hier_level.restype = None
        L.orc_hier_level.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 7
        L.orc_hier_free.restype = None
        L.orc_hier_free.argtypes = [C.c_void_p]
        L.orc_power_iter.restype = C.c_uint64
        L.orc_power_iter.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, f64p, C.c_int,
                                     C.c_double, C.c_uint64, C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
        L.orc_eigensolve.restype = C.c_int
        L.orc_eigensolve.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64, C.c_int,
                                     C.c_double, C.c_uint64, f64p, u64p]
        L.orc_eigensolve_trace.restype = C.c_int
        L.orc_eigensolve_trace.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64, C.c_int,
                                           C.c_double, C.c_uint64, f64p, u64p, C.c_void_p, C.c_void_p]
        L.orc_eigenvalue_log.restype = None
        L.orc_eigenvalue_log.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, f64p, f64p]
        L.orc_compute_partition.restype = C.c_int
        L.orc_compute_partition.argtypes = [C.c_uint32, u32p, u32p, f64p, C.c_int, C.c_int,
                                            u32p, C.POINTER(C.c_uint32)]
        L.orc_partition_graph.restype = C.c_int
        L.orc_partition_graph.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64,
                                          C.POINTER(C.c_uint64), C.c_uint64, C.c_int, C.c_double,
                                          u32p, C.POINTER(C.c_uint32), u32p, u64p, f64p]
        L.orc_map_construct_v3.restype = C.c_uint32
        L.orc_map_construct_v3.argtypes = [C.c_uint32, u32p, u32p]
        L.orc_coarsen_hec3.restype = C.c_uint32
        L.orc_coarsen_hec3.argtypes = [C.c_int, C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, u32p, u32p]
        L.orc_mis2.restype = C.c_int
        L.orc_mis2.argtypes = [C.c_uint32, u32p, u32p, C.c_uint64, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]
        L.orc_coarsen_mis2.restype = C.c_uint32
        L.orc_coarsen_mis2.argtypes = [C.c_uint32, u32p, u32p, C.c_uint64, u32p]
        L.orc_coarsen_match.restype = C.c_uint32
        L.orc_coarsen_match.argtypes = [C.c_int, C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, u32p,
                                        C.POINTER(C.c_int)]
        L.orc_coarsen_mtmetis.restype = C.c_uint32
        L.orc_coarsen_mtmetis.argtypes = L.orc_coarsen_match.argtypes
        L.orc_murmur3_x86_32.restype = C.c_uint32
        L.orc_murmur3_x86_32.argtypes = [C.c_char_p, C.c_int, C.c_uint32]
        L.orc_binsort_u32.restype = None
        L.orc_binsort_u32.argtypes = [C.c_uint32, u32p, u32p]
        L.orc_map_construct_v2.restype = C.c_uint32
        L.orc_map_construct_v2.argtypes = [C.c_uint32, u32p, u32p, u32p, u32p, C.POINTER(C.c_int)]
        L.orc_coarsen_hec2.restype = C.c_uint32
        L.orc_coarsen_hec2.argtypes = L.orc_coarsen_match.argtypes
        L.orc_coarsen_gosh2.restype = C.c_uint32
        L.orc_coarsen_gosh2.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, u32p]
        L.orc_map_construct_prefilled.restype = C.c_uint32
        L.orc_map_construct_prefilled.argtypes = [C.c_uint32, u32p, C.c_uint32, u32p, u32p, C.c_uint32]
        L.orc_coarsen_hec_serial.restype = C.c_uint32
        L.orc_coarsen_hec_serial.argtypes = [C.c_int, C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, u32p, u32p, u32p]
        L.orc_gosh_clusters.restype = C.c_int
        L.orc_gosh_clusters.argtypes = [C.c_uint32, u32p, u32p, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]
        L.orc_coarsen_gosh.restype = C.c_uint32
        L.orc_coarsen_gosh.argtypes = [C.c_uint32, u32p, u32p, u32p]
        L.orc_hierarchy_build_variant.restype = C.c_void_p
        L.orc_hierarchy_build_variant.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, C.c_int]
        L.orc_partition_graph_variant.restype = C.c_int
        L.orc_partition_graph_variant.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64,
                                                  C.POINTER(C.c_uint64), C.c_uint64, C.c_int, C.c_double,
                                                  u32p, C.POINTER(C.c_uint32), u32p, u64p, f64p, C.c_int]
        u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_fm_refine.restype = C.c_uint32
        L.orc_fm_refine.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_void_p, u8p, i64p]
        L.orc_init_gggp.restype = None
        L.orc_init_gggp.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_void_p, u8p]
        L.orc_init_spectral.restype = None
        L.orc_init_spectral.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_void_p, C.c_uint64,
                                        C.c_double, u8p, f64p]
        L.orc_fm_eigensolve.restype = C.c_uint32
        L.orc_fm_eigensolve.argtypes = [C.c_void_p, C.c_uint64, C.c_double, u8p, u32p]
        L.orc_partition_graph_fm.restype = C.c_int
        L.orc_partition_graph_fm.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, C.c_double,
                                             u32p, C.POINTER(C.c_uint32), u8p, u32p]
        L.orc_partition_graph_fm_variant.restype = C.c_int
        L.orc_partition_graph_fm_variant.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, C.c_double,
                                                     u32p, C.POINTER(C.c_uint32), u8p, u32p, C.c_int]


_core = None


def core() -> _Core:
    global _core
    if _core is None:
        _core = _Core()
    return _core


def _wptr(w):
    if w is None:
        return None, None
    w = np.ascontiguousarray(w, dtype=np.uint32)
    return w, w.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- RNGs
def pcg32_stream(state: int, count: int, inc: int = 1):
    """`count` PCG32 outputs from (state, inc); returns (values, new_state)."""
    L = core().L
    st = C.c_uint64(state)
    out = np.empty(count, dtype=np.uint32)
    for i in range(count):
        out[i] = L.orc_pcg32(C.byref(st), inc)
    return out, st.value


def pool_state(seed: int) -> int:
    return core().L.orc_pool_state(seed)


def xs_step(s: int) -> int:
    return core().L.orc_xs_step(s)


def urand64_stream(state: int, count: int):
    """`count` urand64 draws from a pool state; returns (values, new_state)."""
    L = core().L
    st = C.c_uint64(state)
    out = np.empty(count, dtype=np.uint64)
    for i in range(count):
        out[i] = L.orc_xs_urand64(C.byref(st))
    return out, st.value


# ---------------------------------------------------------------- HEC
def generate_permutation(n: int, seed: int):
    """(vperm, keys, state_after) for the pool of `seed`."""
    L = core().L
    st = C.c_uint64(pool_state(seed))
    vperm = np.empty(n, dtype=np.uint32)
    keys = np.empty(n, dtype=np.uint64)
    L.orc_generate_permutation(n, C.byref(st), vperm, keys)
    return vperm, keys, st.value


def heavy_neighbor(level: int, rowptr, col, w, state: int):
    L = core().L
    n = len(rowptr) - 1
    st = C.c_uint64(state)
    hn = np.empty(n, dtype=np.uint32)
    w, wp = _wptr(w)
    L.orc_heavy_neighbor(level, n, rowptr, col, wp, C.byref(st), hn)
    return hn, st.value


def map_construct(vperm, hn):
    """Literal sequential model of the map construction; (vcmap, nc, passes)."""
    L = core().L
    vperm = np.ascontiguousarray(vperm, dtype=np.uint32)
    hn = np.ascontiguousarray(hn, dtype=np.uint32)
    n = len(vperm)
    pos = np.empty(n, dtype=np.uint32)
    pos[vperm] = np.arange(n, dtype=np.uint32)
    vcmap = np.empty(n, dtype=np.uint32)
    passes = C.c_int(0)
    nc = L.orc_map_construct(n, vperm, hn, pos, vcmap, C.byref(passes))
    return vcmap, int(nc), passes.value


def coarsen_hec(level: int, rowptr, col, w, seed: int):
    """One HEC level; returns (vcmap, nc, vperm, hn)."""
    L = core().L
    n = len(rowptr) - 1
    vcmap = np.empty(n, dtype=np.uint32)
    vperm = np.empty(n, dtype=np.uint32)
    hn = np.empty(n, dtype=np.uint32)
    w, wp = _wptr(w)
    nc = L.orc_coarsen_hec(level, n, rowptr, col, wp, seed, vcmap, vperm, hn)
    return vcmap, int(nc), vperm, hn


def build_coarse(rowptr, col, w, vcmap, nc: int, vtx_w=None, builder: int = 0):
    """Coarse CSR (rowptr_c, col_c, w_c, c_vtx_w) with merged rows: sorted by
    column (builder 0, RADIX) or in order of first occurrence (builder 1,
    the HASHMAP build of hec_spec_hashmap)."""
    L = core().L
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vw, vwp = _wptr(vtx_w)
    rp, cp, wc = C.c_void_p(), C.c_void_p(), C.c_void_p()
    cvw = np.zeros(nc + 1, dtype=np.uint32)
    mc = L.orc_build_coarse_order(n, np.ascontiguousarray(rowptr, dtype=np.uint32),
                                  np.ascontiguousarray(col, dtype=np.uint32), wp,
                                  np.ascontiguousarray(vcmap, dtype=np.uint32),
                                  nc, vwp, C.byref(rp), C.byref(cp), C.byref(wc), cvw, builder)
    r = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint32)), shape=(nc + 1,)).copy()
    c = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_uint32)), shape=(max(mc, 1),))[:mc].copy()
    ww = np.ctypeslib.as_array(C.cast(wc, C.POINTER(C.c_uint32)), shape=(max(mc, 1),))[:mc].copy()
    for p in (rp, cp, wc):
        L.orc_free(p)
    return r, c, ww, cvw[:nc]


class Hierarchy:
    """Multilevel hierarchy from the oracle (level 0 = finest)."""

    def __init__(self, rowptr, col, w, seed: int, variant: int = 0, builder: int = 0):
        L = core().L
        self._L = L
        n = len(rowptr) - 1
        w, wp = _wptr(w)
        self._h = L.orc_hierarchy_build_variant(n, rowptr, col, wp, seed, variant | (builder << 8))
        self.nlevels = L.orc_hier_nlevels(self._h)
        self.levels = []
        for l in range(self.nlevels):
            nn, mm = C.c_uint32(), C.c_uint32()
            rp, cp, wp2, vp, vwp = (C.c_void_p() for _ in range(5))
            L.orc_hier_level(self._h, l, C.byref(nn), C.byref(mm), C.byref(rp), C.byref(cp),
                             C.byref(wp2), C.byref(vp), C.byref(vwp))
            n_, m_ = nn.value, mm.value

            def arr(p, k):
                if not p.value or k == 0:
                    return np.zeros(0, dtype=np.uint32)
                return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(k,)).copy()

            self.levels.append(dict(n=n_, m=m_, rowptr=arr(rp, n_ + 1), col=arr(cp, m_),
                                    w=arr(wp2, m_), vcmap=arr(vp, n_) if vp.value else None,
                                    vtx_w=arr(vwp, n_)))

    def eigensolve(self, pcg_state: int, pcg_inc: int = 1, normLap: int = 0,
                   tol: float = 1e-10, iter_cap: int = 0):
        """(evec, iters per level (0 = finest), new pcg state)."""
        st = C.c_uint64(pcg_state)
        n0 = self.levels[0]["n"]
        ev = np.empty(n0, dtype=np.float64)
        iters = np.zeros(self.nlevels, dtype=np.uint64)
        self._L.orc_eigensolve(self._h, C.byref(st), pcg_inc, normLap, tol, iter_cap, ev, iters)
        return ev, iters, st.value

    def eigensolve_trace(self, pcg_state: int, levels, pcg_inc: int = 1, normLap: int = 0,
                         tol: float = 1e-10):
        """eigensolve that also returns, for each level l in `levels`, the
        vector its power iteration starts from and the one it returns:
        (evec, iters, new pcg state, {l: (start, result)})."""
        st = C.c_uint64(pcg_state)
        ev = np.empty(self.levels[0]["n"], dtype=np.float64)
        iters = np.zeros(self.nlevels, dtype=np.uint64)
        trace = {l: (np.empty(self.levels[l]["n"]), np.empty(self.levels[l]["n"])) for l in levels}
        sp = (C.c_void_p * self.nlevels)(*[trace[l][0].ctypes.data if l in trace else None
                                          for l in range(self.nlevels)])
        rp = (C.c_void_p * self.nlevels)(*[trace[l][1].ctypes.data if l in trace else None
                                          for l in range(self.nlevels)])
        self._L.orc_eigensolve_trace(self._h, C.byref(st), pcg_inc, normLap, tol, 0, ev, iters,
                                     C.cast(sp, C.c_void_p), C.cast(rp, C.c_void_p))
        return ev, iters, st.value, trace

    def fm_eigensolve(self, seed: int, tol: float = 1e-10):
        """FM branch of the eigensolve: (finest 0/1 partition, last cut,
        fm_refine calls per level (0 = finest))."""
        part = np.zeros(self.levels[0]["n"] + 1, dtype=np.uint8)
        passes = np.zeros(self.nlevels, dtype=np.uint32)
        cut = self._L.orc_fm_eigensolve(self._h, seed, tol, part, passes)
        return part[:-1].copy(), int(cut), passes

    def __del__(self):
        try:
            self._L.orc_hier_free(self._h)
        except Exception:
            pass


def power_iter(rowptr, col, w, u, normLap: int = 0, tol: float = 1e-10, iter_cap: int = 0):
    """Runs the power iteration in place on a copy of u; (u, niter, max_reached, dot)."""
    L = core().L
    n = len(rowptr) - 1
    u = np.array(u, dtype=np.float64, copy=True)
    w, wp = _wptr(w)
    mr = C.c_int(0)
    d = C.c_double(0)
    it = L.orc_power_iter(n, rowptr, col, wp, u, normLap, tol, iter_cap, C.byref(mr), C.byref(d))
    return u, int(it), bool(mr.value), d.value


def eigenvalue_log(rowptr, col, w, u):
    """Final eigenvalue log (eigensolve_kokkos.h:60-100): (eigenvalue, min, max)."""
    L = core().L
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    out = np.zeros(3, dtype=np.float64)
    L.orc_eigenvalue_log(n, np.ascontiguousarray(rowptr, dtype=np.uint32), np.ascontiguousarray(col, dtype=np.uint32),
                         wp, np.ascontiguousarray(u, dtype=np.float64), out)
    return float(out[0]), float(out[1]), float(out[2])


def compute_partition(rowptr, col, evec, num_partitions: int = 2, imbalance: int = 0):
    L = core().L
    n = len(rowptr) - 1
    part = np.empty(n, dtype=np.uint32)
    ec = C.c_uint32(0)
    L.orc_compute_partition(n, rowptr, col, np.ascontiguousarray(evec, dtype=np.float64),
                            num_partitions, imbalance, part, C.byref(ec))
    return part, ec.value


def map_construct_v3(hn):
    """The HEC v3 map (heuristics.h:766-826) of heavy neighbours hn: (vcmap, nc)."""
    L = core().L
    hn = np.ascontiguousarray(hn, dtype=np.uint32)
    vc = np.empty(len(hn) + 1, dtype=np.uint32)
    nc = L.orc_map_construct_v3(len(hn), hn, vc)
    return vc[:len(hn)].copy(), int(nc)


def mis2(rowptr, col, seed: int):
    """(states (int32: 1 in the set, -1 out), rounds) of mis_2 (heuristics.h:8-100)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    st = np.zeros(n + 1, dtype=np.int32)
    r = L.orc_mis2(n, rowptr, col, seed, st)
    return st[:n].copy(), int(r)


def coarsen_mis2(rowptr, col, seed: int):
    """(vcmap, nc) of one MIS-2 level (sgp_coarsen_mis_2, heuristics.h:218-289)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    vc = np.empty(n + 1, dtype=np.uint32)
    nc = L.orc_coarsen_mis2(n, rowptr, col, seed, vc)
    return vc[:n].copy(), int(nc)


def coarsen_match(level: int, rowptr, col, w, seed: int):
    """(vcmap, nc, rounds) of one heavy-edge matching level (sgp_coarsen_match,
    heuristics.h:1584-1735, -DPUREMATCH); nc = None when the reference has no
    result (a row without entries, or only self-loops left)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vc = np.empty(n + 1, dtype=np.uint32)
    r = C.c_int(0)
    nc = L.orc_coarsen_match(level, n, rowptr, col, wp, seed, vc, C.byref(r))
    return vc[:n].copy(), (None if nc == 0xFFFFFFFF else int(nc)), r.value


def coarsen_mtmetis(level: int, rowptr, col, w, seed: int):
    """(vcmap, nc, rounds) of one level of the mtmetis_* binaries (-DMTMETIS):
    the matching, then the leaf, twin and relative passes
    (heuristics.h:1737-1888); nc = None when the reference has no result."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vc = np.empty(n + 1, dtype=np.uint32)
    r = C.c_int(0)
    nc = L.orc_coarsen_mtmetis(level, n, rowptr, col, wp, seed, vc, C.byref(r))
    return vc[:n].copy(), (None if nc == 0xFFFFFFFF else int(nc)), r.value


def murmur3_x86_32(data: bytes, seed: int = 0) -> int:
    """MurmurHash3_x86_32 of a byte string."""
    return int(core().L.orc_murmur3_x86_32(data, len(data), seed))


def binsort_u32(keys):
    """The order Kokkos::BinSort(BinOp1D(n, 0, 2^32-1)) with in-bin bubble
    passes gives u32 keys on one thread."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    out = np.empty(len(keys) + 1, dtype=np.uint32)
    core().L.orc_binsort_u32(len(keys), keys, out)
    return out[:len(keys)].copy()


def map_construct_v2(vperm, hn):
    """(vcmap, nc, rounds) of parallel_map_construct_v2 (heuristics.h:695-764)
    on one thread; nc = None if a round maps nothing."""
    L = core().L
    vperm = np.ascontiguousarray(vperm, dtype=np.uint32)
    hn = np.ascontiguousarray(hn, dtype=np.uint32)
    n = len(vperm)
    pos = np.empty(n + 1, dtype=np.uint32)
    pos[vperm] = np.arange(n, dtype=np.uint32)
    vc = np.empty(n + 1, dtype=np.uint32)
    r = C.c_int(0)
    nc = L.orc_map_construct_v2(n, vperm, hn, pos, vc, C.byref(r))
    return vc[:n].copy(), (None if nc == 0xFFFFFFFF else int(nc)), r.value


def coarsen_hec2(level: int, rowptr, col, w, seed: int):
    """(vcmap, nc, rounds) of one HEC level with the v2 map (the hec2_* binaries)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vc = np.empty(n + 1, dtype=np.uint32)
    r = C.c_int(0)
    nc = L.orc_coarsen_hec2(level, n, rowptr, col, wp, seed, vc, C.byref(r))
    return vc[:n].copy(), (None if nc == 0xFFFFFFFF else int(nc)), r.value


def coarsen_gosh2(rowptr, col, w, seed: int):
    """(vcmap, nc) of one GOSH v2 level (sgp_coarsen_GOSH_v2, heuristics.h:444-601);
    nc = None when a remaining vertex has no entries."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vc = np.empty(n + 1, dtype=np.uint32)
    nc = L.orc_coarsen_gosh2(n, rowptr, col, wp, seed, vc)
    return vc[:n].copy(), (None if nc == 0xFFFFFFFF else int(nc))


def map_construct_prefilled(vcmap, lst, hn, nc: int):
    """parallel_map_construct_prefilled (heuristics.h:378-442) on one thread:
    (new vcmap, nc)."""
    L = core().L
    vc = np.ascontiguousarray(vcmap, dtype=np.uint32).copy()
    lst = np.ascontiguousarray(lst, dtype=np.uint32)
    hn = np.ascontiguousarray(hn, dtype=np.uint32)
    out = L.orc_map_construct_prefilled(len(vc), lst if len(lst) else np.zeros(1, np.uint32), len(lst), hn, vc, nc)
    return vc, int(out)


def coarsen_hec_serial(level: int, rowptr, col, w, seed: int):
    """(vcmap, nc, vperm, hn) of one level of hec_spec_serial
    (sgp_coarsen_HEC_serial, heuristics.h:828-928); nc = None for a vertex
    without entries."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vc = np.empty(n + 1, dtype=np.uint32)
    vp = np.empty(n + 1, dtype=np.uint32)
    hn = np.empty(n + 1, dtype=np.uint32)
    nc = L.orc_coarsen_hec_serial(level, n, rowptr, col, wp, seed, vc, vp, hn)
    return vc[:n].copy(), (None if nc == 0xFFFFFFFF else int(nc)), vp[:n].copy(), hn[:n].copy()


def gosh_clusters(rowptr, col):
    """(states (int32: 1 centre, -1 next to one), rounds) of GOSH_clusters
    (heuristics.h:102-217)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    st = np.zeros(n + 1, dtype=np.int32)
    r = L.orc_gosh_clusters(n, rowptr, col, st)
    return st[:n].copy(), int(r)


def coarsen_gosh(rowptr, col):
    """(vcmap, nc) of one GOSH level (sgp_coarsen_GOSH, heuristics.h:312-377)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    vc = np.empty(n + 1, dtype=np.uint32)
    nc = L.orc_coarsen_gosh(n, rowptr, col, vc)
    return vc[:n].copy(), int(nc)


def coarsen_hec3(level: int, rowptr, col, w, seed: int):
    """(vcmap, nc, hn) of one HEC level with the v3 map."""
    L = core().L
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vc = np.empty(n + 1, dtype=np.uint32)
    hn = np.empty(n + 1, dtype=np.uint32)
    nc = L.orc_coarsen_hec3(level, n, np.ascontiguousarray(rowptr, dtype=np.uint32),
                            np.ascontiguousarray(col, dtype=np.uint32), wp, seed, vc, hn)
    return vc[:n].copy(), int(nc), hn[:n].copy()


def partition_graph(rowptr, col, w, hec_seed: int, pcg_state: int, pcg_inc: int = 1,
                    normLap: int = 0, tol: float = 1e-10, variant: int = 0, builder: int = 0):
    """Whole bisection; dict(part, edge_cut, levels_n, iters, evec, pcg_state).
    variant: 0 HEC, 1 heavy-edge matching (match_*), 2 matching with the
    mt-metis passes (mtmetis_*), 4 HEC with the v2 map (hec2_*), 7 GOSH v2
    (gosh2_*), 8 the serial HEC (hec_spec_serial), 3 HEC with the v3 map (the
    hec3_* binaries), 5 MIS-2 (mis_*), 6 GOSH (gosh_*); builder 1: the HASHMAP
    build (hec_spec_hashmap)."""
    L = core().L
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    st = C.c_uint64(pcg_state)
    part = np.empty(n, dtype=np.uint32)
    ec = C.c_uint32(0)
    lv = np.zeros(210, dtype=np.uint32)
    it = np.zeros(210, dtype=np.uint64)
    ev = np.empty(n, dtype=np.float64)
    nl = L.orc_partition_graph_variant(n, rowptr, col, wp, hec_seed, C.byref(st), pcg_inc, normLap, tol,
                                       part, C.byref(ec), lv, it, ev, variant | (builder << 8))
    return dict(part=part, edge_cut=ec.value, levels_n=lv[:nl].copy(), iters=it[:nl].copy(),
                evec=ev, pcg_state=st.value)


def _csr(rowptr, col):
    return (np.ascontiguousarray(rowptr, dtype=np.uint32), np.ascontiguousarray(col, dtype=np.uint32))


def fm_refine(rowptr, col, w, vtx_w, part):
    """One fm_refine call on a copy of part (0/1): (new part (uint8), min_cut,
    info dict(swaps, argmin, min_imb, start_cut, maxE))."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vw, vwp = _wptr(vtx_w)
    p = np.zeros(n + 1, dtype=np.uint8)
    p[:n] = np.asarray(part) != 0
    info = np.zeros(5, dtype=np.int64)
    cut = L.orc_fm_refine(n, rowptr, col, wp, vwp, p, info)
    return p[:n].copy(), int(cut), dict(swaps=int(info[0]), argmin=int(info[1]), min_imb=int(info[2]),
                                         start_cut=int(info[3]), maxE=int(info[4]))


def init_gggp(rowptr, col, w, vtx_w=None):
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vw, vwp = _wptr(vtx_w)
    p = np.zeros(n + 1, dtype=np.uint8)
    L.orc_init_gggp(n, rowptr, col, wp, vwp, p)
    return p[:n].copy()


def init_spectral(rowptr, col, w, vtx_w, seed: int, tol: float = 1e-10):
    """(0/1 partition, iterated vector)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    vw, vwp = _wptr(vtx_w)
    p = np.zeros(n + 1, dtype=np.uint8)
    x = np.zeros(n + 1, dtype=np.float64)
    L.orc_init_spectral(n, rowptr, col, wp, vwp, seed, tol, p, x)
    return p[:n].copy(), x[:n].copy()


def partition_graph_fm(rowptr, col, w, hec_seed: int, tol: float = 1e-10, variant: int = 0):
    """Whole bisection of the FM binary: dict(part, edge_cut, fm_part, passes, nlevels)."""
    L = core().L
    rowptr, col = _csr(rowptr, col)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    part = np.empty(n + 1, dtype=np.uint32)
    fmp = np.zeros(n + 1, dtype=np.uint8)
    ps = np.zeros(210, dtype=np.uint32)
    ec = C.c_uint32(0)
    nl = L.orc_partition_graph_fm_variant(n, rowptr, col, wp, hec_seed, tol, part, C.byref(ec), fmp, ps, variant)
    return dict(part=part[:n].copy(), edge_cut=ec.value, fm_part=fmp[:n].copy(), passes=ps[:nl].copy(),
                nlevels=nl)


# ---------------------------------------------------------------- reference (_ref)
class _Ref:
    def __init__(self) -> None:
        L = C.CDLL(REF_SO)
        self.L = L
        L.ref_pcg32.restype = C.c_uint32
        L.ref_pcg32.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
        for f in (L.ref_heap_dedupe, L.ref_hashmap_dedupe):
            f.restype = None
            f.argtypes = [u32p, u32p, u32p, u32p, C.POINTER(C.c_uint32)]
        L.ref_power_iter.restype = C.c_long
        L.ref_power_iter.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, f64p, C.c_int, C.c_int,
                                     C.c_double]
        L.ref_build_coarse.restype = C.c_long
        L.ref_build_coarse.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, u32p, C.c_uint32,
                                       u32p, u32p, u32p]
        L.ref_compute_partition.restype = C.c_int
        L.ref_compute_partition.argtypes = [C.c_uint32, u32p, u32p, f64p, C.c_int, C.c_int, u32p,
                                            C.POINTER(C.c_uint32)]
        L.ref_load_graph.restype = C.c_void_p
        L.ref_load_graph.argtypes = [C.c_char_p]
        L.ref_graph_info.restype = None
        L.ref_graph_info.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.ref_graph_free.restype = None
        L.ref_graph_free.argtypes = [C.c_void_p]
        L.ref_load_config.restype = C.c_int
        L.ref_load_config.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ref_use_partition.restype = C.c_int
        L.ref_use_partition.argtypes = [C.c_uint32, u32p, u32p, u32p, C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p)]
        L.ref_eigenvalue_log.restype = C.c_int
        L.ref_eigenvalue_log.argtypes = [C.c_uint32, u32p, u32p, f64p, f64p]
        L.ref_edit_distance.restype = C.c_int
        L.ref_edit_distance.argtypes = [u32p, u32p, C.c_uint32, C.POINTER(C.c_uint32)]

    def graph_arrays(self, h):
        n, ne = C.c_uint32(), C.c_uint32()
        rp, cp = C.c_void_p(), C.c_void_p()
        self.L.ref_graph_info(h, C.byref(n), C.byref(ne), C.byref(rp), C.byref(cp))
        nn = n.value
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint32)), shape=(nn + 1,)).copy()
        m = int(rowptr[-1])
        col = (np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_uint32)), shape=(m,)).copy()
               if m else np.zeros(0, dtype=np.uint32))
        return rowptr, col

    def pcg32_stream(self, state: int, count: int, inc: int = 1):
        st = C.c_uint64(state)
        out = np.empty(count, dtype=np.uint32)
        for i in range(count):
            out[i] = self.L.ref_pcg32(C.byref(st), inc)
        return out, st.value

    def power_iter(self, rowptr, col, w, u, normLap=0, final_level=0, tol=1e-10):
        u = np.array(u, dtype=np.float64, copy=True)
        w, wp = _wptr(w)
        it = self.L.ref_power_iter(len(rowptr) - 1, rowptr, col, wp, u, normLap, final_level, tol)
        return u, int(it)

    def build_coarse(self, rowptr, col, w, vcmap, nc):
        m = int(rowptr[-1])
        w, wp = _wptr(w)
        orow = np.zeros(nc + 1, dtype=np.uint32)
        ocol = np.zeros(max(m, 1), dtype=np.uint32)
        ow = np.zeros(max(m, 1), dtype=np.uint32)
        mc = self.L.ref_build_coarse(len(rowptr) - 1, rowptr, col, wp,
                                     np.array(vcmap, dtype=np.uint32), nc, orow, ocol, ow)
        return orow, ocol[:mc], ow[:mc]

    def compute_partition(self, rowptr, col, evec, num_partitions=2, imbalance=0):
        n = len(rowptr) - 1
        part = np.empty(n, dtype=np.uint32)
        ec = C.c_uint32(0)
        rc = self.L.ref_compute_partition(n, rowptr, col, np.ascontiguousarray(evec, np.float64),
                                          num_partitions, imbalance, part, C.byref(ec))
        return rc, part, ec.value


_ref = None


def ref():
    """The reference wrapper, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        _ref = _Ref()
    return _ref


# ---------------------------------------------------------------- OpenMP form (CPU baseline)
OMP_SO = os.path.join(HERE, "libsgpar_omp.so")
_omp = None


def omp_available() -> bool:
    return os.path.exists(OMP_SO)


def _omp_lib():
    global _omp
    if _omp is None:
        L = C.CDLL(OMP_SO)
        L.omp_threads.restype = C.c_int
        L.omp_threads.argtypes = []
        L.omp_partition_graph.restype = C.c_int
        L.omp_partition_graph.argtypes = [C.c_uint32, u32p, u32p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                          C.c_uint64, C.c_double, u32p, C.POINTER(C.c_uint32), u32p, u64p]
        _omp = L
    return _omp


def omp_threads() -> int:
    return int(_omp_lib().omp_threads())


def omp_partition_graph(rowptr, col, w, hec_seed: int, pcg_state: int, pcg_inc: int = 1, tol: float = 1e-10):
    """Whole bisection by the OpenMP restatement (sgpar_omp.c), on all
    threads OpenMP gives; dict(part, edge_cut, levels_n, iters, pcg_state)."""
    L = _omp_lib()
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = len(rowptr) - 1
    w, wp = _wptr(w)
    st = C.c_uint64(pcg_state)
    part = np.empty(n, dtype=np.uint32)
    ec = C.c_uint32(0)
    lv = np.zeros(210, dtype=np.uint32)
    it = np.zeros(210, dtype=np.uint64)
    nl = L.omp_partition_graph(n, rowptr, col, wp, hec_seed, C.byref(st), pcg_inc, tol, part, C.byref(ec), lv, it)
    return dict(part=part, edge_cut=ec.value, levels_n=lv[:nl].copy(), iters=it[:nl].copy(), pcg_state=st.value)


def host_cpu_info() -> str:
    """CPU model, sockets, physical cores and logical CPUs of this host (lscpu)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except Exception:
        return "lscpu unavailable"
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        cores = int(kv.get("Core(s) per socket", "0")) * int(kv.get("Socket(s)", "0"))
    except ValueError:
        cores = 0
    return (f"{kv.get('Model name', '?')}, {kv.get('Socket(s)', '?')} socket(s), {cores} physical cores, "
            f"{kv.get('CPU(s)', '?')} logical CPUs, {os.cpu_count()} visible to this process")
