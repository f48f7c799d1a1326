"""Benchmark: multilevel spectral bisection of an R-MAT-24 graph on one
MI355X, plain FP64 and mixed precision.

Workload (the 4-way configuration's graph): R-MAT scale 24 (16,777,216
vertices), edge factor 16, (a, b, c, d) = (0.57, 0.19, 0.19, 0.05), seed 4,
drawn on the GPU; symmetrised, self-loops dropped, duplicates merged with
summed u32 weights, largest connected component kept (8,866,537 vertices,
520,765,078 stored entries).  One step = one whole bisection (HEC coarsening
to <= 50 vertices, P^T A P builds, multilevel FP64 power iteration,
sort-and-split) with the graph already resident in HBM and the partition
left in HBM; HEC and PCG seeds 4.

Prints one JSON line.  `value` is milliseconds per FP64 bisection (lower is
better).  Beside it:
- `config.end_to_end_ms`: the same bisection from host CSR arrays to a host
  partition (upload, bisection, download), what a caller of the reference's
  sgp_partition_graph gets;
- `mixed`: the bisection in mixed precision (FP32 vectors on the levels above
  the finest with at least 2^20 entries; FP64 sums, deflation, stopping test
  and finest level) beside the FP64 one, at the cut each ends with;
- `four_way`: the 4-way partition by recursive bisection (the configuration
  itself), FP64 and mixed;
- `roofline`: the streaming kernel of the power-iteration step
  (k_power_stream) on the level with the most eigensolve time, timed with HIP
  events on the library stream over 50 launches; algorithmic bytes are
  col + w + x of that level (per launch).  `traffic` is the HBM traffic per
  launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by the same counters
  on a copy of known size (profiles/k_power_stream_<workload>_L<l>_traffic.txt),
  or null when that summary is absent;
- `cpu_baseline`: the OpenMP restatement of the path (oracle/sgpar_omp.c) on
  all host cores, whole bisections of the same graph, median of the runs;
- `extras`: R-MAT-20 (seed 1) FP64 and mixed with a per-level precision
  report, the 128^3 weighted 3D grid FP64 and mixed, and the 512-RGG batch
  (mi_bisect_batch) against one call per graph.

    python bench.py [--steps K] [--warmup W]
    python bench.py --scale 20                # R-MAT-20 as the main line
    python bench.py --roofline-only           # only the roofline measurement (for rocprofv3 runs)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_file(workload: str, level: int) -> str:
    return os.path.join(ROOT, "profiles", f"k_power_stream_{workload}_L{level}_traffic.txt")


def measured_traffic(workload: str, level: int):
    """HBM bytes per launch of k_power_stream on that level from the
    committed rocprofv3 counter summary (line 'traffic_bytes_per_launch: X')."""
    try:
        with open(traffic_file(workload, level)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def merge_rate(st: dict) -> dict:
    """Algorithmic bytes of the coarse builds (P^T A P with duplicate merge)
    over their time: per level l -> l+1, read the fine CSR 4(n+1) + 4m + 4m
    (weights; a unit-weight finest level stores none) and vcmap 4n, write
    the coarse CSR 4(nc+1) + 8 m_c."""
    ns, ms = st["level_n"], st["level_m"]
    b = 0.0
    for l in range(len(ns) - 1):
        b += 4 * (ns[l] + 1) + 4 * ms[l] + (4 * ms[l] if (l > 0 or st.get("weighted")) else 0) + 4 * ns[l]
        b += 4 * (ns[l + 1] + 1) + 8 * ms[l + 1]
    t = st["build_ms"]
    return {"alg_bytes": b, "ms": round(t, 3), "gbps": round(b / (t * 1e-3) / 1e9, 1) if t > 0 else None,
            "frac": round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4) if t > 0 else None}


class Runner:
    """Bisections of one graph through the C ABI, partition left in HBM."""

    def __init__(self, mi, ctx, g, seed: int):
        self.mi, self.ctx, self.g, self.seed = mi, ctx, g, seed
        self.L = ctx._L
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.d_part = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.d_part), 4 * g.n) == 0

    def close(self):
        self.hip.hipFree(self.d_part)

    def bisect(self):
        mi, ctx = self.mi, self.ctx
        st = mi.Stats()
        pcg = C.c_uint64(self.seed)
        ec = C.c_uint32()
        cfg = mi.Config(4, 0, 0, 1, 1e-10, 2)
        ctx.check(self.L.mi_partition_graph(ctx.h, self.g.h, C.byref(cfg), 0, self.seed, C.byref(pcg), 1,
                                            self.d_part, 1, C.byref(ec), C.byref(st)), "mi_partition_graph")
        return st

    def four_way(self):
        mi, ctx = self.mi, self.ctx
        pcg = C.c_uint64(self.seed)
        cuts = np.zeros(3, dtype=np.uint32)
        ec = C.c_uint32()
        cfg = mi.Config(4, 0, 0, 1, 1e-10, 4)
        ctx.check(self.L.mi_partition_4way(ctx.h, self.g.h, C.byref(cfg), 0, self.seed, C.byref(pcg), 1,
                                           self.d_part, 1, cuts, C.byref(ec)), "mi_partition_4way")
        return int(ec.value), [int(x) for x in cuts]

    def timed(self, fn, steps: int, warmup: int, tag: str):
        for i in range(warmup):
            r = fn()
            if hasattr(r, "total_ms"):
                log(f"{tag} warmup {i}: {r.total_ms:.2f} ms, cut {r.edge_cut}, levels {r.nlevels}")
        self.hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = fn()
        self.hip.hipDeviceSynchronize()
        return (time.perf_counter() - t0) * 1000.0 / steps, r


def mixed_summary(ms64: float, st64: dict, ms32: float, st32: dict) -> dict:
    """Mixed-precision bisection beside the plain FP64 one on the same graph."""
    return {"ms": round(ms32, 3), "fp64_ms": round(ms64, 3), "speedup": round(ms64 / ms32, 4),
            "edge_cut": st32["edge_cut"], "fp64_edge_cut": st64["edge_cut"],
            "cut_change_pct": round(100.0 * (st32["edge_cut"] - st64["edge_cut"]) / max(st64["edge_cut"], 1), 4),
            "fp32_levels": st32["mixed_levels"], "level_iters": st32["level_iters"],
            "fp64_level_iters": st64["level_iters"],
            "level_eigen_ms": [round(x, 3) for x in st32["level_eigen_ms"]],
            "fp64_level_eigen_ms": [round(x, 3) for x in st64["level_eigen_ms"]],
            "phases_ms": {k: round(st32[k], 3) for k in ("coarsen_ms", "eigen_ms", "partition_ms")}}


def fiedler_quality(rp, col, w, u) -> dict:
    """Rayleigh quotient and relative Fiedler residual ||L u - lam u|| /
    ||lam u|| (L = D - A, u deflated) of a vector, on the host."""
    import scipy.sparse as sp
    n = len(rp) - 1
    A = sp.csr_matrix((w.astype(np.float64), col, rp), shape=(n, n))
    Lap = sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A
    u = u - u.mean()
    lam = float(u @ (Lap @ u) / (u @ u))
    return {"rayleigh": lam, "residual": float(np.linalg.norm(Lap @ u - lam * u) / (abs(lam) * np.linalg.norm(u)))}


def mixed_levels_report(ctx, g, seed: int) -> list:
    """Per level that runs in FP32 (at least 2^20 entries, above the
    finest): the level's power iteration in mixed precision and in FP64 from
    the same start vector (FP64 on the device, which the tests hold to the
    oracle), with iterations, the angle between the two vectors, and each
    one's Rayleigh quotient and relative Fiedler residual."""
    H = g.hierarchy(seed)
    out = []
    for l in range(1, H.nlevels):
        lv = H.level(l)
        if lv["m"] < (1 << 20):
            continue
        n = lv["n"]
        lg = ctx.graph(lv["rowptr"], lv["col"], lv["w"])
        u0 = np.random.default_rng(l).random(n)
        ctx.set_precision(False)
        u64, it64, _ = lg.power_iter(u0)
        ctx.set_precision(True, 1)
        u32, it32, _ = lg.power_iter(u0)
        ctx.set_precision(False)
        lg.close()
        q64 = fiedler_quality(lv["rowptr"], lv["col"], lv["w"], u64)
        q32 = fiedler_quality(lv["rowptr"], lv["col"], lv["w"], u32)
        out.append({"level": l, "n": n, "m": lv["m"], "iters_fp64": it64, "iters_mixed": it32,
                    "cos": round(float(abs(u64 @ u32) / (np.linalg.norm(u64) * np.linalg.norm(u32))), 12),
                    "rayleigh_fp64": q64["rayleigh"], "rayleigh_mixed": q32["rayleigh"],
                    "residual_fp64": round(q64["residual"], 6), "residual_mixed": round(q32["residual"], 6)})
    H.close()
    return out


def bisect_pair(mi, ctx, g, seed: int, steps: int, warmup: int, tag: str):
    """(FP64 ms, FP64 stats, mixed ms, mixed stats) of whole bisections."""
    R = Runner(mi, ctx, g, seed)
    try:
        ms64, st64 = R.timed(R.bisect, steps, warmup, f"{tag} fp64")
        ctx.set_precision(True)
        ms32, st32 = R.timed(R.bisect, steps, warmup, f"{tag} mixed")
    finally:
        ctx.set_precision(False)
        R.close()
    return ms64, st64.as_dict(), ms32, st32.as_dict()


def extra_rmat20(mi, ctx, steps: int) -> dict:
    from ipdps_2021_graph_coarsening_amd import graphs
    t = time.perf_counter()
    rp, col, w = graphs.rmat(20, 16, 0.57, 0.19, 0.19, seed=1)
    log(f"R-MAT-20: generated in {time.perf_counter() - t:.1f} s")
    g = ctx.graph(rp, col, w)
    ms64, st64, ms32, st32 = bisect_pair(mi, ctx, g, 1, steps, 2, "rmat20")
    out = mixed_summary(ms64, st64, ms32, st32)
    out["workload"] = f"rmat20-ef16-lcc-weighted (n={len(rp) - 1}, m={int(rp[-1])}, seed 1)"
    out["level_n"] = st64["level_n"]
    out["merge"] = merge_rate(dict(st64, weighted=True))
    out["per_level"] = mixed_levels_report(ctx, g, 1)
    g.close()
    return out


def extra_grid3d(mi, ctx, steps: int) -> dict:
    """The 128^3 7-point grid with weights U{1..16} (seed 2): FP64 and mixed."""
    from ipdps_2021_graph_coarsening_amd import graphs
    rp, col, w = graphs.grid3d(128, seed=2)
    g = ctx.graph(rp, col, w)
    ms64, st64, ms32, st32 = bisect_pair(mi, ctx, g, 2, steps, 1, "grid3d")
    g.close()
    out = mixed_summary(ms64, st64, ms32, st32)
    out["workload"] = f"grid3d-128 (n={len(rp) - 1}, m={int(rp[-1])}, weights U1..16, seed 2)"
    out["merge"] = merge_rate(dict(st64, weighted=True))
    return out


FM_GRAPHS = {}


def _fm_graph(name: str):
    from ipdps_2021_graph_coarsening_amd import graphs
    if name not in FM_GRAPHS:
        if name == "grid1024":
            rp, col = graphs.grid2d(1024)[:2]
            FM_GRAPHS[name] = (rp, col, None, "grid2d-1024 (n=1048576, 5-point, unit weights)", 1)
        else:
            rp, col, w = graphs.rmat(20, 16, 0.57, 0.19, 0.19, seed=1)
            FM_GRAPHS[name] = (rp, col, w, f"rmat20-ef16-lcc-weighted (n={len(rp) - 1}, m={int(rp[-1])}, seed 1)", 1)
    return FM_GRAPHS[name]


def extra_fm(mi, ctx, steps: int) -> dict:
    """Bisections with FM refinement (the reference's hec_fm binary,
    mi_ctx_set_refinement): whole-bisection time, cut, fm_refine calls and
    time per level."""
    out = {}
    ctx.set_refinement(True)
    try:
        for name, reps in (("grid1024", steps), ("rmat20", 1)):
            rp, col, w, desc, seed = _fm_graph(name)
            g = ctx.graph(rp, col, w)
            R = Runner(mi, ctx, g, seed)
            ms, st = R.timed(R.bisect, reps, 1, f"{name} fm")
            R.close()
            g.close()
            d = st.as_dict()
            L = d["nlevels"]
            out[name] = {"workload": desc, "ms": round(ms, 3), "edge_cut": d["edge_cut"],
                         "coarsen_ms": round(d["coarsen_ms"], 3), "fm_ms": round(d["eigen_ms"], 3),
                         "split_ms": round(d["partition_ms"], 3), "level_n": d["level_n"][:L],
                         "fm_calls_per_level": d["level_iters"][:L],
                         "level_fm_ms": [round(x, 3) for x in d["level_eigen_ms"][:L]]}
            log(f"FM bisection {name}: {ms:.2f} ms, cut {d['edge_cut']}")
    finally:
        ctx.set_refinement(False)
    return out


def cpu_baseline_fm() -> dict:
    """The oracle's FM bisection (scalar C, one core) on the graphs of
    extra_fm: the reference hec_fm binary's algorithm, every fm_refine
    on host arrays."""
    import oracle
    out = {}
    for name in ("grid1024", "rmat20"):
        rp, col, w, desc, seed = _fm_graph(name)
        t = time.perf_counter()
        r = oracle.partition_graph_fm(rp, col, w, seed)
        ms = (time.perf_counter() - t) * 1000.0
        out[name] = {"value": round(ms, 1), "unit": "ms", "cores": 1, "kind": "port", "edge_cut": int(r["edge_cut"]),
                     "sample": f"oracle/sgpar_oracle.c orc_partition_graph_fm on {desc}, one run"}
        log(f"cpu FM baseline {name}: {ms:.0f} ms, cut {r['edge_cut']}")
    return out


def _rgg_job(args):
    from ipdps_2021_graph_coarsening_amd import graphs
    n, seed = args
    return graphs.rgg(int(n), seed=seed)


def extra_rgg_batch(mi, ctx, reps: int = 2) -> dict:
    """512 random geometric graphs (unit square, seed 3, n log-uniform in
    [1000, 200000], radius for mean degree 10, largest component): the
    batched bisection (mi_bisect_batch) against one call per graph, both
    timed on the host clock; the results are identical graph by graph."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    rng = np.random.default_rng(3)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(2e5), 512)).astype(int)
    t = time.perf_counter()
    with ProcessPoolExecutor(8, mp_context=mp.get_context("spawn")) as ex:
        host = list(ex.map(_rgg_job, [(int(n), 3 + i) for i, n in enumerate(ns)], chunksize=8))
    G = [ctx.graph(rp, col) for rp, col, _ in host]
    log(f"512 RGGs: {sum(g.n for g in G)} vertices, {sum(g.m for g in G)} entries, "
        f"generated in {time.perf_counter() - t:.1f} s")
    seeds, states = [1] * len(G), [1] * len(G)
    best, ph = None, None
    for _ in range(reps):
        t = time.perf_counter()
        res, p = mi.bisect_batch(ctx, G, hec_seeds=seeds, pcg_states=states)
        dt = (time.perf_counter() - t) * 1000.0
        if best is None or dt < best:
            best, ph = dt, p
    t = time.perf_counter()
    cuts = [mi.partition_graph(ctx, g, hec_seed=1, pcg_state=1).edge_cut for g in G]
    seq = (time.perf_counter() - t) * 1000.0
    same = cuts == [r.edge_cut for r in res]
    for g in G:
        g.close()
    return {"workload": "512 RGGs, n log-uniform in [1e3, 2e5], seed 3", "graphs": len(G),
            "batch_ms": round(best, 1), "one_call_per_graph_ms": round(seq, 1), "speedup": round(seq / best, 3),
            "batch_phases_ms": {"coarsen": round(ph[0], 1), "eigen": round(ph[1], 1), "split": round(ph[2], 1)},
            "cuts_equal_one_call_per_graph": same}


def end_to_end(mi, ctx, rp, col, w, seed: int, reps: int = 3) -> dict:
    """Host CSR in, host partition out: upload (mi_graph_create), the
    bisection and the download of the partition, median of reps."""
    ts, up = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        g = ctx.graph(rp, col, w)
        t1 = time.perf_counter()
        r = mi.partition_graph(ctx, g, hec_seed=seed, pcg_state=seed)
        t2 = time.perf_counter()
        g.close()
        up.append((t1 - t0) * 1000.0)
        ts.append((t2 - t0) * 1000.0)
        del r
    return {"ms": round(statistics.median(ts), 3), "upload_ms": round(statistics.median(up), 3),
            "upload_gbps": round((4.0 * (len(rp) + len(col) + (len(w) if w is not None else 0)))
                                 / (statistics.median(up) * 1e-3) / 1e9, 2)}


def roofline_of(mi, ctx, g, seed: int, lvl: int, workload: str, name: str) -> dict:
    if lvl == 0:
        gl = g
    else:
        H = g.hierarchy(seed)
        lv = H.level(lvl)
        H.close()
        gl = ctx.graph(lv["rowptr"], lv["col"], lv["w"])
        del lv
    b = gl.bench_step(50)
    achieved = b["stream_bytes"] / (b["stream_ms"] * 1e-3) / 1e9
    log(f"level {lvl} (n={gl.n}, m={gl.m}): step {b['step_ms'] * 1000:.1f} us, "
        f"k_power_stream {b['stream_ms'] * 1000:.1f} us/launch, {b['stream_bytes'] / 1e6:.1f} MB algorithmic, "
        f"{achieved:.0f} GB/s")
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(workload, lvl),
           "kernel": ((f"k_power_stream_lds ({b['column_blocks']} LDS column blocks)" if b["column_blocks"]
                       else "k_power_stream (gathers from global memory)")
                      + f", FP64 power-iteration step, {name} level {lvl}"),
           "level": lvl, "level_n": gl.n, "level_m": gl.m,
           "alg_bytes_per_launch": b["stream_bytes"], "impl_bytes_per_launch": b["impl_bytes"],
           "launch_us": round(b["stream_ms"] * 1000, 2),
           "step_us": round(b["step_ms"] * 1000, 2),
           "step_gbps": round(b["step_bytes"] / (b["step_ms"] * 1e-3) / 1e9, 1)}
    if gl is not g:
        gl.close()
    return out


def cpu_baseline(rp, col, w, seed: int, reps: int, budget_s: float) -> dict | None:
    """The OpenMP restatement of the path (oracle/sgpar_omp.c): whole
    bisections of the same graph on all host cores, median of the runs
    (fewer than `reps` when one run exceeds budget_s / reps)."""
    import oracle
    if not oracle.omp_available():
        log("cpu baseline: oracle/libsgpar_omp.so not built")
        return None
    threads = oracle.omp_threads()
    ts = []
    for i in range(reps):
        t = time.perf_counter()
        r = oracle.omp_partition_graph(rp, col, w, seed, seed)
        ts.append((time.perf_counter() - t) * 1000.0)
        log(f"cpu baseline run {i}: {ts[-1]:.0f} ms, cut {r['edge_cut']}, levels {list(r['levels_n'])}")
        if sum(ts) / 1000.0 > budget_s:
            break
    info = oracle.host_cpu_info()
    return {"value": round(statistics.median(ts), 1), "unit": "ms", "cores": threads, "kind": "port",
            "runs": len(ts), "edge_cut": int(r["edge_cut"]),
            "sample": (f"oracle/sgpar_omp.c (OpenMP restatement of the reference path: HEC coarsening, "
                       f"P^T A P build, FP64 power iteration, sort-and-split) on the whole bisection of the "
                       f"same graph, {threads} threads, median of {len(ts)} run(s); host: {info}")}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24,
                    help="R-MAT scale of the main line: 24 (seed 4, drawn on the GPU, the default) or "
                         "<= 21 (seed 1, drawn on the host)")
    ap.add_argument("--graph-cache", default=None,
                    help="directory that keeps the GPU-drawn R-MAT graph between runs (drawn when absent)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=60.0, help="seconds of CPU baseline at most")
    ap.add_argument("--no-extras", action="store_true", help="no R-MAT-20 / 3D grid / batch extras")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched measurement")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline measurement of the dominant kernel, plus a 256 MiB device copy "
                         "that calibrates rocprofv3's byte counters")
    ap.add_argument("--bisections-only", action="store_true",
                    help="only the timed FP64 bisections (for kernel traces of the whole bisection)")
    ap.add_argument("--extra-only", choices=["rmat20", "grid3d", "batch", "fm"], default=None,
                    help="only one of the extra workloads (for kernel traces of it), printed as JSON")
    ap.add_argument("--bucketed-build", type=int, default=None, metavar="MIN_ENTRIES",
                    help="coarse builds of levels with at least MIN_ENTRIES entries take the bucketed form "
                         "(default: the radix form everywhere)")
    ap.add_argument("--roofline-level", type=int, default=-1,
                    help="level for the roofline (default: the level with the most eigensolve time; "
                         "1 with --roofline-only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent graphs would shard one per GPU, with no collectives).", file=sys.stderr)
        return 2

    import ipdps_2021_graph_coarsening_amd as mi
    from ipdps_2021_graph_coarsening_amd import graphs

    if args.extra_only:
        ctx = mi.Context(0)
        if args.bucketed_build is not None:
            ctx.set_bucketed_build(args.bucketed_build)
        ctx.graph(np.array([0, 1, 2], dtype=np.uint32), np.array([1, 0], dtype=np.uint32)).close()
        steps = max(1, min(args.steps, 5))
        fn = {"rmat20": lambda: extra_rmat20(mi, ctx, steps), "grid3d": lambda: extra_grid3d(mi, ctx, steps),
              "batch": lambda: extra_rgg_batch(mi, ctx), "fm": lambda: extra_fm(mi, ctx, steps)}[args.extra_only]
        res = {"extra": args.extra_only, "result": fn()}
        ctx.close()
        if args.extra_only == "fm" and not args.no_cpu_baseline:
            res["cpu_baseline"] = cpu_baseline_fm()
        print(json.dumps(res), flush=True)
        return 0

    t = time.perf_counter()
    if args.scale >= 22:
        seed = 4
        rp, col, w = graphs.rmat_device(args.scale, 16, 0.57, 0.19, 0.19, seed=seed, cache_dir=args.graph_cache)
    else:
        seed = 1
        rp, col, w = graphs.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=seed)
    name = f"R-MAT-{args.scale}"
    workload = f"rmat{args.scale}"
    n, m = len(rp) - 1, int(rp[-1])
    log(f"{name}: n={n} m={m} generated in {time.perf_counter() - t:.1f} s")

    ctx = mi.Context(0)
    if args.bucketed_build is not None:
        ctx.set_bucketed_build(args.bucketed_build)
    ctx.graph(np.array([0, 1, 2], dtype=np.uint32), np.array([1, 0], dtype=np.uint32)).close()  # first-use costs
    t = time.perf_counter()
    g = ctx.graph(rp, col, w)  # host -> HBM upload of the CSR (synchronous)
    upload_ms = (time.perf_counter() - t) * 1000.0
    log(f"graph upload (host arrays -> HBM): {upload_ms:.2f} ms")

    if args.roofline_only:
        lvl = 1 if args.roofline_level < 0 else args.roofline_level
        roof = roofline_of(mi, ctx, g, seed, lvl, workload, name)
        # calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE: a device copy of known size
        hip = C.CDLL("libamdhip64.so")
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipFree.argtypes = [C.c_void_p]
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        nb = 256 << 20
        a_, b_ = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(a_), nb) == 0 and hip.hipMalloc(C.byref(b_), nb) == 0
        for _ in range(3):
            hip.hipMemcpy(b_, a_, nb, 3)
        hip.hipDeviceSynchronize()
        hip.hipFree(a_)
        hip.hipFree(b_)
        print(json.dumps({"roofline": roof, "calibration_copy_bytes": nb}), flush=True)
        g.close()
        ctx.close()
        return 0

    R = Runner(mi, ctx, g, seed)
    ms, st = R.timed(R.bisect, args.steps, args.warmup, f"{name} fp64")
    stats = st.as_dict()
    log(f"timed {args.steps} FP64 bisections: {ms:.2f} ms, cut {stats['edge_cut']}, "
        f"levels {stats['level_n']}, iterations {stats['level_iters']}")
    if args.bisections_only:
        print(json.dumps({"workload": workload, "ms_per_bisection": round(ms, 3), "stats": stats}), flush=True)
        R.close()
        g.close()
        ctx.close()
        return 0
    ctx.set_precision(True)
    ms32, st32 = R.timed(R.bisect, args.steps, args.warmup, f"{name} mixed")
    ctx.set_precision(False)
    mixed = mixed_summary(ms, stats, ms32, st32.as_dict())
    log(f"mixed precision: {ms32:.2f} ms, cut {st32.edge_cut}, iterations {list(st32.level_iters[:st32.nlevels])}")
    steps4 = max(1, min(args.steps, 5))
    ms4, (cut4, cuts4) = R.timed(R.four_way, steps4, 1, f"{name} 4-way fp64")
    ctx.set_precision(True)
    ms4m, (cut4m, cuts4m) = R.timed(R.four_way, steps4, 1, f"{name} 4-way mixed")
    ctx.set_precision(False)
    four = {"fp64_ms": round(ms4, 3), "fp64_edge_cut": cut4, "fp64_cuts": cuts4,
            "mixed_ms": round(ms4m, 3), "mixed_edge_cut": cut4m, "mixed_cuts": cuts4m,
            "speedup": round(ms4 / ms4m, 4),
            "cut_change_pct": round(100.0 * (cut4m - cut4) / max(cut4, 1), 4)}
    log(f"4-way: FP64 {ms4:.2f} ms cut {cut4}, mixed {ms4m:.2f} ms cut {cut4m}")
    R.close()

    # finest-level Fiedler residual of both precisions' final vectors
    H = g.hierarchy(seed)
    ev64, _, _ = H.fiedler(seed)
    ctx.set_precision(True)
    ev32, _, _ = H.fiedler(seed)
    ctx.set_precision(False)
    H.close()
    q64 = fiedler_quality(rp, col, w, ev64)
    q32 = fiedler_quality(rp, col, w, ev32)
    mixed["finest"] = {"cos": round(float(abs(ev64 @ ev32) / (np.linalg.norm(ev64) * np.linalg.norm(ev32))), 12),
                       "rayleigh_fp64": q64["rayleigh"], "rayleigh_mixed": q32["rayleigh"],
                       "residual_fp64": round(q64["residual"], 6), "residual_mixed": round(q32["residual"], 6)}
    del ev64, ev32

    e2e = end_to_end(mi, ctx, rp, col, w, seed)
    log(f"end to end (host CSR -> host partition): {e2e['ms']:.1f} ms, upload {e2e['upload_ms']:.1f} ms")

    lvl = args.roofline_level if args.roofline_level >= 0 else int(np.argmax(stats["level_eigen_ms"]))
    roof = roofline_of(mi, ctx, g, seed, lvl, workload, name)

    extras = None
    if not args.no_extras:
        extras = {}
        if args.scale != 20:
            extras["rmat20"] = extra_rmat20(mi, ctx, max(1, min(args.steps, 5)))
            log(f"R-MAT-20: FP64 {extras['rmat20']['fp64_ms']:.2f} ms, mixed {extras['rmat20']['ms']:.2f} ms")
        extras["grid3d"] = extra_grid3d(mi, ctx, max(1, min(args.steps, 5)))
        log(f"grid3d-128: FP64 {extras['grid3d']['fp64_ms']:.2f} ms, mixed {extras['grid3d']['ms']:.2f} ms")
        extras["fm"] = extra_fm(mi, ctx, max(1, min(args.steps, 3)))
        if not args.no_batch:
            extras["batched_rgg512"] = extra_rgg_batch(mi, ctx)
            b = extras["batched_rgg512"]
            log(f"512 RGGs: batch {b['batch_ms']} ms, one call per graph {b['one_call_per_graph_ms']} ms")

    g.close()
    ctx.close()

    cpu = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        cpu = cpu_baseline(rp, col, w, seed, args.cpu_reps, args.cpu_budget)
        if cpu:
            log(f"cpu baseline {cpu['value']} ms ({time.perf_counter() - t:.1f} s)")
        if cpu is not None and extras is not None and "fm" in extras:
            cpu["fm"] = cpu_baseline_fm()

    out = {
        "metric": "bisection ms + SpMV HBM GB/s: FP32-mixed/batched vs plain FP64 at equal cut",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{workload}-ef16-lcc-weighted plain-fp64 bisection",
                   "n": n, "m": m, "seed": seed, "levels": stats["nlevels"],
                   "upload_ms": round(upload_ms, 3),
                   "end_to_end": e2e,
                   "level_n": stats["level_n"], "level_m": stats["level_m"],
                   "level_iters": stats["level_iters"],
                   "edge_cut": stats["edge_cut"],
                   "phases_ms": {k: round(stats[k], 3) for k in
                                 ("coarsen_ms", "eigen_ms", "partition_ms", "permute_ms", "heavy_ms",
                                  "map_ms", "build_ms")},
                   "level_eigen_ms": [round(x, 3) for x in stats["level_eigen_ms"]],
                   "level_us_per_iteration": [round(1000.0 * t / max(i, 1), 2) for t, i in
                                              zip(stats["level_eigen_ms"], stats["level_iters"])],
                   "merge": merge_rate(dict(stats, weighted=True))},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "four_way": four,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
