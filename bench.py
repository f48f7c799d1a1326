"""Benchmark: plain FP64 multilevel spectral bisection of an R-MAT graph on
one MI355X.

Workload: R-MAT scale 20 (1,048,576 vertices), edge factor 16,
(a, b, c, d) = (0.57, 0.19, 0.19, 0.05), seed 1; symmetrised, self-loops
dropped, duplicates merged with summed u32 weights, largest connected
component kept.  One step = one whole bisection (HEC coarsening to <= 50
vertices, P^T A P builds, multilevel FP64 power iteration, sort-and-split)
with the graph already resident in HBM and the partition left in HBM.

Prints one JSON line.  `value` is milliseconds per bisection (lower is
better).  `roofline` is the kernel that dominates the bisection: the
streaming kernel of the FP64 power-iteration step (k_power_stream) on the
level with the most eigensolve time (level 1 of R-MAT-20, which takes about
400 iterations), timed with HIP events on the library stream over 50
launches; its algorithmic bytes are col + w + x + wdeg + y of that level.
`traffic` is the HBM traffic per launch of that kernel measured with
rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by the same counters on a copy of
known size (profiles/k_power_stream_rmat20_L1_traffic.txt), or null when
that summary is absent.  `cpu_baseline` times the CPU oracle (a scalar
restatement of the reference's Serial path) on a bounded sample and scales
it to one whole bisection.  `mixed` is the same bisection in mixed precision
(FP32 vectors on the levels above the finest with at least 2^20 entries,
FP64 sums, deflation, stopping test and finest level) beside the FP64 one,
with `per_level` comparing each FP32 level's power iteration with the FP64
one from the same start (iterations, angle, Rayleigh quotient, residual);
`mixed_grid3d` the same comparison on the 128^3 7-point grid with weights
U{1..16}.  `batched_rgg512` times the batched bisection of 512 random
geometric graphs (mi_bisect_batch) against one call per graph.

    python bench.py [--steps K] [--warmup W]
    python bench.py --roofline-only      # only the roofline measurement (for rocprofv3 runs)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SEED = 1
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_power_stream_rmat20_L1_traffic.txt")


def measured_traffic():
    """HBM bytes per launch of k_power_stream on R-MAT-20 level 1 from the
    committed rocprofv3 counter summary (line 'traffic_bytes_per_launch: X')."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(rp, col, w, gpu_stats, budget_s: float = 20.0) -> dict:
    """Times the oracle on a bounded sample of the same bisection: the full
    coarsening, then per level a capped number of power iterations (time per
    iteration), then the split; scales the per-iteration times by the GPU
    run's iteration counts (the oracle follows the same iteration)."""
    import oracle

    t0 = time.perf_counter()
    H = oracle.Hierarchy(rp, col, w, SEED)
    t_coarsen = time.perf_counter() - t0
    iters_gpu = gpu_stats["level_iters"]
    per_iter = []
    rng = np.random.default_rng(0)
    spent = t_coarsen
    for l in range(H.nlevels):
        L = H.levels[l]
        cap = 50 if L["n"] > 100000 else min(400, max(1, int(iters_gpu[l]) if l < len(iters_gpu) else 400))
        u0 = rng.random(L["n"])
        t = time.perf_counter()
        _, it, _, _ = oracle.power_iter(L["rowptr"], L["col"], L["w"], u0, 0, 1e-10, cap)
        dt = time.perf_counter() - t
        per_iter.append(dt / max(it, 1))
        spent += dt
        if spent > budget_s:
            break
    while len(per_iter) < H.nlevels:
        per_iter.append(per_iter[-1])
    ev = rng.standard_normal(len(rp) - 1)
    t = time.perf_counter()
    oracle.compute_partition(rp, col, ev)
    t_split = time.perf_counter() - t
    est = t_coarsen + sum(p * int(i) for p, i in zip(per_iter, iters_gpu)) + t_split
    return {"value": round(est * 1000.0, 1), "unit": "ms", "cores": 1, "kind": "port",
            "sample": (f"oracle (scalar C restatement of the reference Serial path) on the same graph: "
                       f"full coarsening {t_coarsen:.2f} s + capped power iterations per level "
                       f"(per-iteration time x the GPU run's iteration counts) + split {t_split:.2f} s")}


def merge_rate(st: dict) -> dict:
    """Algorithmic bytes of the coarse builds (P^T A P with duplicate merge)
    over their time: per level l -> l+1, read the fine CSR 4(n+1) + 4m + 4m
    (weights; the unit-weight finest level stores none) and vcmap 4n, write
    the coarse CSR 4(nc+1) + 8 m_c."""
    ns, ms = st["level_n"], st["level_m"]
    b = 0.0
    for l in range(len(ns) - 1):
        b += 4 * (ns[l] + 1) + 4 * ms[l] + (4 * ms[l] if l > 0 else 0) + 4 * ns[l]
        b += 4 * (ns[l + 1] + 1) + 8 * ms[l + 1]
    t = st["build_ms"]
    return {"alg_bytes": b, "ms": round(t, 3), "gbps": round(b / (t * 1e-3) / 1e9, 1) if t > 0 else None}


def mixed_summary(ms64: float, st64: dict, ms32: float, st32: dict) -> dict:
    """Mixed-precision bisection beside the plain FP64 one on the same graph."""
    return {"ms": round(ms32, 3), "fp64_ms": round(ms64, 3), "speedup": round(ms64 / ms32, 4),
            "edge_cut": st32["edge_cut"], "fp64_edge_cut": st64["edge_cut"],
            "cut_change_pct": round(100.0 * (st32["edge_cut"] - st64["edge_cut"]) / max(st64["edge_cut"], 1), 4),
            "fp32_levels": st32["mixed_levels"], "level_iters": st32["level_iters"],
            "fp64_level_iters": st64["level_iters"],
            "level_eigen_ms": [round(x, 3) for x in st32["level_eigen_ms"]]}


def mixed_levels_report(mi, ctx, g, seed: int) -> list:
    """Per level that runs in FP32 (at least 2^20 entries, above the
    finest): the level's power iteration in mixed precision and in FP64 from
    the same start vector (FP64 on the device, which the tests hold to the
    oracle), with iterations, the angle between the two vectors, and each
    one's Rayleigh quotient and relative Fiedler residual ||L u - lam u|| /
    ||lam u|| (L = D - A, computed on the host)."""
    import scipy.sparse as sp
    H = g.hierarchy(seed)
    out = []
    for l in range(1, H.nlevels):
        lv = H.level(l)
        if lv["m"] < (1 << 20):
            continue
        n = lv["n"]
        A = sp.csr_matrix((lv["w"].astype(np.float64), lv["col"], lv["rowptr"]), shape=(n, n))
        Lap = sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A
        lg = ctx.graph(lv["rowptr"], lv["col"], lv["w"])
        u0 = np.random.default_rng(l).random(n)
        ctx.set_precision(False)
        u64, it64, _ = lg.power_iter(u0)
        ctx.set_precision(True, 1)
        u32, it32, _ = lg.power_iter(u0)
        ctx.set_precision(False)
        lg.close()

        def quality(u):
            u = u - u.mean()
            lam = float(u @ (Lap @ u) / (u @ u))
            return lam, float(np.linalg.norm(Lap @ u - lam * u) / (abs(lam) * np.linalg.norm(u)))

        l64, r64 = quality(u64)
        l32, r32 = quality(u32)
        out.append({"level": l, "n": n, "m": lv["m"], "iters_fp64": it64, "iters_mixed": it32,
                    "cos": round(float(abs(u64 @ u32) / (np.linalg.norm(u64) * np.linalg.norm(u32))), 12),
                    "rayleigh_fp64": l64, "rayleigh_mixed": l32, "residual_fp64": round(r64, 6),
                    "residual_mixed": round(r32, 6)})
    H.close()
    return out


def grid3d_mixed(mi, ctx, steps: int) -> dict:
    """The 128^3 7-point grid with weights U{1..16} (seed 2): FP64 and
    mixed-precision bisections, timed on the host clock around `steps` calls."""
    from ipdps_2021_graph_coarsening_amd import graphs
    rp, col, w = graphs.grid3d(128, seed=2)
    g = ctx.graph(rp, col, w)
    res = {}
    for mixed in (False, True):
        ctx.set_precision(mixed)
        r = mi.partition_graph(ctx, g, hec_seed=2, pcg_state=2)
        t0 = time.perf_counter()
        for _ in range(steps):
            r = mi.partition_graph(ctx, g, hec_seed=2, pcg_state=2)
        res[mixed] = ((time.perf_counter() - t0) * 1000.0 / steps, r.stats)
    ctx.set_precision(False)
    g.close()
    out = mixed_summary(res[False][0], res[False][1], res[True][0], res[True][1])
    out["workload"] = f"grid3d-128 (n={len(rp) - 1}, m={int(rp[-1])}, weights U1..16, seed 2)"
    return out


def rgg_batch(mi, ctx, reps: int = 2) -> dict:
    """512 random geometric graphs (unit square, seed 3, n log-uniform in
    [1000, 200000], radius for mean degree 10, largest component): the
    batched bisection (mi_bisect_batch) against one call per graph, both
    timed on the host clock; the results are identical graph by graph."""
    from ipdps_2021_graph_coarsening_amd import graphs
    rng = np.random.default_rng(3)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(2e5), 512)).astype(int)
    t = time.perf_counter()
    G = [ctx.graph(*graphs.rgg(int(n), seed=3 + i)[:2]) for i, n in enumerate(ns)]
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20,
                    help="R-MAT scale; from 22 on the graph is drawn on the GPU (graphs.rmat_device, seed 4)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched measurement")
    ap.add_argument("--grid2d", type=int, default=0,
                    help="bisect a k x k grid instead of the R-MAT graph (implies --no-extras)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the plain FP64 R-MAT bisection and its roofline (no mixed, 3D grid or batch)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline measurement of the dominant kernel, plus a 256 MiB device copy "
                         "that calibrates rocprofv3's byte counters")
    ap.add_argument("--roofline-level", type=int, default=-1,
                    help="level for the roofline (default: the level with the most eigensolve time; "
                         "1 with --roofline-only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent graphs would shard one per GPU, with no collectives).", file=sys.stderr)
        return 2

    import ctypes as C
    import ipdps_2021_graph_coarsening_amd as mi
    from ipdps_2021_graph_coarsening_amd import graphs

    t = time.perf_counter()
    if args.grid2d:
        # (a graph whose hierarchy is mostly small levels, for profiling those)
        rp, col, w = graphs.grid2d(args.grid2d)
        name = f"grid2d-{args.grid2d}"
        args.no_extras = True
    elif args.scale >= 22:
        # (drawn on the GPU: the host generator takes minutes at these sizes)
        rp, col, w = graphs.rmat_device(args.scale, 16, 0.57, 0.19, 0.19, seed=4)
        name = f"R-MAT-{args.scale}"
    else:
        rp, col, w = graphs.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=SEED)
        name = f"R-MAT-{args.scale}"
    n, m = len(rp) - 1, int(rp[-1])
    log(f"{name}: n={n} m={m} generated in {time.perf_counter() - t:.1f} s")

    ctx = mi.Context(0)
    ctx.graph(np.array([0, 1, 2], dtype=np.uint32), np.array([1, 0], dtype=np.uint32)).close()  # first-use costs
    t = time.perf_counter()
    g = ctx.graph(rp, col, w)  # host -> HBM upload of the CSR (synchronous)
    upload_ms = (time.perf_counter() - t) * 1000.0
    log(f"graph upload (host arrays -> HBM): {upload_ms:.2f} ms")
    L = ctx._L
    cfg = mi.Config(4, 0, 0, 1, 1e-10, 2)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    ms = None
    stats = None
    if not args.roofline_only:
        d_part = C.c_void_p()
        assert hip.hipMalloc(ctypes.byref(d_part), 4 * n) == 0

        def step():
            st = mi.Stats()
            pcg = C.c_uint64(SEED)
            ec = C.c_uint32()
            ctx.check(L.mi_partition_graph(ctx.h, g.h, C.byref(cfg), 0, SEED, C.byref(pcg), 1, d_part, 1,
                                           C.byref(ec), C.byref(st)), "mi_partition_graph")
            return st

        def timed():
            for i in range(args.warmup):
                st = step()
                log(f"warmup {i}: {st.total_ms:.2f} ms, cut {st.edge_cut}, levels {st.nlevels}")
            hip.hipDeviceSynchronize()
            t0 = time.perf_counter()
            for i in range(args.steps):
                st = step()
            hip.hipDeviceSynchronize()
            return (time.perf_counter() - t0) * 1000.0 / args.steps, st.as_dict()

        ms, stats = timed()
        log(f"timed {args.steps} steps: {ms:.2f} ms/bisection, cut {stats['edge_cut']}, "
            f"iterations per level {stats['level_iters']}")
        # the same bisection in mixed precision (FP32 on the levels above the
        # finest with at least 2^20 entries, FP64 finest level)
        mixed = None
        if not args.no_extras:
            ctx.set_precision(True)
            ms_mixed, st_mixed = timed()
            ctx.set_precision(False)
            mixed = mixed_summary(ms, stats, ms_mixed, st_mixed)
            if args.scale < 22:
                mixed["per_level"] = mixed_levels_report(mi, ctx, g, SEED)
            log(f"mixed precision: {ms_mixed:.2f} ms/bisection, cut {st_mixed['edge_cut']}, "
                f"iterations per level {st_mixed['level_iters']}")
        hip.hipFree(d_part)

    # roofline of the dominant kernel: the streaming kernel of the power step
    # on the level with the most eigensolve time
    lvl = args.roofline_level
    if lvl < 0:
        lvl = 1 if stats is None else int(np.argmax(stats["level_eigen_ms"]))
    if lvl == 0:
        gl = g
    else:
        H = g.hierarchy(SEED)
        lv = H.level(lvl)
        H.close()
        gl = ctx.graph(lv["rowptr"], lv["col"], lv["w"])
    b = gl.bench_step(50)
    achieved = b["stream_bytes"] / (b["stream_ms"] * 1e-3) / 1e9
    log(f"level {lvl} (n={gl.n}, m={gl.m}): step {b['step_ms'] * 1000:.1f} us, "
        f"k_power_stream {b['stream_ms'] * 1000:.1f} us/launch, {b['stream_bytes'] / 1e6:.1f} MB algorithmic, "
        f"{achieved:.0f} GB/s")
    traffic = measured_traffic() if (lvl == 1 and args.scale == 20 and not args.grid2d) else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "kernel": ((f"k_power_stream_lds ({b['column_blocks']} LDS column blocks)" if b["column_blocks"]
                            else "k_power_stream (gathers from global memory)")
                           + f", FP64 power-iteration step, {name} level {lvl}"),
                "level_n": gl.n, "level_m": gl.m,
                "alg_bytes_per_launch": b["stream_bytes"], "impl_bytes_per_launch": b["impl_bytes"],
                "launch_us": round(b["stream_ms"] * 1000, 2),
                "step_us": round(b["step_ms"] * 1000, 2),
                "step_gbps": round(b["step_bytes"] / (b["step_ms"] * 1e-3) / 1e9, 1)}
    if args.roofline_only:
        # calibration for rocprofv3 FETCH_SIZE / WRITE_SIZE: a device copy of known size
        nb = 256 << 20
        a_, b_ = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(a_), nb) == 0 and hip.hipMalloc(ctypes.byref(b_), nb) == 0
        for _ in range(3):
            hip.hipMemcpy(b_, a_, nb, 3)
        hip.hipDeviceSynchronize()
        hip.hipFree(a_)
        hip.hipFree(b_)
        print(json.dumps({"roofline": roofline, "calibration_copy_bytes": nb}), flush=True)
        return 0

    g3 = None
    if not args.roofline_only and not args.no_extras:
        t = time.perf_counter()
        g3 = grid3d_mixed(mi, ctx, max(1, min(args.steps, 5)))
        log(f"grid3d-128: FP64 {g3['fp64_ms']:.2f} ms, mixed {g3['ms']:.2f} ms, cut change "
            f"{g3['cut_change_pct']}% ({time.perf_counter() - t:.1f} s)")

    batch = None
    if not args.roofline_only and not args.no_batch and not args.no_extras:
        batch = rgg_batch(mi, ctx)
        log(f"512 RGGs: batch {batch['batch_ms']} ms, one call per graph {batch['one_call_per_graph_ms']} ms")

    cpu = None
    if not args.no_cpu_baseline and args.scale < 22 and not args.grid2d:  # (the oracle would take hours beyond)
        t = time.perf_counter()
        cpu = cpu_baseline(rp, col, w, stats)
        log(f"cpu baseline {cpu['value']} ms (sampled in {time.perf_counter() - t:.1f} s)")

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
        "config": {"workload": (f"grid2d-{args.grid2d} unit-weight plain-fp64 bisection" if args.grid2d else
                                f"rmat{args.scale}-ef16-lcc-weighted plain-fp64 bisection"),
                   "n": n, "m": m, "seed": SEED, "levels": stats["nlevels"],
                   "upload_ms": round(upload_ms, 3),
                   "level_n": stats["level_n"], "level_iters": stats["level_iters"],
                   "edge_cut": stats["edge_cut"],
                   "phases_ms": {k: round(stats[k], 3) for k in
                                 ("coarsen_ms", "eigen_ms", "partition_ms", "permute_ms", "heavy_ms",
                                  "map_ms", "build_ms")},
                   "level_eigen_ms": [round(x, 3) for x in stats["level_eigen_ms"]],
                   "level_us_per_iteration": [round(1000.0 * t / max(i, 1), 2) for t, i in
                                              zip(stats["level_eigen_ms"], stats["level_iters"])],
                   "merge": merge_rate(stats)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "mixed_grid3d": g3,
        "batched_rgg512": batch,
    }
    print(json.dumps(out), flush=True)
    g.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
