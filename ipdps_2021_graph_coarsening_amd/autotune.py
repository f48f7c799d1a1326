"""Autotuner for the level policies of the eigensolve.

Two policies decide how each level of a hierarchy runs its power iteration:
- which levels iterate inside one workgroup (`mi_ctx_set_small_levels`: at
  most 16384 vertices and at most `small_max_entries` stored entries), and
- which levels above the finest run in mixed precision
  (`mi_ctx_set_precision`: at least `mixed_min_entries` entries).

Their best thresholds depend on the graphs and the device, so `tune` times
whole bisections of the caller's graphs (or a sample of them) for each
candidate and keeps the fastest.  Mixed precision is only kept when it is
faster and every graph's cut stays within `cut_rel` of the FP64 cut (the
comparison is "at equal cut").  The result is a small dict that `apply`
sets on a context and `save` / `load` keep as JSON.

    python -m ipdps_2021_graph_coarsening_amd.autotune --rmat 20 --out tuning.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Callable, Iterable, Sequence

SMALL_CANDIDATES = (0, 1 << 15, 1 << 16, 1 << 17, 1 << 18)
MIXED_CANDIDATES = (1 << 18, 1 << 20, 1 << 22)
DEFAULT = {"small_max_entries": 1 << 17, "mixed": False, "mixed_min_entries": 1 << 20}


def apply(ctx, tuning: dict) -> None:
    """Sets a tuning (as `tune` returns it) on a context."""
    ctx.set_small_levels(int(tuning["small_max_entries"]))
    ctx.set_precision(bool(tuning["mixed"]), int(tuning["mixed_min_entries"]))


def save(tuning: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(tuning, f, indent=1, sort_keys=True)


def load(path: str) -> dict:
    with open(path) as f:
        t = json.load(f)
    for k in DEFAULT:
        if k not in t:
            raise ValueError(f"tuning file {path} lacks {k}")
    return t


def _measure(run: Callable[[], Sequence[int]], reps: int) -> tuple[float, list[int]]:
    """Best wall time (ms) over reps runs after one warm-up, and the cuts."""
    cuts = list(run())
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        cuts = list(run())
        best = min(best, (time.perf_counter() - t) * 1000.0)
    return best, cuts


def choose(timings: Iterable[dict], cut_rel: float, margin: float = 0.01) -> dict:
    """Starting from the default setting, a candidate is taken only when it
    is faster than the current choice by more than `margin` (timing noise is
    about 1 %): first among the FP64 ones, then a mixed one that also keeps
    every cut within cut_rel of the FP64 cuts of its small-level setting.
    Each timing: {"small_max_entries", "mixed", "mixed_min_entries", "ms",
    "cuts"}."""
    rows = list(timings)
    fp64 = [r for r in rows if not r["mixed"]]
    if not fp64:
        raise ValueError("no FP64 candidate timed")
    best = next((r for r in fp64 if r["small_max_entries"] == DEFAULT["small_max_entries"]), fp64[0])
    for r in sorted(fp64, key=lambda r: r["ms"]):
        if r["ms"] < best["ms"] * (1 - margin):
            best = r
    for r in sorted((r for r in rows if r["mixed"]), key=lambda r: r["ms"]):
        if r["ms"] >= best["ms"] * (1 - margin):
            continue
        ref = next(f for f in fp64 if f["small_max_entries"] == r["small_max_entries"])
        if all(abs(a - b) <= cut_rel * max(b, 1) for a, b in zip(r["cuts"], ref["cuts"])):
            best = r
            break
    return {k: best[k] for k in ("small_max_entries", "mixed", "mixed_min_entries")}


def tune(ctx, graphs: Sequence, reps: int = 3, allow_mixed: bool = True, cut_rel: float = 0.01,
         hec_seed: int = 1, pcg_state: int = 1, log: Callable[[str], None] | None = None) -> dict:
    """Times bisections of `graphs` (device graphs of `ctx`) for every
    candidate setting and returns the chosen one, with the table of
    timings under "table".  The context is left with the chosen setting."""
    import ipdps_2021_graph_coarsening_amd as mi

    def run():
        return [mi.partition_graph(ctx, g, hec_seed=hec_seed, pcg_state=pcg_state).edge_cut for g in graphs]

    table = []
    for small in SMALL_CANDIDATES:
        ctx.set_small_levels(small)
        ctx.set_precision(False)
        ms, cuts = _measure(run, reps)
        table.append({"small_max_entries": small, "mixed": False, "mixed_min_entries": DEFAULT["mixed_min_entries"],
                      "ms": ms, "cuts": cuts})
        if log:
            log(f"small<= {small}: {ms:.2f} ms")
    if allow_mixed:
        small = min((r for r in table), key=lambda r: r["ms"])["small_max_entries"]
        ctx.set_small_levels(small)
        for mm in MIXED_CANDIDATES:
            ctx.set_precision(True, mm)
            ms, cuts = _measure(run, reps)
            table.append({"small_max_entries": small, "mixed": True, "mixed_min_entries": mm, "ms": ms,
                          "cuts": cuts})
            if log:
                log(f"small<= {small}, mixed from {mm} entries: {ms:.2f} ms")
    best = choose(table, cut_rel)
    apply(ctx, best)
    best["table"] = table
    return best


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--rmat", type=int, default=0, help="tune on an R-MAT graph of this scale (ef 16, seed 1)")
    ap.add_argument("--grid", type=int, default=0, help="tune on a k x k grid")
    ap.add_argument("--csr", default="", help="tune on a graph in the reference's binary CSR format")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    import ipdps_2021_graph_coarsening_amd as mi
    from ipdps_2021_graph_coarsening_amd import graphs as gen

    if a.csr:
        rp, col, w = gen.read_csr(a.csr)
    elif a.rmat:
        rp, col, w = gen.rmat(a.rmat, 16, seed=1)
    else:
        rp, col, w = gen.grid2d(a.grid or 256)
    with mi.Context(0) as ctx:
        g = ctx.graph(rp, col, w)
        best = tune(ctx, [g], reps=a.reps, allow_mixed=not a.fp64_only, log=lambda s: print(s, flush=True))
        g.close()
    table = best.pop("table")
    print(json.dumps({"chosen": best, "table": table}))
    if a.out:
        save(best, a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
