"""Choice logic of the autotuner (CPU: timings given), and the tuner itself
on a small graph (GPU)."""
import json
import os

import numpy as np
import pytest

from ipdps_2021_graph_coarsening_amd import autotune


def _row(small, mixed, ms, cuts, mm=1 << 20):
    return {"small_max_entries": small, "mixed": mixed, "mixed_min_entries": mm, "ms": ms, "cuts": cuts}


def test_choose_fastest_fp64():
    t = [_row(0, False, 10.0, [100]), _row(1 << 17, False, 6.0, [100]), _row(1 << 18, False, 7.0, [101])]
    assert autotune.choose(t, 0.01) == {"small_max_entries": 1 << 17, "mixed": False, "mixed_min_entries": 1 << 20}


def test_choose_mixed_only_when_faster_at_equal_cut():
    base = [_row(0, False, 10.0, [1000, 50]), _row(1 << 17, False, 8.0, [1000, 50])]
    slower = base + [_row(1 << 17, True, 8.5, [1000, 50], 1 << 18)]
    assert autotune.choose(slower, 0.01)["mixed"] is False
    worse_cut = base + [_row(1 << 17, True, 7.0, [1000, 52], 1 << 18)]  # 4 % off on the second graph
    assert autotune.choose(worse_cut, 0.01)["mixed"] is False
    good = base + [_row(1 << 17, True, 7.5, [1005, 50], 1 << 22), _row(1 << 17, True, 7.0, [995, 50], 1 << 20)]
    assert autotune.choose(good, 0.01) == {"small_max_entries": 1 << 17, "mixed": True,
                                           "mixed_min_entries": 1 << 20}


def test_choose_keeps_default_within_noise():
    t = [_row(0, False, 20.0, [9]), _row(1 << 17, False, 11.3, [9]), _row(1 << 15, False, 11.25, [9]),
         _row(1 << 15, True, 11.2, [9], 1 << 18)]
    assert autotune.choose(t, 0.01) == autotune.DEFAULT


def test_save_load_roundtrip(tmp_path):
    p = os.path.join(tmp_path, "t.json")
    autotune.save({"small_max_entries": 65536, "mixed": True, "mixed_min_entries": 4096}, p)
    assert autotune.load(p) == {"small_max_entries": 65536, "mixed": True, "mixed_min_entries": 4096}
    json.dump({"mixed": True}, open(p, "w"))
    with pytest.raises(ValueError):
        autotune.load(p)


@pytest.mark.gpu
def test_tune_small_graph(ctx):
    from ipdps_2021_graph_coarsening_amd import graphs
    import ipdps_2021_graph_coarsening_amd as mi
    rp, col, w = graphs.grid2d(96)
    g = ctx.graph(rp, col, w)
    try:
        best = autotune.tune(ctx, [g], reps=1)
        assert len(best["table"]) == len(autotune.SMALL_CANDIDATES) + len(autotune.MIXED_CANDIDATES)
        r1 = mi.partition_graph(ctx, g, hec_seed=1, pcg_state=1)
        autotune.apply(ctx, best)
        r2 = mi.partition_graph(ctx, g, hec_seed=1, pcg_state=1)
        assert np.array_equal(r1.part, r2.part)
    finally:
        autotune.apply(ctx, autotune.DEFAULT)
        g.close()
