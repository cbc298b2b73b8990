"""Host logic of the launch-shape autotuner (mobileraytracer_amd/tuning.py):
the coordinate descent on synthetic cost tables (known optimum), its noise
threshold and single measurement per configuration, workload classes, the
table's JSON form, and that every searched value is one mrt_ctx_set_tuning
accepts."""
from __future__ import annotations

import itertools
import json

import pytest

from mobileraytracer_amd import abi, tuning


def separable(costs):
    """measure(cfg) = sum of per-knob costs, counting calls."""
    calls = []

    def measure(cfg):
        calls.append(dict(cfg))
        return 10.0 + sum(costs[k][v] for k, v in cfg.items())
    return measure, calls


def test_descent_finds_separable_optimum():
    space = {"a": (1, 2, 3), "b": (10, 20), "c": (5, 6, 7, 8)}
    costs = {"a": {1: 3.0, 2: 0.0, 3: 1.0}, "b": {10: 2.0, 20: 0.0}, "c": {5: 1.0, 6: 2.0, 7: 0.0, 8: 4.0}}
    measure, calls = separable(costs)
    best, t, trials = tuning.coordinate_descent(measure, space, {"a": 1, "b": 10, "c": 5}, passes=2)
    brute = min(itertools.product(*space.values()),
                key=lambda v: sum(costs[k][x] for k, x in zip(space, v)))
    assert tuple(best[k] for k in space) == brute
    assert t == 10.0
    # Each configuration is measured once, and the trials list them all.
    ids = [tuple(sorted(c.items())) for c in calls]
    assert len(ids) == len(set(ids)) == len(trials)


def test_descent_second_pass_finds_coupled_optimum():
    """A cost where knob a's best value depends on b: one pass stops at a
    worse point, a second pass reaches the optimum."""
    space = {"a": (0, 1), "b": (0, 1)}
    table = {(0, 0): 5.0, (1, 0): 6.0, (0, 1): 4.0, (1, 1): 1.0}
    best1, t1, _ = tuning.coordinate_descent(lambda c: table[(c["a"], c["b"])], space, {"a": 0, "b": 0}, passes=1)
    best2, t2, _ = tuning.coordinate_descent(lambda c: table[(c["a"], c["b"])], space, {"a": 0, "b": 0}, passes=2)
    assert (best1, t1) == ({"a": 0, "b": 1}, 4.0)
    assert (best2, t2) == ({"a": 1, "b": 1}, 1.0)


def test_descent_ignores_gains_below_threshold():
    space = {"a": (0, 1, 2)}
    t = {0: 100.0, 1: 99.0, 2: 97.0}
    best, ms, _ = tuning.coordinate_descent(lambda c: t[c["a"]], space, {"a": 0}, min_gain=0.02)
    assert best == {"a": 2} and ms == 97.0
    best, ms, _ = tuning.coordinate_descent(lambda c: t[c["a"]], space, {"a": 0}, min_gain=0.05)
    assert best == {"a": 0} and ms == 100.0


def test_workload_classes():
    k = tuning.workload_key
    assert k(abi.BVH, 990_000, 1920, 1080, 16) == k(abi.BVH, 1_000_000, 1920, 1080, 64)
    assert k(abi.BVH, 1_000_000, 1920, 1080, 16) != k(abi.REGULAR_GRID, 1_000_000, 1920, 1080, 16)
    assert k(abi.BVH, 1_000_000, 256, 256, 16) != k(abi.BVH, 1_000_000, 1920, 1080, 16)
    assert k(abi.BVH, 1_000_000, 256, 256, 1) != k(abi.BVH, 1_000_000, 256, 256, 4)
    assert k(abi.BVH, 100, 256, 256, 4) != k(abi.BVH, 1_000_000, 256, 256, 4)
    # Classes are powers of two: (2^19, 2^20] is one class.
    assert k(abi.BVH, 524_289, 64, 64, 2) == k(abi.BVH, 1 << 20, 64, 64, 2) != k(abi.BVH, (1 << 20) + 1, 64, 64, 2)


def test_search_space_within_library_limits():
    """Every searched value passes mrt_ctx_set_tuning's checks (mirrored
    here: the library refuses the rest with MRT_E_INVALID)."""
    limits = {"refill_lanes": lambda v: 0 <= v <= 64,
              "fetch_chunk": lambda v: v == 0 or 64 <= v <= 4096,
              "launch_share_pct": lambda v: 0 <= v <= 100,
              "max_slot_lanes": lambda v: v in (0, 1, 2, 4, 8, 16),
              "max_groups": lambda v: 0 <= v <= 3,
              "overlap_loads": lambda v: 0 <= v <= 2,
              "graph_loop": lambda v: 0 <= v <= 2,
              "service_lanes": lambda v: 0 <= v <= 64}
    assert set(limits) == set(tuning.KNOBS)
    for shape, space in tuning.SEARCH_SPACE.items():
        for knob, values in space.items():
            assert knob in tuning.KNOBS, (shape, knob)
            assert all(limits[knob](v) for v in values), (shape, knob, values)
            assert len(set(values)) == len(values)


def test_table_json_round_trip(tmp_path):
    t = tuning.Autotuner(ctx=None)
    t.table["a0-p20-px21-s4"] = tuning.TuneResult("a0-p20-px21-s4", tuning.MEGAKERNEL, {"service_lanes": 24}, 131.0,
                                                 134.0, [{"shape": "megakernel", "knobs": {}, "ms": 134.0}])
    t.table["a1-p20-px21-s1"] = tuning.TuneResult("a1-p20-px21-s1", tuning.WAVEFRONT, {"max_groups": 2}, 50.0, 52.0)
    p = tmp_path / "t.json"
    t.save(p)
    u = tuning.Autotuner(ctx=None)
    u.load(p)
    assert {k: r.to_json() for k, r in u.table.items()} == {k: r.to_json() for k, r in t.table.items()}
    bad = json.loads(p.read_text())
    bad["a0-p20-px21-s4"]["knobs"] = {"no_such_knob": 1}
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError):
        tuning.Autotuner(ctx=None).load(p)
