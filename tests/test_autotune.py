"""Coordinate-descent search of the autotuner (CPU: synthetic objectives, no device)."""
import importlib
import json

autotune = importlib.import_module("wavefront-renderer_amd.autotune")


def test_finds_separable_optimum_and_reuses_measurements():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30], "c": ["x", "y"]}
    target = {"a": 3, "b": 20, "c": "y"}
    calls = []

    def measure(s):
        calls.append(dict(s))
        return sum(0 if s[k] == target[k] else 1 for k in s) + 0.1 * s["a"]

    res = autotune.coordinate_descent(measure, space, {"a": 1, "b": 10, "c": "x"}, rounds=3)
    assert res["best"] == target
    # every measured setting is distinct (repeats come from the memo)
    keys = [tuple(sorted(c.items())) for c in calls]
    assert len(keys) == len(set(keys)) == len(res["trials"])
    # a round that improves nothing ends the search early
    assert len(calls) <= 1 + 2 * (3 + 2 + 1)


def test_coupled_objective_improves_monotonically():
    space = {"w": [1, 2, 4, 8], "l": [8, 16, 32, 64]}

    def measure(s):     # best at w * l == 128, a ridge the axis-wise search must follow
        return abs(s["w"] * s["l"] - 128) + 0.01 * s["w"]

    res = autotune.coordinate_descent(measure, space, {"w": 1, "l": 8}, rounds=4)
    assert measure(res["best"]) <= min(t for _, t in res["trials"])
    assert res["best"]["w"] * res["best"]["l"] == 128


def test_cache_round_trip(tmp_path):
    p = tmp_path / "tune.json"
    autotune._store(p, "k", {"best": {"a": 1}, "seconds": 0.5, "trials": 3})
    assert autotune._cached(p, "k")["best"] == {"a": 1}
    assert autotune._cached(p, "other") is None
    p.write_text("not json")
    assert autotune._cached(p, "k") is None
    autotune._store(p, "k", {"best": {}, "seconds": 1.0, "trials": 1})
    assert json.loads(p.read_text())["k"]["seconds"] == 1.0


def test_start_outside_space_moves_to_nearest_candidate():
    seen = []

    def measure(s):
        seen.append(dict(s))
        return 1.0

    res = autotune.coordinate_descent(measure, {"w": [1, 2], "m": ["a", "b"]},
                                      {"w": 4, "m": "z", "other": 7}, rounds=1)
    assert seen[0] == {"w": 2, "m": "a", "other": 7}
    assert res["best"]["w"] in (1, 2) and res["best"]["m"] in ("a", "b")
