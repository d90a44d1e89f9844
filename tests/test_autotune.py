"""The autotuner's coordinate search, on synthetic cost functions (CPU)."""
import importlib

import pytest

A = importlib.import_module("fnec-rust_amd.tuning")


def test_search_finds_separable_optimum():
    space = {"a": (1, 2, 3, 4), "b": (10, 20, 30), "c": (0, 1)}
    cost = lambda k: (k["a"] - 3) ** 2 + 0.1 * (k["b"] - 20) ** 2 + 5.0 * k["c"] + 1.0  # noqa: E731
    best, t, trials = A.search(space, {"a": 1, "b": 10, "c": 1}, lambda k: (cost(k), True))
    assert best == {"a": 3, "b": 20, "c": 0}
    assert t == pytest.approx(1.0)
    # no setting is measured twice
    keys = [tuple(sorted(k.items())) for k, _, _ in trials]
    assert len(keys) == len(set(keys))


def test_search_keeps_the_incumbent_within_the_noise_margin():
    space = {"a": (1, 2, 3)}
    cost = {1: 1.0, 2: 0.99, 3: 0.9}
    best, t, _ = A.search(space, {"a": 1}, lambda k: (cost[k["a"]], True), min_gain=0.02)
    assert best == {"a": 3} and t == pytest.approx(0.9)
    cost[3] = 0.985  # every candidate within 2% of the start: no move
    best, t, _ = A.search(space, {"a": 1}, lambda k: (cost[k["a"]], True), min_gain=0.02)
    assert best == {"a": 1} and t == pytest.approx(1.0)


def test_search_never_picks_a_rejected_candidate():
    space = {"a": (1, 2, 3)}
    best, t, trials = A.search(space, {"a": 1}, lambda k: (1.0 / k["a"], k["a"] != 3))
    assert best == {"a": 2}
    assert any(k == {"a": 3} and not ok for k, _, ok in trials)


def test_search_rejects_failing_start():
    with pytest.raises(RuntimeError):
        A.search({"a": (1, 2)}, {"a": 1}, lambda k: (1.0, False))


def test_default_space_matches_engine_ranges():
    # values the engine accepts (fnec_hip_set_tunable ranges)
    assert set(A.SPACE["update_tile_m"]) <= {64, 128}
    assert set(A.SPACE["update_k128"]) <= {0, 1}
    assert all(0 <= v <= 65536 for v in A.SPACE["gang_n"])
    assert all(1 <= v <= 64 for v in A.SPACE["gang_max"])
    assert all(0 <= v <= 8192 for v in A.SPACE["solve1_max_n"])
    assert all(1 <= v <= 12 for v in A.SPACE["mix_first_chunk"])


def test_package_autotune_is_the_function():
    F = importlib.import_module("fnec-rust_amd")
    assert callable(F.autotune) and F.autotune.__module__ == F.__name__
