"""Autotuner search logic (CPU): coordinate descent on synthetic costs."""
import pytest

from gpma_bfs_amd.tune import coordinate_descent


def test_coordinate_descent_finds_separable_minimum():
    space = {"a": [0, 1, 2, 3], "b": [10, 20, 30], "c": [5, 6]}
    target = {"a": 2, "b": 30, "c": 5}

    def cost(cfg):
        return sum((cfg[k] - target[k]) ** 2 for k in cfg) + 1.0

    best, t, n = coordinate_descent(cost, space)
    assert best == target and t == 1.0
    assert n <= 4 * 3 * 2  # never more than the grid


def test_invalid_candidates_are_skipped():
    space = {"a": [0, 1, 2], "b": [0, 1]}

    def cost(cfg):
        if cfg["a"] == 2:
            return None  # fails its correctness check
        return 10 - cfg["a"] - cfg["b"]

    best, t, _ = coordinate_descent(cost, space)
    assert best == {"a": 1, "b": 1} and t == 8


def test_failing_start_raises():
    with pytest.raises(RuntimeError):
        coordinate_descent(lambda cfg: None, {"a": [0, 1]})
