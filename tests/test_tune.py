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


def test_batched_steps_follow_the_sliding_window():
    """Each fused batch inserts the next window edges and deletes the oldest,
    half and half, in order; sources lie in [0, N)."""
    import numpy as np

    from gpma_bfs_amd.tune import batched_steps

    keys = np.arange(40_000_000, dtype=np.uint64)
    N = 1000
    plan, ins, dele = batched_steps(N, keys, steps=3, fused=8, nsrc=64, seed=2)
    half = len(keys) // 2
    nxt_ins, nxt_del = half, 0
    for k, v, offs, starts in plan:
        assert len(offs) == 9 and offs[0] == 0 and int(offs[-1]) == len(k) == len(v)
        assert starts.shape == (64,) and int(starts.max()) < N
        for b in range(8):
            a0, a1 = int(offs[b]), int(offs[b + 1])
            h = (a1 - a0) // 2
            assert 500 <= h <= 500_000 and (a1 - a0) % 2 == 0
            assert np.array_equal(k[a0: a0 + h], keys[nxt_ins: nxt_ins + h]) and v[a0: a0 + h].all()
            assert np.array_equal(k[a0 + h: a1], keys[nxt_del: nxt_del + h]) and not v[a0 + h: a1].any()
            nxt_ins += h
            nxt_del += h
    assert (ins, dele) == (nxt_ins, nxt_del)
