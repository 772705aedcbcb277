"""GPU parity of gpma_step, the batched update -> rebalance -> BFS step run as
one captured HIP graph: every step must leave the same array (layout
included) and return the same BFS levels and reachable counts as the host
path (gpma_update_batches then gpma_bfs_multi) on a twin handle, including
the steps that continue on the host after the replay (bulk loads, root
resizes, BFS deeper than the captured levels) and the steps that never take
a graph.
"""
import numpy as np
import pytest

import gpma_bfs_amd as G
from oracle import oracle as O
from tests.helpers import assert_invariants

pytestmark = pytest.mark.gpu


def _uneven_step(rng, keys, ins, dele, fused, lo, hi):
    """`fused` batches of log-uniform sizes in [lo, hi], each half inserts of
    the next window edges and half deletes of the oldest."""
    ks, vs, offs = [], [], [0]
    for _ in range(fused):
        h = max(1, int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) // 2)
        ks += [keys[ins: ins + h], keys[dele: dele + h]]
        vs += [np.ones(h, np.uint8), np.zeros(h, np.uint8)]
        ins += h
        dele += h
        offs.append(offs[-1] + 2 * h)
    assert ins <= len(keys), "not enough edges for the steps"
    return np.concatenate(ks), np.concatenate(vs), np.array(offs, np.uint64), ins, dele


def _same_state(a, b, msg):
    ka, va, ra = a.copy_layout()
    kb, vb, rb = b.copy_layout()
    assert np.array_equal(ka, kb) and np.array_equal(va, vb) and np.array_equal(ra, rb), msg


def test_step_graph_equals_host_path():
    """Uneven fused batches over two size classes, 64 sources per step."""
    N = 20000
    src, dst = G.gen_rmat(15, 1500000, seed=12)
    m = (src < N) & (dst < N)
    keys = G.make_keys(src[m], dst[m])
    half = len(keys) // 2
    a, b = G.GPMA(N), G.GPMA(N)
    for g in (a, b):
        G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    rng = np.random.default_rng(4)
    ins, dele = half, 0
    for step in range(12):
        # size classes 2^16 .. 2^18
        k, v, offs, ins, dele = _uneven_step(rng, keys, ins, dele, 8, 100, 8000 if step < 8 else 60000)
        starts = rng.integers(0, N, 64).astype(np.uint32)
        starts[3] = starts[4]
        rc, lv, reach = G.gpma_step(a, k, v, offs, starts, return_levels=True)
        assert rc == G.GPMA_OK
        G.update_gpma(b, k, v)
        lv_b, reach_b = G.gpma_bfs_multi(b, starts)
        assert np.array_equal(lv, lv_b) and np.array_equal(reach, reach_b), f"step {step}"
        _same_state(a, b, f"step {step}")
    win = keys[dele: ins]
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    expect = np.sort(np.concatenate([win, guards, np.array([G.KEY_MAX, G.KEY_MAX], np.uint64)]))
    assert np.array_equal(a.export_live(), expect)
    s = (win >> np.uint64(32)).astype(np.uint32)
    d = (win & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    olv, oreach = O.bfs_edges(N, s, d, int(starts[0]))
    assert np.array_equal(lv[0], olv) and reach[0] == oreach
    st = G.gpma_step_stats(a)
    assert st["steps"] == 12 and st["replays"] >= 10 and st["captures"] <= 8, st
    assert_invariants(a)


def test_step_device_levels_update_only_and_key_range():
    """Levels kept in a device buffer, a step without sources, and a step
    with out-of-range keys (GPMA_ERR_KEY_RANGE, the valid records applied)."""
    N = 5000
    src, dst = G.gen_rmat(13, 60000, seed=3)
    m = (src < N) & (dst < N)
    keys = G.make_keys(src[m], dst[m])
    half = len(keys) // 2
    a, b = G.GPMA(N), G.GPMA(N)
    for g in (a, b):
        G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    rng = np.random.default_rng(9)
    dl = G.DeviceBuffer(16 * N * 4)
    try:
        ins, dele = half, 0
        k, v, offs, ins, dele = _uneven_step(rng, keys, ins, dele, 4, 50, 5000)
        starts = rng.integers(0, N, 16).astype(np.uint32)
        dk, dv = G.DeviceBuffer.from_numpy(k), G.DeviceBuffer.from_numpy(v)
        rc, lv, reach = G.gpma_step(a, dk.ptr, dv.ptr, offs, starts, levels_device_ptr=dl.ptr)
        dk.free()
        dv.free()
        assert rc == G.GPMA_OK and lv is None
        G.update_gpma(b, k, v)
        lv_b, reach_b = G.gpma_bfs_multi(b, starts)
        assert np.array_equal(dl.to_numpy(np.uint32, 16 * N).reshape(16, N), lv_b)
        assert np.array_equal(reach, reach_b)
        # update only
        k, v, offs, ins, dele = _uneven_step(rng, keys, ins, dele, 3, 50, 5000)
        rc, _, r0 = G.gpma_step(a, k, v, offs)
        assert rc == G.GPMA_OK and len(r0) == 0
        G.update_gpma(b, k, v)
        _same_state(a, b, "update-only step")
        # out-of-range keys: reported, the rest applied
        k, v, offs, ins, dele = _uneven_step(rng, keys, ins, dele, 3, 50, 5000)
        bad = np.array([(np.uint64(N + 3) << np.uint64(32)) | np.uint64(1), np.uint64(7) << np.uint64(32) | np.uint64(N)],
                       np.uint64)
        k = np.concatenate([k, bad])
        v = np.concatenate([v, np.ones(2, np.uint8)])
        offs = np.append(offs, offs[-1] + 2)
        rc, lv, reach = G.gpma_step(a, k, v, offs, starts, return_levels=True, strict=False)
        assert rc == G.GPMA_ERR_KEY_RANGE
        assert G.update_gpma(b, k, v, strict=False) == G.GPMA_ERR_KEY_RANGE
        lv_b, reach_b = G.gpma_bfs_multi(b, starts)
        assert np.array_equal(lv, lv_b) and np.array_equal(reach, reach_b)
        _same_state(a, b, "key-range step")
        assert a.stats()["invalid_skipped"] == 2
    finally:
        dl.free()


def test_step_host_continuations():
    """Work past what a graph holds continues on the host with the same
    results: a bulk batch, growth through root resizes (new layouts capture
    new graphs), a BFS far deeper than the captured levels, and a skewed
    batch that turns the bucket sort off for the following steps."""
    N = 600
    # a chain 0 -> 1 -> ... -> N-1 makes every BFS from 0 N levels deep
    chain = G.make_keys(np.arange(N - 1, dtype=np.uint32), np.arange(1, N, dtype=np.uint32))
    rng = np.random.default_rng(21)
    extra = np.unique(G.make_keys(rng.integers(0, N, 40000).astype(np.uint32),
                                  rng.integers(0, N, 40000).astype(np.uint32)))
    extra = extra[~np.isin(extra, chain)]
    rng.shuffle(extra)
    a, b = G.GPMA(N, bulk_threshold=6000), G.GPMA(N, bulk_threshold=6000)
    starts = np.array([0, 1, 2, N // 2], np.uint32)
    plan = [chain[: N // 2], chain[N // 2:], extra[:3000], extra[3000:12000], extra[12000:12500], extra[12500:20000]]
    for i, k in enumerate(plan):
        v = np.ones(len(k), np.uint8)
        offs = np.array([0, len(k) // 3, len(k)], np.uint64)
        rc, lv, reach = G.gpma_step(a, k, v, offs, starts, return_levels=True)
        assert rc == G.GPMA_OK
        G.update_gpma(b, k, v)
        lv_b, reach_b = G.gpma_bfs_multi(b, starts)
        assert np.array_equal(lv, lv_b) and np.array_equal(reach, reach_b), f"step {i}"
        _same_state(a, b, f"step {i}")
    # deletes of most extra edges: the chain keeps the BFS deep
    k = extra[:15000]
    v = np.zeros(len(k), np.uint8)
    rc, lv, reach = G.gpma_step(a, k, v, np.array([0, len(k)], np.uint64), starts, return_levels=True)
    G.update_gpma(b, k, v)
    lv_b, reach_b = G.gpma_bfs_multi(b, starts)
    assert np.array_equal(lv, lv_b) and np.array_equal(reach, reach_b)
    _same_state(a, b, "deletes")
    st = G.gpma_step_stats(a)
    assert st["update_tails"] >= 1 and st["bfs_tails"] >= 1 and st["captures"] >= 2, st
    assert a.stats()["resizes"] >= 1 and a.stats()["bulk_loads"] >= 1
    # one source crowding 20000 records into few sort buckets, then ordinary steps
    Nb = 40000
    c, d = G.GPMA(Nb), G.GPMA(Nb)
    hub = (np.uint64(5) << np.uint64(32)) | rng.permutation(Nb)[:20000].astype(np.uint64)
    rest = np.unique((rng.integers(0, Nb, 30000).astype(np.uint64) << np.uint64(32)) |
                     rng.integers(0, Nb, 30000).astype(np.uint64))
    for i, k in enumerate([hub, rest[:10000], rest[10000:20000]]):
        v = np.ones(len(k), np.uint8)
        offs = np.array([0, len(k)], np.uint64)
        rc, lv, reach = G.gpma_step(c, k, v, offs, starts, return_levels=True)
        G.update_gpma(d, k, v)
        lv_b, reach_b = G.gpma_bfs_multi(d, starts)
        assert np.array_equal(lv, lv_b) and np.array_equal(reach, reach_b), f"skewed step {i}"
        _same_state(c, d, f"skewed step {i}")
    assert G.gpma_step_stats(c)["host_steps"] >= 2
    assert_invariants(a)
    assert_invariants(c)


def test_step_rebalance_levels_grow():
    """A step whose inserts climb above the captured rebalance levels is
    finished on the host with the same results; later captures hold the
    higher levels, so a like step then needs no host continuation."""
    N = 5000
    src, dst = G.gen_rmat(13, 60000, seed=7)
    m = (src < N) & (dst < N)
    keys = np.unique(G.make_keys(src[m], dst[m]))
    a, b = G.GPMA(N), G.GPMA(N)
    for g in (a, b):
        G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    rng = np.random.default_rng(13)
    starts = rng.integers(0, N, 8).astype(np.uint32)
    tails = []
    for row in (1001, 4003):  # far apart: each climbs to about level 6 on its own
        have = keys[(keys >> np.uint64(32)) == np.uint64(row)] & np.uint64(0xFFFFFFFF)
        new = np.setdiff1d(rng.permutation(N)[:200].astype(np.uint64), have)[:100]
        k = (np.uint64(row) << np.uint64(32)) | np.sort(new)
        v = np.ones(len(k), np.uint8)
        rc, lv, reach = G.gpma_step(a, k, v, np.array([0, len(k)], np.uint64), starts, return_levels=True)
        assert rc == G.GPMA_OK
        G.update_gpma(b, k, v)
        lv_b, reach_b = G.gpma_bfs_multi(b, starts)
        assert np.array_equal(lv, lv_b) and np.array_equal(reach, reach_b), f"row {row}"
        _same_state(a, b, f"row {row}")
        tails.append((G.gpma_step_stats(a)["update_tails"], b.stats()["last_levels"]))
    assert 5 < tails[0][1] - 1 <= 8, tails  # the first step climbed above five levels
    assert tails[0][0] == 1 and tails[1][0] == 1 and tails[1][1] <= tails[0][1], tails
    assert_invariants(a)
