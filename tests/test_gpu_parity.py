"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Layout checks are exact: keys[], live flags and row_offset[] of the GPU
array equal the oracle's restatement of the reference after every batch.
Live-key sets equal the closed form of the sliding window; BFS levels equal
the oracle and the reference's own cpu_baseline fixtures.
"""
import os
import subprocess

import numpy as np
import pytest

import gpma_bfs_amd as G
from oracle import oracle as O
from tests.helpers import assert_invariants, assert_layout_equal, closed_form_live, protocol_batches

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_four_node_notes_traces():
    """Notes pp.2-4: N=4, insert 1->2, then +1->3 / -1->2."""
    g = G.GPMA(4)
    o = O.OracleGPMA(4)
    assert_layout_equal(g, o, "guards")
    k, v, r = g.copy_layout()
    F = 0xFFFFFFFF
    expect = [(0 << 32) | F, (1 << 32) | F, (2 << 32) | F, G.KEY_NONE, (3 << 32) | F, G.KEY_MAX, G.KEY_MAX, G.KEY_NONE]
    assert [int(x) for x in k] == expect
    assert r.tolist() == [0, 0, 1, 2, 4]
    for batch, vals in [([(1 << 32) | 2], [1]), ([(1 << 32) | 3, (1 << 32) | 2], [1, 0])]:
        b = np.array(batch, np.uint64)
        vv = np.array(vals, np.uint8)
        G.update_gpma(g, b, vv)
        o.update(b, vv)
        assert_layout_equal(g, o, f"batch {batch}")
    k, v, r = g.copy_layout()
    assert r.tolist() == [0, 0, 4, 5, 8]
    assert g.get_size() == 16 and g.segment_length == 4 and g.tree_height == 2


@pytest.mark.parametrize("seed,N,draw_scale", [(3, 200, 12), (5, 1000, 14), (11, 64, 10)])
def test_small_protocol_layout_exact(seed, N, draw_scale):
    rng_src, rng_dst = G.gen_rmat(int(np.ceil(np.log2(N))), 1 << draw_scale, seed=seed)
    m = (rng_src < N) & (rng_dst < N)
    src, dst = rng_src[m], rng_dst[m]
    keys = G.make_keys(src, dst)
    g = G.GPMA(N)
    o = O.OracleGPMA(N)
    assert_layout_equal(g, o, "guards")
    half = len(keys) // 2
    ones = np.ones(half, np.uint8)
    G.update_gpma(g, keys[:half], ones)
    o.update(keys[:half], ones)
    assert_layout_equal(g, o, "window")
    assert np.array_equal(g.export_live(), closed_form_live(keys, N, 0, num_slide=20))
    for i, b, vals in protocol_batches(keys, 20, num_slide=20):
        G.update_gpma(g, b, vals)
        o.update(b, vals)
        assert_layout_equal(g, o, f"slide {i}")
        assert np.array_equal(g.export_live(), closed_form_live(keys, N, i + 1, num_slide=20)), f"slide {i}"
    assert_invariants(g)
    for start in (0, 1, N - 1):
        lv, reach = G.gpma_bfs(g, start)
        olv, oreach = o.bfs(start)
        assert reach == oreach and np.array_equal(lv, olv), f"bfs from {start}"


def _model_check(g, model, N, tag):
    """The live keys and their stored value bytes against a dict model
    (guards and the two KEY_MAX sentinels added)."""
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    expect = np.sort(np.concatenate([np.array(list(model), np.uint64), guards,
                                     np.array([G.KEY_MAX, G.KEY_MAX], np.uint64)]))
    assert np.array_equal(g.export_live(), expect), f"{tag}: live set"
    k, v, _ = g.copy_layout()
    sel = (k != np.uint64(G.KEY_NONE)) & (v != 0) & ((k & np.uint64(0xFFFFFFFF)) != np.uint64(0xFFFFFFFF)) & \
        (k != np.uint64(G.KEY_MAX))
    stored = dict(zip(k[sel].tolist(), v[sel].tolist()))
    assert stored == model, f"{tag}: stored value bytes"


def test_random_mixed_batches_layout_exact():
    """Random batches mixing every kind of record: inserts of new keys,
    modifications of live keys to other non-zero values, deletes of live
    keys, deletes of absent keys and revivals of tombstones (gpma.cuh:188-199
    and the leaf merge). One record per key per batch: the layout equals
    the oracle slot for slot after every batch, and the live set and stored
    value bytes equal a model. Then batches with repeated keys (the last
    record of a key applies, a documented difference from the reference's
    parallel locate) and fused batches (gpma_update_batches), against the
    model."""
    rng = np.random.default_rng(41)
    N = 300
    g = G.GPMA(N)
    o = O.OracleGPMA(N)
    model = {}

    def rand_keys(n):
        a = rng.integers(0, N, n).astype(np.uint64)
        b = rng.integers(0, N, n).astype(np.uint64)
        return ((a << np.uint64(32)) | b)[a != b]

    def batch(dups):
        live = np.array(sorted(model), np.uint64)
        ks, vs = [], []
        new = rand_keys(int(rng.integers(1, 500)))
        ks.append(new)
        vs.append(rng.choice(np.array([1, 2, 7], np.uint8), len(new)))
        if len(live):
            d = rng.choice(live, size=min(len(live), int(rng.integers(0, 350))), replace=False)
            ks.append(d)
            vs.append(np.zeros(len(d), np.uint8))
            m = rng.choice(live, size=min(len(live), int(rng.integers(0, 60))), replace=False)
            ks.append(m)
            vs.append(rng.choice(np.array([3, 9], np.uint8), len(m)))
        a = rand_keys(int(rng.integers(0, 60)))
        ks.append(a)
        vs.append(np.zeros(len(a), np.uint8))
        k, v = np.concatenate(ks), np.concatenate(vs)
        if not dups:
            _, first = np.unique(k, return_index=True)
            first.sort()
            k, v = k[first], v[first]
        else:  # repeat some of the batch's keys with other values
            r = rng.choice(len(k), size=min(len(k), 40), replace=False)
            k = np.concatenate([k, k[r]])
            v = np.concatenate([v, rng.choice(np.array([0, 1, 5], np.uint8), len(r))])
        p = rng.permutation(len(k))
        return k[p], v[p]

    def apply(k, v):
        for kk, vv in zip(k.tolist(), v.tolist()):
            if vv:
                model[kk] = vv
            else:
                model.pop(kk, None)

    for step in range(40):
        k, v = batch(dups=False)
        G.update_gpma(g, k, v)
        o.update(k, v)
        apply(k, v)
        assert_layout_equal(g, o, f"batch {step}")
        _model_check(g, model, N, f"batch {step}")
    for step in range(10):
        k, v = batch(dups=True)
        assert len(np.unique(k)) < len(k)
        G.update_gpma(g, k, v)
        apply(k, v)
        _model_check(g, model, N, f"repeated keys {step}")
    for step in range(5):
        parts = [batch(dups=True) for _ in range(int(rng.integers(2, 9)))]
        G.update_gpma_batches(g, [p[0] for p in parts], [p[1] for p in parts])
        for k, v in parts:
            apply(k, v)
        _model_check(g, model, N, f"fused {step}")
    assert_invariants(g)
    g.close()


def test_c1_protocol():
    """R-MAT 2^16 / 2^20 (seed 1), full 100-slide protocol: live set after every
    batch equals the closed form; layout equals the oracle every 10 slides;
    BFS levels equal the reference cpu_baseline fixture."""
    N = 1 << 16
    src, dst = G.gen_rmat(16, 1 << 20, seed=1)
    keys = G.make_keys(src, dst)
    g = G.GPMA(N)
    o = O.OracleGPMA(N)
    half = len(keys) // 2
    ones = np.ones(half, np.uint8)
    G.update_gpma(g, keys[:half], ones)
    o.update(keys[:half], ones)
    assert_layout_equal(g, o, "window")
    fx = np.load(os.path.join(GOLDEN, "c1_rmat16_seed1.npz"))
    lv, reach = G.gpma_bfs(g, 0)
    assert np.array_equal(lv, fx["levels_before"]) and reach == int(fx["reach_before"])
    for i, b, vals in protocol_batches(keys, 100):
        G.update_gpma(g, b, vals)
        o.update(b, vals)
        live = g.export_live()
        assert np.array_equal(live, closed_form_live(keys, N, i + 1)), f"slide {i}"
        if i % 10 == 9:
            assert_layout_equal(g, o, f"slide {i}")
    assert_invariants(g)
    lv, reach = G.gpma_bfs(g, 0)
    assert np.array_equal(lv, fx["levels_after"]) and reach == int(fx["reach_after"])


def test_reference_demo_protocol_through_the_facade(tmp_path):
    """The reference demo's protocol (gpma_bfs_demo.cu:6-111), compiled against
    the C++ facade (GPMA_multidev<c,1>::update_batch, gpma_bfs(multidev, ...))
    and run on C1's graph written as an "N E" text file: its printed
    reachable counts before and after the 100 slides equal the reference
    cpu_baseline fixture."""
    demo = os.path.join(os.path.dirname(GOLDEN), "..", "gpma_bfs_amd", "demo", "gpma_bfs_demo")
    src, dst = G.gen_rmat(16, 1 << 20, seed=1)
    f = tmp_path / "c1.txt"
    G.write_edge_text(f, 1 << 16, src, dst)
    p = subprocess.run([demo, str(f), "0"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert f"node_num: {1 << 16}, edge_num: {len(src)}" in p.stdout
    reach = [int(line.rsplit(":", 1)[1]) for line in p.stdout.splitlines() if "number of reachable nodes" in line]
    fx = np.load(os.path.join(GOLDEN, "c1_rmat16_seed1.npz"))
    assert reach == [int(fx["reach_before"]), int(fx["reach_after"])], p.stdout


def test_update_semantics_edge_cases():
    N = 50
    g = G.GPMA(N)
    o = O.OracleGPMA(N)
    K = lambda s, d: (s << 32) | d  # noqa: E731
    # empty batch
    G.update_gpma(g, np.zeros(0, np.uint64), np.zeros(0, np.uint8))
    # delete of an absent key is a no-op (U1)
    b = np.array([K(1, 2)], np.uint64)
    G.update_gpma(g, b, np.zeros(1, np.uint8))
    o.update(b, np.zeros(1, np.uint8))
    assert_layout_equal(g, o, "absent delete")
    # insert, delete (tombstone), revive in place (U2)
    for vals in ([1], [0], [1]):
        G.update_gpma(g, b, np.array(vals, np.uint8))
        o.update(b, np.array(vals, np.uint8))
        assert_layout_equal(g, o, f"op {vals}")
    live = g.export_live()
    assert K(1, 2) in set(int(x) for x in live)
    # duplicate keys in one batch: the last record wins
    d = np.array([K(3, 4), K(3, 4)], np.uint64)
    G.update_gpma(g, d, np.array([1, 0], np.uint8))
    assert K(3, 4) not in set(int(x) for x in g.export_live())
    G.update_gpma(g, d, np.array([0, 1], np.uint8))
    assert K(3, 4) in set(int(x) for x in g.export_live())
    # out-of-range keys are skipped and reported; the rest of the batch applies
    bad = np.array([K(N, 1), K(2, N), K(5, 6)], np.uint64)
    rc = G.update_gpma(g, bad, np.ones(3, np.uint8), strict=False)
    assert rc == G.GPMA_ERR_KEY_RANGE
    live = set(int(x) for x in g.export_live())
    assert K(5, 6) in live and K(N, 1) not in live
    assert g.stats()["invalid_skipped"] == 2
    assert_invariants(g)
    # BFS argument errors
    with pytest.raises(G.GpmaError):
        G.gpma_bfs(g, N)


def test_legacy_csr_bfs_matches():
    N = 300
    src, dst = G.gen_rmat(9, 4000, seed=21)
    m = (src < N) & (dst < N)
    keys = G.make_keys(src[m], dst[m])
    g = G.GPMA(N)
    G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    c = g.csr()
    lv_legacy = G.gpma_csr_bfs(c.keys, c.values, c.row_offset, N, len(keys), 0)
    lv, _ = G.gpma_bfs(g, 0)
    assert np.array_equal(lv, lv_legacy)
    olv, _ = O.bfs_edges(N, src[m], dst[m], 0)
    assert np.array_equal(lv, olv)


def test_bulk_path_matches_oracle():
    """Bulk load (reference significant_insert) with a lowered threshold."""
    N = 2048
    src, dst = G.gen_rmat(11, 40000, seed=9)
    keys = G.make_keys(src, dst)
    g = G.GPMA(N, bulk_threshold=5000)
    half = len(keys) // 2
    ones = np.ones(half, np.uint8)
    G.update_gpma(g, keys[:half], ones)
    assert g.stats()["bulk_loads"] == 1
    assert np.array_equal(g.export_live(), closed_form_live(keys, N, 0, num_slide=10))
    assert_invariants(g)
    for i, b, vals in protocol_batches(keys, 10, num_slide=10):
        G.update_gpma(g, b, vals)
        assert np.array_equal(g.export_live(), closed_form_live(keys, N, i + 1, num_slide=10))
    lv, reach = G.gpma_bfs(g, 0)
    E = len(keys)
    step = half // 10
    olv, oreach = O.bfs_edges(N, src[10 * step: half + 10 * step], dst[10 * step: half + 10 * step], 0)
    assert reach == oreach and np.array_equal(lv, olv)


def test_million_record_batches():
    """Batches in the millions through the plain path: a 4M-record window
    (under the bulk threshold, so it takes the ordinary sort, locate and
    rebalance) and three slides of 1M inserts + 1M deletes on 2^20 vertices;
    the live set equals the closed form after every batch and the final BFS
    equals the oracle."""
    N = 1 << 20
    src, dst = G.gen_rmat(20, 1 << 23, seed=17)
    keys = G.make_keys(src, dst)
    half = len(keys) // 2
    assert 4_000_000 < half < 5_000_000
    g = G.GPMA(N)
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    assert g.stats()["bulk_loads"] == 0
    assert np.array_equal(g.export_live(), closed_form_live(keys, N, 0, num_slide=4))
    for i, b, vals in protocol_batches(keys, 3, num_slide=4):
        G.update_gpma(g, b, vals)
        assert np.array_equal(g.export_live(), closed_form_live(keys, N, i + 1, num_slide=4)), f"slide {i}"
    assert_invariants(g)
    step = half // 4
    lv, reach = G.gpma_bfs(g, 0)
    olv, oreach = O.bfs_edges(N, src[3 * step: half + 3 * step], dst[3 * step: half + 3 * step], 0)
    assert reach == oreach and np.array_equal(lv, olv)


def test_sort_buckets_past_the_lds_size():
    """Batches of about 3.2M records: their 4,096 sort buckets average ~785
    items, and many pass the local pass's 1,024-item in-LDS size (those take
    its LSD passes over the bucket's range). Two slides through update_gpma
    and one as 8 fused batches through gpma_update_batches; the live set
    equals the closed form after each, and BFS equals the oracle."""
    N = 1 << 20
    src, dst = G.gen_rmat(20, 1 << 24, seed=19)
    keys = G.make_keys(src, dst)
    half = len(keys) // 2
    step = half // 5
    assert 2 * step > 640 * 4096 and 2 * step <= 4 << 20
    g = G.GPMA(N)
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    batches = list(protocol_batches(keys, 3, num_slide=5))
    for i, b, vals in batches[:2]:
        G.update_gpma(g, b, vals)
        assert np.array_equal(g.export_live(), closed_form_live(keys, N, i + 1, num_slide=5)), f"slide {i}"
    _, b, vals = batches[2]
    ins, dele = b[:step], b[step:]
    parts, pv = [], []
    for j in range(8):  # uneven inserts, then uneven deletes, in 8 batches
        a0, a1 = (j * step) // 8, ((j + 1) * step) // 8
        parts.append(np.concatenate([ins[a0:a1], dele[a0:a1]]))
        pv.append(np.concatenate([np.ones(a1 - a0, np.uint8), np.zeros(a1 - a0, np.uint8)]))
    G.update_gpma_batches(g, parts, pv)
    assert np.array_equal(g.export_live(), closed_form_live(keys, N, 3, num_slide=5)), "fused slide"
    assert_invariants(g)
    lv, reach = G.gpma_bfs(g, 0)
    olv, oreach = O.bfs_edges(N, src[3 * step: half + 3 * step], dst[3 * step: half + 3 * step], 0)
    assert reach == oreach and np.array_equal(lv, olv)
    g.close()


def test_wide_window_levels_match_oracle_layout():
    """The 4.2M-record window on 2^20 vertices (under the bulk threshold)
    climbs through rebalance levels whose windows (8,192 slots and more) are
    wider than one workgroup; all windows of such a level go through one set
    of device-wide passes. The layout (keys, live flags, row_offset) equals
    the oracle slot for slot after the window and after a 2M-record slide,
    and the oracle's level statistics confirm windows tried at every level
    9..17 and placed at some of them."""
    N = 1 << 20
    src, dst = G.gen_rmat(20, 1 << 23, seed=17)
    keys = G.make_keys(src, dst)
    half = len(keys) // 2
    g = G.GPMA(N)
    o = O.OracleGPMA(N)
    ones = np.ones(half, np.uint8)
    G.update_gpma(g, keys[:half], ones)
    o.update(keys[:half], ones)
    assert g.stats()["bulk_loads"] == 0
    assert_layout_equal(g, o, "window")
    st = o.level_stats()
    assert all(st[L][0] > 0 for L in range(9, 18)), st[9:18]
    assert st[9:18, 1].sum() > 0, st[9:18]
    step = half // 4
    b = np.concatenate([keys[half: half + step], keys[:step]])
    v = np.concatenate([np.ones(step, np.uint8), np.zeros(step, np.uint8)])
    G.update_gpma(g, b, v)
    o.update(b, v)
    assert_layout_equal(g, o, "slide")
    assert_invariants(g)
    g.close()


@pytest.mark.parametrize("pull_permille", [0, 1, 1000])
def test_bfs_hubs_and_direction_modes(pull_permille):
    """Hub vertices (ranges far above the 2048-slot split) and both expansion
    modes: 1 forces the streaming pull step at every level, 1000 disables it."""
    N = 20000
    rng = np.random.default_rng(17)
    hubs = np.array([0, 5, 77], np.uint32)
    src = [np.full(9000, h, np.uint32) for h in hubs]
    dst = [rng.choice(N, 9000, replace=False).astype(np.uint32) for _ in hubs]
    s2, d2 = G.gen_rmat(15, 60000, seed=23)
    m = (s2 < N) & (d2 < N)
    src = np.concatenate(src + [s2[m]])
    dst = np.concatenate(dst + [d2[m]])
    keys = np.unique(G.make_keys(src, dst))
    keys = keys[(keys >> np.uint64(32)) != (keys & np.uint64(0xFFFFFFFF))]
    src = (keys >> np.uint64(32)).astype(np.uint32)
    dst = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    g = G.GPMA(N, pull_permille=pull_permille)
    G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    # tombstone some edges so skipped slots are exercised
    G.update_gpma(g, keys[::7], np.zeros(len(keys[::7]), np.uint8))
    keep = np.ones(len(keys), bool)
    keep[::7] = False
    for start in (0, 5, 1234, N - 1):
        lv, reach = G.gpma_bfs(g, start)
        olv, oreach = O.bfs_edges(N, src[keep], dst[keep], start)
        assert reach == oreach and np.array_equal(lv, olv), f"start {start}"


@pytest.mark.parametrize("pull_permille", [0, 1, 1000])
def test_multi_source_bfs_matches_single_source(pull_permille):
    """64 sources in one traversal: every source's levels and reachable count
    equal the single-source BFS (and the oracle), including duplicate
    sources, hubs, tombstoned edges and both expansion modes."""
    N = 20000
    rng = np.random.default_rng(5)
    hubs = np.array([3, 900], np.uint32)
    src = [np.full(7000, h, np.uint32) for h in hubs]
    dst = [rng.choice(N, 7000, replace=False).astype(np.uint32) for _ in hubs]
    s2, d2 = G.gen_rmat(15, 90000, seed=31)
    m = (s2 < N) & (d2 < N)
    keys = np.unique(G.make_keys(np.concatenate(src + [s2[m]]), np.concatenate(dst + [d2[m]])))
    keys = keys[(keys >> np.uint64(32)) != (keys & np.uint64(0xFFFFFFFF))]
    g = G.GPMA(N, pull_permille=pull_permille)
    G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    G.update_gpma(g, keys[::9], np.zeros(len(keys[::9]), np.uint8))
    live = np.setdiff1d(keys, keys[::9])
    srcs = (live >> np.uint64(32)).astype(np.uint32)
    dsts = (live & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    starts = rng.integers(0, N, 64).astype(np.uint32)
    starts[5] = starts[6]  # a duplicate source
    starts[0], starts[1] = 3, 900
    lv, reach = G.gpma_bfs_multi(g, starts)
    for i, s in enumerate(starts):
        one, r1 = G.gpma_bfs(g, int(s))
        assert np.array_equal(lv[i], one) and reach[i] == r1, f"source {i} ({s})"
        if i % 8 == 0:
            olv, _ = O.bfs_edges(N, srcs, dsts, int(s))
            assert np.array_equal(lv[i], olv)
    lv2, reach2 = G.gpma_bfs_multi(g, starts[:7])
    assert np.array_equal(lv2, lv[:7]) and np.array_equal(reach2, reach[:7])


def test_multi_source_bfs_large_dense_levels():
    """4.2M vertices (R-MAT scale 22, 2^24 draws): on the widest levels every
    vertex's mask is scanned and each workgroup's share of vertices exceeds
    its LDS queue, so the next frontier is flushed in several pieces; every
    source still equals the single-source BFS, two equal the oracle."""
    N = 1 << 22
    s, d = G.gen_rmat(22, 1 << 24, seed=41)
    keys = np.unique(G.make_keys(s, d))
    keys = keys[(keys >> np.uint64(32)) != (keys & np.uint64(0xFFFFFFFF))]
    g = G.GPMA(N)
    G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    starts = np.random.default_rng(8).integers(0, N, 64).astype(np.uint32)
    starts[0] = 0  # the R-MAT hub: the widest traversal
    lv, reach = G.gpma_bfs_multi(g, starts)
    assert reach[0] > N // 4
    for i, st in enumerate(starts):
        one, r1 = G.gpma_bfs(g, int(st))
        assert np.array_equal(lv[i], one) and reach[i] == r1, f"source {i} ({st})"
    srcs = (keys >> np.uint64(32)).astype(np.uint32)
    dsts = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    for i in (0, 1):
        olv, oreach = O.bfs_edges(N, srcs, dsts, int(starts[i]))
        assert np.array_equal(lv[i], olv) and reach[i] == oreach


def test_fused_batches_equal_sequential():
    """gpma_update_batches over uneven batches leaves the same live keys (and
    BFS levels) as applying the batches one by one, including a key inserted
    in one batch and deleted in a later one, and a delete+reinsert."""
    N = 4096
    src, dst = G.gen_rmat(12, 80000, seed=8)
    keys = G.make_keys(src, dst)
    half = len(keys) // 2
    a = G.GPMA(N)
    b = G.GPMA(N)
    for g in (a, b):
        G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    rng = np.random.default_rng(3)
    ins, dele = half, 0
    bk, bv = [], []
    for size in rng.integers(10, 3000, 8):
        k = np.concatenate([keys[ins: ins + size], keys[dele: dele + size]])
        v = np.concatenate([np.ones(size, np.uint8), np.zeros(size, np.uint8)])
        ins += size
        dele += size
        bk.append(k)
        bv.append(v)
    # a key inserted early and deleted later, and one deleted then re-inserted
    bk.append(np.array([keys[half], keys[0]], np.uint64))
    bv.append(np.array([0, 1], np.uint8))
    for k, v in zip(bk, bv):
        G.update_gpma(a, k, v)
    G.update_gpma_batches(b, bk, bv)
    la, lb = a.export_live(), b.export_live()
    assert np.array_equal(la, lb)
    assert keys[half] not in set(la.tolist()) and keys[0] in set(la.tolist())
    for s in (0, 1, 100):
        assert np.array_equal(G.gpma_bfs(a, s)[0], G.gpma_bfs(b, s)[0])


def test_skewed_batches_sort_paths():
    """Batches whose updates crowd into few sort buckets (one hub source with
    thousands of edges, then a batch over two sources only): every sort path
    (comparison ranking, in-bucket digit passes, and the switch to the full
    LSD passes after a bucket above 16384 items) must leave the same live
    set as the oracle, layout included."""
    N = 40000
    rng = np.random.default_rng(7)
    g = G.GPMA(N)
    o = O.OracleGPMA(N)
    # hub 5 gets 2000 edges, then 20000 edges spread over sources 7 and 8
    hub = (np.uint64(5) << np.uint64(32)) | rng.permutation(N)[:2000].astype(np.uint64)
    two = np.unique((rng.integers(7, 9, 30000).astype(np.uint64) << np.uint64(32)) |
                    rng.integers(0, N, 30000).astype(np.uint64))[:20000]
    rng.shuffle(two)
    mixed = (rng.integers(0, N, 5000).astype(np.uint64) << np.uint64(32)) | rng.integers(0, N, 5000).astype(np.uint64)
    batches = [(hub, np.ones(len(hub), np.uint8)), (two, np.ones(len(two), np.uint8)),
               (mixed, np.ones(len(mixed), np.uint8)),
               (np.concatenate([hub[:700], two[:9000]]), np.zeros(9700, np.uint8))]
    for i, (k, v) in enumerate(batches):
        G.update_gpma(g, k, v)
        o.update(k, v)
        assert np.array_equal(g.export_live(), o.live()), f"batch {i}"
        assert_layout_equal(g, o, f"batch {i}")
        assert_invariants(g)


@pytest.mark.parametrize("knobs", [{"segment_length": 8}, {"segment_length": 32}, {"segment_length": 4},
                                   {"device_levels": 1}, {"device_levels": 8},
                                   {"sort_bucket_bits": 4}, {"sort_bucket_bits": 12},
                                   {"lower_leaf_permille": 150},
                                   {"upper_leaf_permille": 800, "upper_root_permille": 700, "lower_root_permille": 350},
                                   {"lower_leaf_permille": 63, "lower_root_permille": 250, "upper_root_permille": 600,
                                    "upper_leaf_permille": 980},
                                   {"lower_leaf_permille": 100, "lower_root_permille": 450, "upper_root_permille": 900,
                                    "upper_leaf_permille": 950, "segment_length": 32},
                                   {"sort_items_per_lane": 4}, {"sort_items_per_lane": 16},
                                   {"push_tile": 16}, {"ms_tile": 16}, {"push_tile": 32, "ms_tile": 40}])
def test_tuning_knobs_keep_results(knobs):
    """Tuning knobs change timing (and, for segment_length and the density
    thresholds, the layout) but never the live-key set or BFS levels: 20
    protocol slides on an R-MAT graph with the closed-form live set after
    every slide, and single- and 64-source BFS equal to a textbook BFS over
    the final window."""
    src, dst = G.gen_rmat(14, 1 << 17, seed=5)
    N = 1 << 14
    keys = G.make_keys(src, dst)
    g = G.GPMA(N, tuning=knobs)
    half = len(keys) // 2
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    assert np.array_equal(g.export_live(), closed_form_live(keys, N, 0, num_slide=20)), "window load"
    assert_invariants(g)
    for i, b, v in protocol_batches(keys, 20, num_slide=20):
        G.update_gpma(g, b, v)
        assert np.array_equal(g.export_live(), closed_form_live(keys, N, i + 1, num_slide=20)), f"slide {i}"
    assert_invariants(g)
    if knobs.get("segment_length"):
        assert g.segment_length == knobs["segment_length"]
    step = half // 20
    win = keys[20 * step: half + 20 * step]
    ws, wd = (win >> np.uint64(32)).astype(np.uint32), (win & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    for s0 in (0, 7, 1234):
        lv, reach = G.gpma_bfs(g, s0)
        elv, ereach = O.bfs_edges(N, ws, wd, s0)
        assert np.array_equal(lv, elv) and reach == ereach, f"start {s0}"
    starts = np.random.default_rng(9).integers(0, N, 64).astype(np.uint32)
    mlv, mreach = G.gpma_bfs_multi(g, starts)
    elv, ereach = O.bfs_edges_multi(N, ws, wd, starts)
    assert np.array_equal(mlv, elv) and np.array_equal(mreach, ereach)


@pytest.mark.parametrize("name", ["mi355x_powerlaw_1p6m.json", "mi355x_powerlaw_1p6m_batched.json"])
def test_committed_tuning_files_keep_results(name):
    """The autotuner's committed results (segment length, device levels, sort
    shape, density thresholds, BFS tiles) keep the closed-form live set over
    20 slides and equal textbook BFS, single- and 64-source."""
    from gpma_bfs_amd import tune

    knobs = tune.load(os.path.join(os.path.dirname(GOLDEN), "..", "gpma_bfs_amd", "tuned", name))
    test_tuning_knobs_keep_results(knobs)


def test_unsupported_tuning_is_rejected():
    for bad in ({"segment_length": 2}, {"segment_length": 12}, {"device_levels": 9}, {"sort_bucket_bits": 3},
                {"lower_leaf_permille": 60}, {"lower_root_permille": 500}, {"upper_root_permille": 950},
                {"upper_leaf_permille": 1001}, {"sort_items_per_lane": 5}, {"push_tile": 8}, {"ms_tile": 65},
                {"no_such_knob": 1}):
        with pytest.raises(G.GpmaError) as e:
            G.GPMA(64, tuning=bad)
        assert e.value.code == G.GPMA_ERR_ARG


def _chain_graph():
    """A 70,000-vertex chain 0 -> 1 -> ... -> 69,999 (depth 70,000 from 0)
    with 2,000 side branches (chain vertex 35k -> side vertex 70,000 + k ->
    chain vertex 35k + 1, never a shortcut) and 500 backward edges."""
    n_chain, n_side = 70000, 2000
    N = n_chain + n_side
    c = np.arange(n_chain - 1, dtype=np.uint32)
    k = np.arange(n_side, dtype=np.uint32)
    rng = np.random.default_rng(70)
    a = rng.integers(1, n_chain, 500).astype(np.uint32)
    b = (a - rng.integers(1, 1000, 500).clip(max=a)).astype(np.uint32)
    src = np.concatenate([c, 35 * k, n_chain + k, a])
    dst = np.concatenate([c + 1, n_chain + k, np.minimum(35 * k + 1, n_chain - 1), b])
    keys = np.unique(G.make_keys(src, dst))
    keys = keys[(keys >> np.uint64(32)) != (keys & np.uint64(0xFFFFFFFF))]
    return N, keys


def _edges(keys):
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)


@pytest.mark.timeout(600)
def test_multi_source_bfs_deeper_than_u16_levels():
    """Sources whose BFS depth passes 65,535 levels: the 64-source traversal
    (gpma_bfs_multi), the captured step (gpma_step, whose graph holds at most
    64 levels and continues on the host path) and a 2-request server group
    must all equal the oracle's textbook BFS, and gpma_bfs for a few sources."""
    N, keys = _chain_graph()
    g = G.GPMA(N)
    G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    rng = np.random.default_rng(71)
    starts = rng.integers(0, N, 64).astype(np.uint32)
    starts[:6] = [0, 1, 70000, 35000, 69999, 71999]
    starts[10] = starts[0]  # duplicate source
    s, d = _edges(keys)
    olv, oreach = O.bfs_edges_multi(N, s, d, starts)
    assert olv[0].max() == 70000
    lv, reach = G.gpma_bfs_multi(g, starts)
    for i in range(64):
        assert np.array_equal(lv[i], olv[i]) and reach[i] == oreach[i], f"source {i} ({starts[i]})"
    for i in (0, 2, 5):
        one, r1 = G.gpma_bfs(g, int(starts[i]))
        assert np.array_equal(one, lv[i]) and r1 == reach[i], f"gpma_bfs source {i}"
    # captured step: delete 100 backward edges, insert 100 new ones
    back = keys[(keys >> np.uint64(32)) > (keys & np.uint64(0xFFFFFFFF))][:100]
    a = rng.integers(2000, 69000, 100).astype(np.uint64)
    new = np.setdiff1d((a << np.uint64(32)) | (a - np.uint64(7)), keys)
    bk = np.concatenate([new, back])
    bv = np.concatenate([np.ones(len(new), np.uint8), np.zeros(len(back), np.uint8)])
    rc, lv2, reach2 = G.gpma_step(g, bk, bv, np.array([0, len(bk)], np.uint64), starts, return_levels=True)
    assert rc == G.GPMA_OK
    keys2 = np.union1d(np.setdiff1d(keys, back), new)
    s2, d2 = _edges(keys2)
    olv2, oreach2 = O.bfs_edges_multi(N, s2, d2, starts)
    assert np.array_equal(lv2, olv2) and np.array_equal(reach2, oreach2)
    assert G.gpma_step_stats(g)["bfs_tails"] + G.gpma_step_stats(g)["host_steps"] >= 1
    # server: a 2-request BFS group behind a long no-op update (deletes of absent keys)
    i = np.arange(1_000_000, dtype=np.uint64)
    absent = np.setdiff1d(((i % np.uint64(N)) << np.uint64(32)) | (np.uint64(N - 1) - i // np.uint64(N)), keys2)
    assert len(absent) > 900_000
    grouped = False
    for attempt in range(3):
        with G.GpmaServer(g, max_bfs_group=2, order="snapshot") as srv:
            tu = srv.submit_update(absent, np.zeros(len(absent), np.uint8))
            tb = [srv.submit_bfs(0), srv.submit_bfs(1)]
            assert srv.wait(tu, timeout_s=120)["status"] == G.GPMA_OK
            res = [srv.wait(t, timeout_s=300) for t in tb]
        for r, i in zip(res, (0, 1)):
            assert r["status"] == G.GPMA_OK
            assert np.array_equal(r["levels"], olv2[i]) and r["reachable"] == oreach2[i], f"server source {i}"
        grouped = all(r["group"] == 2 for r in res)
        if grouped:
            break
    assert grouped, "the two BFS requests never shared a traversal"
    g.close()


def test_multi_source_bfs_many_levels_per_vertex():
    """Vertices reached at more than the 8 event slots' distinct levels
    (the u16 overflow rows, depth below 65,535): 64 sources spread along a
    3,000-vertex chain with side branches, so a chain vertex is reached at up
    to 64 levels. N = 3,000 leaves a partial last 64-vertex group, so the
    output pass's full-group path and its per-lane path both run; 63 sources
    take the per-lane path everywhere. Every row equals the oracle."""
    n_chain, n_side = 2900, 100
    N = n_chain + n_side
    c = np.arange(n_chain - 1, dtype=np.uint32)
    k = np.arange(n_side, dtype=np.uint32)
    src = np.concatenate([c, 29 * k, n_chain + k])
    dst = np.concatenate([c + 1, n_chain + k, np.minimum(29 * k + 2, n_chain - 1)])
    keys = np.unique(G.make_keys(src, dst))
    g = G.GPMA(N)
    G.update_gpma(g, keys, np.ones(len(keys), np.uint8))
    s, d = _edges(keys)
    starts = (np.arange(64, dtype=np.uint32) * 45).astype(np.uint32)
    starts[63] = n_chain + 7  # a side vertex
    olv, oreach = O.bfs_edges_multi(N, s, d, starts)
    per_vertex = np.array([len(np.unique(col[col > 0])) for col in olv.T])
    assert per_vertex.max() > 8 and olv.max() < 65535
    lv, reach = G.gpma_bfs_multi(g, starts)
    assert np.array_equal(lv, olv) and np.array_equal(reach, oreach)
    lv63, reach63 = G.gpma_bfs_multi(g, starts[:63])
    assert np.array_equal(lv63, olv[:63]) and np.array_equal(reach63, oreach[:63])
    g.close()


def test_rebuild_allocation_failure_keeps_the_handle_usable():
    """A bulk rebuild whose new arrays cannot be allocated: (a) the new arrays
    do not fit beside the old ones, so the old go first and the rebuild
    succeeds; (b) they never fit: the live items stay at the old capacity,
    GPMA_ERR_OOM is returned, and the handle stays usable (re-applying the
    batch gives the closed form); (c) not even the old shape can be restored:
    every later call returns GPMA_ERR_OOM."""
    N = 3000
    src, dst = G.gen_rmat(12, 120000, seed=29)
    m = (src < N) & (dst < N)
    keys = G.make_keys(src[m], dst[m])
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    kmax = np.array([G.KEY_MAX, G.KEY_MAX], np.uint64)

    def expect(k):
        return np.sort(np.concatenate([k, guards, kmax]))

    def prepared(first, n_big):
        g = G.GPMA(N, bulk_threshold=5000)
        G.update_gpma(g, keys[:first], np.ones(first, np.uint8))
        # size every scratch buffer the big batch needs (a no-op batch of absent deletes) and the
        # rebuild's compaction buffer, so the injected failures hit the new arrays
        absent = (np.arange(n_big, dtype=np.uint64) % np.uint64(N) << np.uint64(32)) | np.uint64(N - 1)
        absent = np.setdiff1d(absent, keys)[:n_big]
        absent = np.resize(absent, n_big)
        G.update_gpma(g, absent, np.zeros(n_big, np.uint8))
        g.export_live()
        return g

    first, big = 2000, 30000
    batch = keys[first: first + big]
    ones = np.ones(len(batch), np.uint8)
    # (a)
    g = prepared(first, len(batch))
    c0 = g.get_size()
    g.test_fail_allocs(0, 1)
    G.update_gpma(g, batch, ones)
    g.test_fail_allocs(0, 0)
    assert g.get_size() > c0 and g.stats()["bulk_loads"] == 1
    assert np.array_equal(g.export_live(), expect(keys[: first + big]))
    assert_invariants(g)
    g.close()
    # (b)
    g = prepared(first, len(batch))
    g.test_fail_allocs(0, 2)
    with pytest.raises(G.GpmaError) as e:
        G.update_gpma(g, batch, ones)
    assert e.value.code == G.GPMA_ERR_OOM
    g.test_fail_allocs(0, 0)
    assert g.get_size() == c0
    live = g.export_live()
    assert np.all(np.isin(expect(keys[:first]), live)), "a live key was lost"
    assert_invariants(g)
    G.update_gpma(g, batch, ones)
    assert np.array_equal(g.export_live(), expect(keys[: first + big]))
    lv, reach = G.gpma_bfs(g, 0)
    s, d = _edges(keys[: first + big])
    olv, oreach = O.bfs_edges(N, s, d, 0)
    assert reach == oreach and np.array_equal(lv, olv)
    g.close()
    # (c)
    g = prepared(first, len(batch))
    g.test_fail_allocs(0, 3)
    with pytest.raises(G.GpmaError) as e:
        G.update_gpma(g, batch, ones)
    assert e.value.code == G.GPMA_ERR_OOM
    g.test_fail_allocs(0, 0)
    for call in (lambda: G.update_gpma(g, batch[:10], ones[:10]), lambda: G.gpma_bfs(g, 0), g.export_live,
                 g.live_digest):
        with pytest.raises(G.GpmaError) as e:
            call()
        assert e.value.code == G.GPMA_ERR_OOM
    g.close()
