"""CPU tests that pin the oracle (oracle/gpma_oracle.cpp) before it is trusted.

Each check is an independent known answer, not a retyping of the oracle's
formulas: the reference notes' hand-traced layouts, closed-form thresholds and
array sizes at the survey's shapes, properties the redistribution must have
by construction, the closed-form sliding-window live set, scipy's shortest
paths, and golden vectors from the reference's own cpu_baseline.
"""
import hashlib
import os

import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.csgraph

import gpma_bfs_amd as G
from oracle import oracle as O
from tests.helpers import closed_form_live, protocol_batches

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
F = 0xFFFFFFFF
NONE = 0xFFFFFFFFFFFFFFFF
KMAX = 0xFFFFFFFFFFFFFFFE


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_fls_matches_bit_length():
    rng = np.random.default_rng(0)
    xs = [0, 1, 2, 3, 0x80000000, 0xFFFFFFFF, 0xF0000001, 0x50000001, 0x101] + list(rng.integers(0, 2**32, 200))
    for x in xs:
        assert O.fls(int(x)) == int(x).bit_length()


def test_density_table_known_values():
    """Bounds at S=16, h=21 (the 2^25-slot array) from the reference formulas."""
    lo, hi = O.density_table(16, 21)
    assert (lo[0], hi[0]) == (2, 14)
    assert (lo[1], hi[1]) == (4, 28)
    assert (lo[2], hi[2]) == (8, 56)
    assert (lo[3], hi[3]) == (17, 112)
    assert (lo[21], hi[21]) == (14092862, 28185722)
    # structural facts of any table: non-decreasing, lower <= upper, doubling caps
    for S, h in [(2, 1), (4, 3), (8, 10), (16, 16), (16, 25)]:
        lo, hi = O.density_table(S, h)
        assert np.all(lo <= hi)
        assert np.all(lo[1:] >= 2 * lo[:-1]) and np.all(hi[1:] <= 2 * hi[:-1])


@pytest.mark.parametrize("N,C,S,h", [(65536, 1 << 17, 16, 13), (1632803, 1 << 21, 16, 17), (4, 8, 4, 1)])
def test_guard_init_sizes(N, C, S, h):
    o = O.OracleGPMA(N)
    inf = o.info()
    assert (inf["capacity"], inf["segment_length"], inf["tree_height"]) == (C, S, h)
    assert np.array_equal(o.live(), closed_form_live(np.zeros(0, np.uint64), N, 0))


def test_four_node_notes_traces():
    """gpma.bfs.notes.pdf pp.2-4 (re-derived by hand in the survey)."""
    o = O.OracleGPMA(4)
    k, v, r = o.layout()
    assert [int(x) for x in k] == [(0 << 32) | F, (1 << 32) | F, (2 << 32) | F, NONE, (3 << 32) | F, KMAX, KMAX, NONE]
    assert r.tolist() == [0, 0, 1, 2, 4]
    o.update(np.array([(1 << 32) | 2], np.uint64), np.array([1], np.uint8))
    k, v, r = o.layout()
    assert o.info()["capacity"] == 16 and o.info()["segment_length"] == 4 and o.info()["tree_height"] == 2
    expect = {0: (0 << 32) | F, 1: (1 << 32) | 2, 4: (1 << 32) | F, 5: (2 << 32) | F, 8: (3 << 32) | F, 9: KMAX,
              12: KMAX}
    assert [int(x) for x in k] == [expect.get(i, NONE) for i in range(16)]
    assert r.tolist() == [0, 0, 4, 5, 8]
    # +1->3 and -1->2 in one batch: 1->2 becomes a tombstone and the leaf is
    # rebuilt with 1->3 packed after guard 0 (level-0 bounds [1, 3]).
    o.update(np.array([(1 << 32) | 3, (1 << 32) | 2], np.uint64), np.array([1, 0], np.uint8))
    k, v, r = o.layout()
    expect[1] = (1 << 32) | 3
    assert [int(x) for x in k] == [expect.get(i, NONE) for i in range(16)]


def test_redistribution_properties_brute_force():
    """Item i of m goes to slot (f*i/m)*S + (f*i%m)/f: slots strictly increase,
    every segment's items sit packed at its front, segment sizes differ by at
    most one, and segment s starts at item ceil(s*m/f)."""
    for S in (2, 4, 8, 16):
        for W in (S, 2 * S, 4 * S, 8 * S, 64 * S):
            f = W // S
            for m in range(1, W + 1, max(1, W // 97)):
                slots = [O.redispatch_slot(W, S, m, i) for i in range(m)]
                assert all(b > a for a, b in zip(slots, slots[1:]))
                seg = [s // S for s in slots]
                lane = [s % S for s in slots]
                counts = np.bincount(seg, minlength=f)
                assert counts.max() - counts.min() <= 1
                for i in range(m):
                    first = -(-seg[i] * m // f)
                    assert lane[i] == i - first


def test_faithful_root_overflow_reproduces_reference_drop():
    """With the fix off, N=50 loses every guard (gpma.cuh:1034-1044 ignores
    the failed root rebalance); with the product's fix the guards are kept."""
    o = O.OracleGPMA(50, fix_root=False)
    assert o.info()["dropped"] == 50
    o2 = O.OracleGPMA(50, fix_root=True)
    assert o2.info()["dropped"] == 0
    assert np.array_equal(o2.live(), closed_form_live(np.zeros(0, np.uint64), 50, 0))


def _rmat(N, scale, ndraws, seed):
    src, dst = G.gen_rmat(scale, ndraws, seed=seed)
    m = (src < N) & (dst < N)
    return src[m], dst[m]


@pytest.mark.parametrize("N,scale,ndraws,seed", [(256, 8, 3000, 4), (1000, 10, 20000, 5), (4096, 12, 60000, 2)])
def test_live_set_closed_form_every_batch(N, scale, ndraws, seed):
    src, dst = _rmat(N, scale, ndraws, seed)
    keys = G.make_keys(src, dst)
    o = O.OracleGPMA(N)
    half = len(keys) // 2
    o.update(keys[:half], np.ones(half, np.uint8))
    assert np.array_equal(o.live(), closed_form_live(keys, N, 0))
    for i, b, vals in protocol_batches(keys, 100):
        o.update(b, vals)
        assert np.array_equal(o.live(), closed_form_live(keys, N, i + 1)), f"slide {i}"
        k, v, r = o.layout()
        occ = k[k != np.uint64(NONE)]
        assert np.all((occ[1:] > occ[:-1]) | ((occ[1:] == occ[:-1]) & (occ[1:] == np.uint64(KMAX))))
        exp = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(F)
        assert np.array_equal(k[r[1:].astype(np.int64)], exp)
        # from 16 slots per segment no segment is left empty (segment first
        # keys strictly increase), which the product's locate relies on
        S = o.info()["segment_length"]
        if S >= 16:
            assert np.all(k[::S] != np.uint64(NONE)), f"slide {i}: empty segment at S={S}"


def test_bfs_matches_scipy_shortest_paths():
    N = 3000
    src, dst = _rmat(N, 12, 40000, 13)
    lv, reach = O.bfs_edges(N, src, dst, 0)
    A = scipy.sparse.csr_matrix((np.ones(len(src)), (src, dst)), shape=(N, N))
    d = scipy.sparse.csgraph.shortest_path(A, method="D", unweighted=True, indices=0)
    expect = np.where(np.isinf(d), 0, d + 1).astype(np.uint32)
    assert np.array_equal(lv, expect)
    assert reach == int(np.count_nonzero(expect))
    # the layout BFS of the oracle GPMA agrees with the edge-list BFS
    o = O.OracleGPMA(N)
    keys = G.make_keys(src, dst)
    o.update(keys, np.ones(len(keys), np.uint8))
    lv2, reach2 = o.bfs(0)
    assert np.array_equal(lv2, lv) and reach2 == reach


def test_multi_start_bfs_matches_scipy_and_chain_depth():
    """bfs_edges_multi (one adjacency, many starts) against scipy's shortest
    paths for several starts, and the closed form of a chain deeper than
    65,535 levels (level of chain vertex i from 0 is i + 1)."""
    N = 3000
    src, dst = _rmat(N, 12, 40000, 17)
    starts = np.array([0, 5, 5, 2999, 1234], np.uint32)
    lv, reach = O.bfs_edges_multi(N, src, dst, starts)
    A = scipy.sparse.csr_matrix((np.ones(len(src)), (src, dst)), shape=(N, N))
    d = scipy.sparse.csgraph.shortest_path(A, method="D", unweighted=True, indices=starts.astype(np.int64))
    expect = np.where(np.isinf(d), 0, d + 1).astype(np.uint32)
    assert np.array_equal(lv, expect)
    assert reach.tolist() == np.count_nonzero(expect, axis=1).tolist()
    n = 70000
    c_src = np.arange(n - 1, dtype=np.uint32)
    lvc, rc = O.bfs_edges_multi(n, c_src, c_src + 1, np.array([0, 60000], np.uint32))
    assert np.array_equal(lvc[0], np.arange(1, n + 1, dtype=np.uint32)) and rc[0] == n
    assert lvc[1][59999] == 0 and lvc[1][n - 1] == n - 60000 and rc[1] == n - 60000


@pytest.mark.parametrize("name", ["tiny_rmat8_seed4.npz", "small_rmat12_seed2.npz", "c1_rmat16_seed1.npz"])
def test_oracle_against_reference_cpu_baseline_golden(name):
    fx = np.load(os.path.join(GOLDEN, name))
    N = int(fx["N"])
    src, dst = _rmat(N, int(fx["scale"]), int(fx["ndraws"]), int(fx["seed"]))
    keys = G.make_keys(src, dst)
    assert sha(keys) == str(fx["edges_sha256"]), "generator output changed"
    start = int(fx["start"])
    o = O.OracleGPMA(N)
    half = len(keys) // 2
    o.update(keys[:half], np.ones(half, np.uint8))
    lv, reach = o.bfs(start)
    assert np.array_equal(lv, fx["levels_before"]) and reach == int(fx["reach_before"])
    for i, b, vals in protocol_batches(keys, int(fx["slides"])):
        o.update(b, vals)
    lv, reach = o.bfs(start)
    assert np.array_equal(lv, fx["levels_after"]) and reach == int(fx["reach_after"])
    live = o.live()
    edges = live[(live & np.uint64(F)) != np.uint64(F)]
    edges = edges[edges != np.uint64(KMAX)]
    assert len(edges) == int(fx["live_count"]) and sha(edges) == str(fx["live_sha256"])
    if "live_after" in fx:
        assert np.array_equal(edges, fx["live_after"])


def test_c1_resize_path_sizes():
    """C1's half window (477,545 edges < 5M) takes the level loop and a root
    resize to 2^20 slots (S=16, h=16)."""
    src, dst = G.gen_rmat(16, 1 << 20, seed=1)
    keys = G.make_keys(src, dst)
    o = O.OracleGPMA(1 << 16)
    half = len(keys) // 2
    o.update(keys[:half], np.ones(half, np.uint8))
    inf = o.info()
    assert (inf["capacity"], inf["segment_length"], inf["tree_height"]) == (1 << 20, 16, 16)
    assert inf["bulk_loads"] == 0 and inf["resizes"] == 2
