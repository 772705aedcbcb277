"""GPU parity at the benchmark workloads' full sizes.

Each test runs one of the measured configurations end to end through the C
ABI and checks what the reference's protocol fixes (gpma_bfs_demo.cu:50-111):

* the live-key set after every batch equals the closed form of the window
  (guards and the two KEY_MAX sentinels included): by (count, sum mod 2^64,
  xor) after every batch, computed on the device, and exactly (sorted keys)
  at chosen points;
* BFS levels and reachable counts equal the oracle's textbook BFS over the
  window's edges, and the multi-source traversal equals single-source BFS.

Sizes: the power-law graph with 1,632,803 nodes / 30,622,564 edges (plain
protocol, then the batched captured steps), R-MAT 2^25 nodes / 2^29 draws
(bulk load, 2^21-record slides, a 64-source BFS) and R-MAT 2^24 / 2^28
(request server, every BFS answer replayed on the plain path).
"""
import os
import threading

import numpy as np
import pytest

import gpma_bfs_amd as G
from oracle import oracle as O
from tests.helpers import assert_invariants, closed_form_live, protocol_batches

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KMAX = np.array([G.KEY_MAX, G.KEY_MAX], np.uint64)


def _log(msg):
    """Progress on stderr (seen with pytest -s): the large cases run for minutes."""
    import sys

    print(f"[workloads] {msg}", file=sys.stderr, flush=True)


def _fixed_digest(N):
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    return G.combine_digests(G.key_digest(guards), G.key_digest(KMAX))


def _edges(keys):
    return (keys >> np.uint64(32)).astype(np.uint32), (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)


class RunningDigest:
    """Digest of a key multiset under inserts and deletes of present keys."""

    def __init__(self, keys):
        self.c, self.s, self.x = G.key_digest(keys)

    def apply(self, ins, dele):
        ci, si, xi = G.key_digest(ins)
        cd, sd, xd = G.key_digest(dele)
        self.c += ci - cd
        self.s = (self.s + si - sd) % (1 << 64)
        self.x ^= xi ^ xd

    def value(self):
        return self.c, self.s, self.x


def _powerlaw():
    N, E = 1632803, 30622564
    src, dst = G.gen_powerlaw(N, E, alpha=2.1, soft_cap=2800.0, p1=0.24, seed=0)
    return N, G.make_keys(src, dst)


@pytest.fixture(scope="module")
def powerlaw():
    return _powerlaw()


@pytest.mark.timeout(600)
def test_c2_powerlaw_full_protocol(powerlaw):
    """Plain drop-in at 1,632,803 nodes / 30,622,564 edges: bulk window, BFS
    from node 0, 100 slides of 153,112 inserts + 153,112 deletes; the live
    set after every slide (digest) and exactly after slides 0, 50 and 100;
    BFS from node 0 before and after, and 64 sources after, against the
    oracle."""
    N, keys = powerlaw
    E = len(keys)
    half, step = E // 2, (E // 2) // 100
    fixed = _fixed_digest(N)
    g = G.GPMA(N)
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    st = g.stats()
    assert st["bulk_loads"] == 1 and (st["capacity"], st["segment_length"], st["tree_height"]) == (1 << 25, 16, 21)
    assert np.array_equal(g.export_live(), closed_form_live(keys, N, 0))
    s, d = _edges(keys[:half])
    lv, reach = G.gpma_bfs(g, 0)
    olv, oreach = O.bfs_edges(N, s, d, 0)
    assert reach == oreach and np.array_equal(lv, olv), "BFS from node 0 before the slides"
    _log("C2: window and first BFS checked")
    run = RunningDigest(keys[:half])
    for i, b, vals in protocol_batches(keys, 100):
        G.update_gpma(g, b, vals)
        run.apply(b[:step], b[step:])
        assert g.live_digest() == G.combine_digests(run.value(), fixed), f"slide {i}"
        if i + 1 in (50, 100):
            assert np.array_equal(g.export_live(), closed_form_live(keys, N, i + 1)), f"slide {i}"
            _log(f"C2: slide {i + 1} exact")
    assert_invariants(g)
    win = keys[100 * step: half + 100 * step]
    s, d = _edges(win)
    lv, reach = G.gpma_bfs(g, 0)
    olv, oreach = O.bfs_edges(N, s, d, 0)
    assert reach == oreach and np.array_equal(lv, olv), "BFS from node 0 after 100 slides"
    starts = np.random.default_rng(2).integers(0, N, 64).astype(np.uint32)
    starts[0] = 0
    mlv, mreach = G.gpma_bfs_multi(g, starts)
    olv, oreach = O.bfs_edges_multi(N, s, d, starts)
    assert np.array_equal(mlv, olv) and np.array_equal(mreach, oreach), "64-source BFS after 100 slides"
    g.close()


def _uneven_step(rng, keys, ins, dele, fused=8, lo=1e3, hi=1e6):
    """`fused` batches of log-uniform sizes in [lo, hi] (even), each half
    inserts of the next window edges and half deletes of the oldest."""
    ks, vs, offs = [], [], [0]
    for _ in range(fused):
        h = max(1, int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) // 2)
        ks += [keys[ins: ins + h], keys[dele: dele + h]]
        vs += [np.ones(h, np.uint8), np.zeros(h, np.uint8)]
        ins += h
        dele += h
        offs.append(offs[-1] + 2 * h)
    assert ins <= len(keys)
    return np.concatenate(ks), np.concatenate(vs), np.array(offs, np.uint64), ins, dele


@pytest.mark.timeout(600)
def test_c3_batched_captured_steps(powerlaw):
    """Batched form on the same graph with the autotuned knobs: 12 steps of 8
    fused log-uniform 1e3..1e6-record batches and a 64-source BFS, through
    the captured graph (gpma_step). After every step the live set equals the
    closed form (digest), every source equals gpma_bfs; the last step's
    first two sources equal the oracle, and the final live set is exact."""
    from gpma_bfs_amd import tune

    N, keys = powerlaw
    half = len(keys) // 2
    knobs = tune.load(os.path.join(ROOT, "gpma_bfs_amd", "tuned", "mi355x_powerlaw_1p6m_batched.json"))
    g = G.GPMA(N, tuning=knobs)
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    fixed = _fixed_digest(N)
    rng = np.random.default_rng(2)
    ins, dele = half, 0
    for step in range(12):
        k, v, offs, ins, dele = _uneven_step(rng, keys, ins, dele)
        starts = rng.integers(0, N, 64).astype(np.uint32)
        rc, lv, reach = G.gpma_step(g, k, v, offs, starts, return_levels=True)
        assert rc == G.GPMA_OK
        assert g.live_digest() == G.combine_digests(G.key_digest(keys[dele:ins]), fixed), f"step {step}"
        _log(f"C3: step {step} ({len(k)} records) live set ok")
        for i, s0 in enumerate(starts):
            one, r1 = G.gpma_bfs(g, int(s0))
            assert np.array_equal(lv[i], one) and reach[i] == r1, f"step {step} source {i}"
    s, d = _edges(keys[dele:ins])
    olv, oreach = O.bfs_edges_multi(N, s, d, starts[:2])
    assert np.array_equal(lv[:2], olv) and np.array_equal(reach[:2], oreach)
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    assert np.array_equal(g.export_live(), np.sort(np.concatenate([keys[dele:ins], guards, KMAX])))
    sst = G.gpma_step_stats(g)
    assert sst["replays"] >= 10, sst
    assert_invariants(g)
    g.close()


@pytest.mark.timeout(900)
def test_c4_rmat25_bulk_slides_and_64_source_bfs():
    """R-MAT scale 25 (2^25 nodes), 2^29 draws (seed 3): bulk load of the half
    window (C = 2^29), 3 slides of 2^20 inserts + 2^20 deletes as one batch
    each, then a 64-source BFS. Live set digest after the load and every
    slide; every source equals gpma_bfs; two sources equal the oracle."""
    N = 1 << 25
    src, dst = G.gen_rmat(25, 1 << 29, seed=3)
    keys = G.make_keys(src, dst)
    del src, dst
    E = len(keys)
    _log(f"C4: {E} edges generated")
    half, step = E // 2, 1 << 20
    fixed = _fixed_digest(N)
    g = G.GPMA(N)
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    st = g.stats()
    # the 2^25 guards (above the 5,000,000 bulk threshold) and the window both take the bulk path
    assert st["bulk_loads"] == 2 and st["capacity"] == 1 << 29
    run = RunningDigest(keys[:half])
    assert g.live_digest() == G.combine_digests(run.value(), fixed), "bulk window"
    for i in range(3):
        ins = keys[half + i * step: half + (i + 1) * step]
        dele = keys[i * step: (i + 1) * step]
        b = np.concatenate([ins, dele])
        G.update_gpma(g, b, np.concatenate([np.ones(step, np.uint8), np.zeros(step, np.uint8)]))
        run.apply(ins, dele)
        assert g.live_digest() == G.combine_digests(run.value(), fixed), f"slide {i}"
        _log(f"C4: slide {i} live set ok")
    starts = np.random.default_rng(3).integers(0, N, 64).astype(np.uint32)
    starts[0] = 0
    lv, reach = G.gpma_bfs_multi(g, starts)
    _log("C4: 64-source BFS done")
    for i, s0 in enumerate(starts):
        one, r1 = G.gpma_bfs(g, int(s0))
        assert np.array_equal(lv[i], one) and reach[i] == r1, f"source {i}"
    _log("C4: every source equals gpma_bfs")
    win = keys[3 * step: half + 3 * step]
    del keys
    s, d = _edges(win)
    del win
    olv, oreach = O.bfs_edges_multi(N, s, d, starts[:2])
    assert np.array_equal(lv[:2], olv) and np.array_equal(reach[:2], oreach)
    g.close()


def _levels_hash(lv):
    try:
        import xxhash

        return xxhash.xxh3_128_hexdigest(np.ascontiguousarray(lv).view(np.uint8))
    except ImportError:
        import hashlib

        return hashlib.blake2b(np.ascontiguousarray(lv).view(np.uint8)).hexdigest()


@pytest.mark.timeout(900)
def test_c5_served_bfs_answers_equal_plain_replay():
    """Request server on R-MAT scale 24 (2^24 nodes), 2^28 draws (seed 4),
    half loaded: 2,000 requests arriving as a Poisson stream (70% update
    batches of log-uniform 1..1e5 records over a circular sliding window,
    30% single-source BFS from uniform sources). Every BFS answer (levels
    and reachable count) equals a plain-path replay: a second handle applies
    the update requests one by one in ticket order and runs gpma_bfs after
    the update each answer's state_ticket names. The served handle's final
    live set equals the closed form (digest)."""
    N = 1 << 24
    src, dst = G.gen_rmat(24, 1 << 28, seed=4)
    keys = G.make_keys(src, dst)
    del src, dst
    E = len(keys)
    half = E // 2
    rng = np.random.default_rng(4)
    R = 2000
    is_upd = rng.random(R) < 0.7
    sizes = np.where(is_upd, np.exp(rng.uniform(0, np.log(1e5), R)).astype(np.int64), 0) // 2 * 2
    sizes[is_upd & (sizes == 0)] = 2
    starts = rng.integers(0, N, R).astype(np.uint32)
    arrivals = np.cumsum(rng.exponential(1.0 / 5000.0, R))

    def circ(a, h):
        a %= E
        return keys[a: a + h] if a + h <= E else np.concatenate([keys[a:], keys[: a + h - E]])

    reqs, ins, dele = [], half, 0
    for j in range(R):
        if is_upd[j]:
            h = int(sizes[j] // 2)
            reqs.append((np.concatenate([circ(ins, h), circ(dele, h)]),
                         np.concatenate([np.ones(h, np.uint8), np.zeros(h, np.uint8)])))
            ins += h
            dele += h
        else:
            reqs.append(None)
    g = G.GPMA(N)
    G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    _log(f"C5: {E} edges, half loaded; serving {R} requests")
    tickets = [None] * R
    results = [None] * R
    ready = threading.Semaphore(0)

    def waiter():
        for j in range(R):
            ready.acquire()
            r = srv.wait(tickets[j], timeout_s=600)
            if "levels" in r:
                r["hash"] = _levels_hash(r.pop("levels"))
            results[j] = r

    import time

    with G.GpmaServer(g, order="snapshot") as srv:
        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for j in range(R):
            wait = arrivals[j] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            tickets[j] = srv.submit_update(*reqs[j]) if is_upd[j] else srv.submit_bfs(int(starts[j]))
            ready.release()
        th.join()
        sst = srv.stats()
    assert all(r["status"] == G.GPMA_OK for r in results)
    assert sst["completed"] == R
    _log(f"C5: served, {sst}")
    win = circ(dele, ins - dele)
    assert g.live_digest() == G.combine_digests(G.key_digest(win), _fixed_digest(N))
    g.close()
    # replay on the plain path, in ticket order
    h = G.GPMA(N)
    G.update_gpma(h, keys[:half], np.ones(half, np.uint8))
    order = np.argsort(np.array(tickets, np.uint64))
    by_state = {}
    for j in range(R):
        if not is_upd[j]:
            by_state.setdefault(results[j]["state_ticket"], []).append(j)
    checked = 0

    def check(t):
        nonlocal checked
        for j in by_state.get(t, ()):
            lv, r1 = G.gpma_bfs(h, int(starts[j]))
            assert r1 == results[j]["reachable"] and _levels_hash(lv) == results[j]["hash"], \
                f"request {j} (ticket {tickets[j]}, state {t})"
            checked += 1

    check(0)
    for j in order:
        if is_upd[j]:
            G.update_gpma(h, *reqs[j])
            check(tickets[j])
            if checked and checked % 100 == 0:
                _log(f"C5: {checked} BFS answers replayed")
    assert checked == int((~is_upd).sum())
    h.close()
