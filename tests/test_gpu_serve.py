"""GPU tests of the request server (include/gpma_serve.h): every request's
result equals a sequential replay of the same requests in ticket order with
the plain entry points (update_gpma, gpma_bfs), whatever the fusing and
grouping the server chose."""
import threading

import numpy as np
import pytest

import gpma_bfs_amd as G

pytestmark = pytest.mark.gpu


def _stream(seed, N, nreq, edges):
    """Requests of one submitter: ('u', keys, values) or ('b', start)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nreq):
        if rng.random() < 0.6:
            n = int(rng.integers(1, 3000))
            idx = rng.integers(0, len(edges), n)
            out.append(("u", edges[idx], (rng.random(n) < 0.7).astype(np.uint8)))
        else:
            out.append(("b", int(rng.integers(0, N))))
    return out


@pytest.mark.parametrize("order,group,fused", [("strict", 0, 0), ("strict", 4, 5000), ("snapshot", 0, 0),
                                               ("snapshot", 4, 5000)])
def test_server_equals_sequential_replay(order, group, fused):
    N = 2048
    src, dst = G.gen_rmat(11, 40000, seed=9)
    edges = G.make_keys(src, dst)
    base = edges[: len(edges) // 2]
    g = G.GPMA(N)
    G.update_gpma(g, base, np.ones(len(base), np.uint8))
    streams = [_stream(100 + t, N, 40, edges) for t in range(3)]
    log = []  # (ticket, request, result)
    lock = threading.Lock()
    with G.GpmaServer(g, max_bfs_group=group, max_fused_records=fused, order=order) as srv:
        def submitter(reqs):
            mine = []
            for r in reqs:
                t = srv.submit_update(r[1], r[2]) if r[0] == "u" else srv.submit_bfs(r[1])
                mine.append((t, r))
            for t, r in mine:
                res = srv.wait(t, timeout_s=60)
                with lock:
                    log.append((t, r, res))
        th = [threading.Thread(target=submitter, args=(s,)) for s in streams]
        for x in th:
            x.start()
        for x in th:
            x.join()
        st = srv.stats()
    assert st["completed"] == st["submitted"] == sum(len(s) for s in streams)
    assert st["update_launches"] + st["bfs_launches"] <= st["completed"]
    log.sort(key=lambda x: x[0])
    upd = [(t, r) for t, r, _ in log if r[0] == "u"]
    upd_tickets = [t for t, _ in upd]
    for t, r, res in log:
        assert res["status"] == G.GPMA_OK
        if r[0] == "b":
            before = max([u for u in upd_tickets if u < t], default=0)
            if order == "strict":
                assert res["state_ticket"] == before, f"ticket {t}"
            else:
                assert res["state_ticket"] >= before, f"ticket {t}"
    # replay: apply updates in ticket order up to each BFS's state_ticket, then compare its result
    ref = G.GPMA(N)
    G.update_gpma(ref, base, np.ones(len(base), np.uint8))
    applied = 0
    for t, r, res in sorted([x for x in log if x[1][0] == "b"], key=lambda x: x[2]["state_ticket"]):
        while applied < len(upd) and upd[applied][0] <= res["state_ticket"]:
            G.update_gpma(ref, upd[applied][1][1], upd[applied][1][2])
            applied += 1
        lv, reach = G.gpma_bfs(ref, r[1])
        assert res["reachable"] == reach, f"ticket {t}"
        assert np.array_equal(res["levels"], lv), f"ticket {t}"
    while applied < len(upd):
        G.update_gpma(ref, upd[applied][1][1], upd[applied][1][2])
        applied += 1
    assert np.array_equal(g.export_live(), ref.export_live())


def test_server_admission_and_key_range():
    N = 512
    g = G.GPMA(N)
    with G.GpmaServer(g, max_fused_records=2000) as srv:
        big = np.arange(3000, dtype=np.uint64)
        with pytest.raises(G.GpmaError) as e:
            srv.submit_update(big, np.ones(3000, np.uint8))
        assert e.value.code == G.GPMA_ERR_CAPACITY
        keys = np.array([(1 << 32) | 2, (N << 32) | 1, (3 << 32) | 4], np.uint64)
        t = srv.submit_update(keys, np.ones(3, np.uint8))
        ok = srv.submit_update(np.array([(5 << 32) | 6], np.uint64), np.ones(1, np.uint8))
        b = srv.submit_bfs(1)
        assert srv.wait(t)["status"] == G.GPMA_ERR_KEY_RANGE
        assert srv.wait(ok)["status"] == G.GPMA_OK
        r = srv.wait(b)
        assert r["reachable"] == 2 and r["levels"][2] == 2
        assert srv.stats()["rejected"] == 1
    live = g.export_live()
    assert (1 << 32) | 2 in set(live.tolist()) and (3 << 32) | 4 in set(live.tolist())


def _replay_check(N, base, log):
    """Every result equals the same request run alone on the plain path: a
    fresh handle applies the update requests one at a time in ticket order
    and runs each BFS by itself after the update its state_ticket names.
    Returns the replayed handle."""
    log = sorted(log, key=lambda x: x[0])
    upd = [(t, r) for t, r, _ in log if r[0] == "u"]
    ref = G.GPMA(N)
    G.update_gpma(ref, base, np.ones(len(base), np.uint8))
    applied = 0
    for t, r, res in sorted([x for x in log if x[1][0] == "b"], key=lambda x: x[2]["state_ticket"]):
        while applied < len(upd) and upd[applied][0] <= res["state_ticket"]:
            G.update_gpma(ref, upd[applied][1][1], upd[applied][1][2])
            applied += 1
        lv, reach = G.gpma_bfs(ref, r[1])
        assert res["reachable"] == reach and np.array_equal(res["levels"], lv), f"ticket {t}"
    while applied < len(upd):
        G.update_gpma(ref, upd[applied][1][1], upd[applied][1][2])
        applied += 1
    return ref


@pytest.mark.parametrize("order", ["strict", "snapshot"])
def test_server_pipeline_same_alone_and_in_the_mix(order):
    """Four submitters mix host and device update requests with BFS requests;
    the server stages each round's records while the round before runs and
    its per-round records limit (6,000) defers requests to later rounds.
    Every result, whatever round and group it was served in, equals the same
    request run alone on the plain path, and the final arrays are equal."""
    N = 20000
    src, dst = G.gen_rmat(15, 300000, seed=19)
    m = (src < N) & (dst < N)
    edges = G.make_keys(src[m], dst[m])
    base = edges[: len(edges) // 2]
    g = G.GPMA(N)
    G.update_gpma(g, base, np.ones(len(base), np.uint8))
    log, lock, keep = [], threading.Lock(), []
    with G.GpmaServer(g, max_fused_records=6000, order=order) as srv:
        def submitter(seed):
            rng = np.random.default_rng(seed)
            mine = []
            for _ in range(60):
                if rng.random() < 0.6:
                    n = int(rng.integers(1, 5000))
                    k = edges[rng.integers(0, len(edges), n)]
                    v = (rng.random(n) < 0.6).astype(np.uint8)
                    if rng.random() < 0.5:
                        dk, dv = G.DeviceBuffer.from_numpy(k), G.DeviceBuffer.from_numpy(v)
                        with lock:
                            keep.extend([dk, dv])
                        t = srv.submit_update(dk.ptr, dv.ptr, n)
                    else:
                        t = srv.submit_update(k, v)
                    mine.append((t, ("u", k, v)))
                else:
                    s0 = int(rng.integers(0, N))
                    mine.append((srv.submit_bfs(s0), ("b", s0)))
            for t, r in mine:
                res = srv.wait(t, timeout_s=120)
                with lock:
                    log.append((t, r, res))
        th = [threading.Thread(target=submitter, args=(200 + i,)) for i in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        st = srv.stats()
    for b in keep:
        b.free()
    assert all(res["status"] == G.GPMA_OK for _, _, res in log)
    assert st["completed"] == st["submitted"] == 240
    assert st["deferred"] > 0 and st["staged_ahead"] > 0, st
    assert st["max_fused_records"] == 6000 and st["growth_reserve_bytes"] > 0
    if order == "snapshot":
        assert any(res["group"] > 1 for _, r, res in log if r[0] == "b"), "no BFS group was shared"
    ref = _replay_check(N, base, log)
    assert np.array_equal(g.export_live(), ref.export_live())
    ref.close()
    g.close()


def test_server_budget_reserves_rebuild_peak():
    """The default per-round limit follows the budget minus the rebuild
    reserve (28 bytes per slot of the array) and the BFS group; a budget that
    barely covers the reserve leaves the minimum of 1,024 records."""
    N = 4096
    g = G.GPMA(N)
    C = g.get_size()
    bfs = 1 * N * 4 + N * 384  # one BFS per group: its levels and the traversal scratch
    with G.GpmaServer(g, max_bfs_group=1, hbm_budget_bytes=28 * C + bfs + 146 * 50000) as srv:
        st = srv.stats()
        assert st["growth_reserve_bytes"] == 28 * C
        assert st["max_fused_records"] == 50000
    with G.GpmaServer(g, max_bfs_group=1, hbm_budget_bytes=28 * C + bfs) as srv:
        st = srv.stats()
        assert st["max_fused_records"] == 1024
        with pytest.raises(G.GpmaError) as e:
            srv.submit_update(np.zeros(1025, np.uint64), np.ones(1025, np.uint8))
        assert e.value.code == G.GPMA_ERR_CAPACITY
    g.close()
