"""GPU: phase timing (gpma_profile / gpma_get_profile) reports time for the
phases that ran, honours the phase mask, and leaves results unchanged."""
import numpy as np
import pytest

import gpma_bfs_amd as G
from tests.helpers import closed_form_live, protocol_batches

pytestmark = pytest.mark.gpu


def test_phase_timing_reports_and_keeps_results():
    N = 5000
    src, dst = G.gen_rmat(13, 80000, seed=5)
    m = (src < N) & (dst < N)
    keys = G.make_keys(src[m], dst[m])
    half = len(keys) // 2
    a, b = G.GPMA(N), G.GPMA(N)
    for g in (a, b):
        G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
    a.profile(True, reset=True)
    slides = list(protocol_batches(keys, 10, num_slide=10))
    for i, k, v in slides[:5]:
        G.update_gpma(a, k, v)
        G.update_gpma(b, k, v)
    rep = a.profile_report()
    # with 16 or 32 slots per segment the locate and the leaf rebalance run
    # as one kernel, timed as the locate phase
    fused = a.stats()["segment_length"] in (16, 32)
    for phase in ("sort", "locate", "compact") + (() if fused else ("leaf",)):
        assert rep[phase]["ms"] > 0 and rep[phase]["launches"] >= 5, (phase, rep[phase])
    if fused:
        assert rep["leaf"]["launches"] == 0, rep["leaf"]
    # only the sort phase from here on
    a.profile_phases(["sort"])
    before = a.profile_report()
    for i, k, v in slides[5:]:
        G.update_gpma(a, k, v)
        G.update_gpma(b, k, v)
    after = a.profile_report()
    assert after["sort"]["ms"] > before["sort"]["ms"]
    for phase in ("locate", "compact", "leaf"):
        assert after[phase] == before[phase], phase
    a.profile_phases(None)
    lv, reach = G.gpma_bfs(a, 0)
    assert a.profile_report()["bfs_expand"]["ms"] > 0
    a.profile(False)
    lv_b, reach_b = G.gpma_bfs(b, 0)
    assert reach == reach_b and np.array_equal(lv, lv_b)
    expect = closed_form_live(keys, N, 10, num_slide=10)
    assert np.array_equal(a.export_live(), expect) and np.array_equal(b.export_live(), expect)
    # multi-source traversal while profiling
    a.profile(True)
    starts = np.arange(0, 64 * 7, 7, dtype=np.uint32)
    lv_m, reach_m = G.gpma_bfs_multi(a, starts)
    a.profile(False)
    lv_mb, reach_mb = G.gpma_bfs_multi(b, starts)
    assert np.array_equal(lv_m, lv_mb) and np.array_equal(reach_m, reach_mb)
