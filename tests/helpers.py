"""Shared test helpers: the sliding-window protocol and layout comparison."""
import numpy as np

import gpma_bfs_amd as G
from oracle import oracle as O

KEY_NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def protocol_batches(keys, slides, num_slide=100):
    """Reference protocol (gpma_bfs_demo.cu:50-104): window = first half; slide
    i inserts keys[half+i*step, +step) (value 1) followed by deletes of
    keys[i*step, +step) (value 0), one batch per slide."""
    E = len(keys)
    half = E // 2
    step = half // num_slide
    vals = np.concatenate([np.ones(step, np.uint8), np.zeros(step, np.uint8)])
    for i in range(slides):
        b = np.concatenate([keys[half + i * step: half + (i + 1) * step], keys[i * step: (i + 1) * step]])
        yield i, b, vals


def closed_form_live(keys, N, slides_done, num_slide=100):
    """Live keys after `slides_done` slides: the window plus guards plus two KEY_MAX."""
    E = len(keys)
    half = E // 2
    step = half // num_slide
    win = keys[slides_done * step: half + slides_done * step]
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    kmax = np.array([G.KEY_MAX, G.KEY_MAX], np.uint64)
    return np.sort(np.concatenate([win, guards, kmax]))


def assert_layout_equal(g, o, tag=""):
    gk, gv, gr = g.copy_layout()
    ok, ov, orr = o.layout()
    ok, ov = O.normalize_layout(ok, ov)
    assert len(gk) == len(ok), f"{tag}: capacity {len(gk)} vs oracle {len(ok)}"
    if not np.array_equal(gk, ok):
        bad = np.nonzero(gk != ok)[0]
        raise AssertionError(f"{tag}: {len(bad)} key slots differ, first at {bad[:5]}: "
                             f"gpu {gk[bad[:5]]} oracle {ok[bad[:5]]}")
    live_g = (gk != KEY_NONE) & (gv != 0)
    live_o = (ok != KEY_NONE) & (ov != 0)
    assert np.array_equal(live_g, live_o), f"{tag}: live/tombstone flags differ"
    assert np.array_equal(gr, orr), f"{tag}: row_offset differs"
    st = g.stats()
    inf = o.info()
    assert (st["segment_length"], st["tree_height"]) == (inf["segment_length"], inf["tree_height"]), tag


def assert_invariants(g):
    """Size-independent properties of any valid layout."""
    k, v, r = g.copy_layout()
    occ = k[k != KEY_NONE]
    inc = occ[1:] > occ[:-1]
    kmax_pair = (occ[1:] == occ[:-1]) & (occ[1:] == np.uint64(G.KEY_MAX))
    assert np.all(inc | kmax_pair), "occupied keys not strictly increasing (apart from the KEY_MAX pair)"
    N = g.row_num
    guard_slots = r[1:].astype(np.int64)
    expect = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    assert np.array_equal(k[guard_slots], expect), "row_offset[v+1] does not point at guard v"
    assert r[0] == 0
    # with 16+ slots per segment every redistribution gives each segment at
    # least one item (density >= 0.08 > 1/16), so segment first keys increase
    S = g.stats()["segment_length"]
    if S >= 16:
        assert np.all(k[::S] != KEY_NONE), "a segment's first slot is empty"
