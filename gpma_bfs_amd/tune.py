"""Autotuner for the library's tuning knobs (gpma_config): segment length,
device-driven rebalance levels, the batch sort's bucket bits and items per
lane, the four density thresholds (reference gpma.cuh:40-43), the BFS
push/pull threshold and the push tile sizes of the single- and multi-source
BFS.

None of the knobs changes results: every candidate is checked against the
closed-form live-key set of the sliding window it ran (and BFS levels against
the default configuration) before its time counts. The search is coordinate
descent: one knob at a time over its candidates, the others held at the best
values so far, repeated until a sweep changes nothing. The result is a small
JSON file that GPMA(..., tuning=load(path)) takes.

    python -m gpma_bfs_amd.tune --out tuned.json            # power-law 1.63M-node graph
    python -m gpma_bfs_amd.tune --rmat-scale 20 --out t.json
    python -m gpma_bfs_amd.tune --batched --out tb.json     # captured batched steps

--batched tunes the batched form instead: steps of 8 fused log-uniform
1e3..1e6-record batches plus a 64-source BFS, each one gpma_step (a captured
HIP graph), checked against the closed-form live set and the levels of the
default configuration.
"""
import argparse
import json
import time

import numpy as np

# 0 is the default of every knob: the reference formula for the segment
# length (16 at these sizes), the reference thresholds (80 / 420 / 920 / 840
# thousandths), sizes from the batch for the sort, 64-vertex push tiles.
# Threshold combinations the library refuses (bounds that do not nest) are
# skipped by the search.
DENSITY_SPACE = {
    "lower_leaf_permille": [0, 120, 200],
    "upper_leaf_permille": [0, 850, 980],
    "lower_root_permille": [0, 300],
    "upper_root_permille": [0, 700, 900],
}
UPDATE_SPACE = {
    "segment_length": [0, 8, 32],
    "device_levels": [1, 2, 3, 4, 5, 6],
    "sort_bucket_bits": [0, 9, 10, 11, 12],
    "sort_items_per_lane": [0, 4, 8, 16],
    **DENSITY_SPACE,
}
BFS_SPACE = {"pull_permille": [0, 50, 100, 500], "push_tile": [0, 16, 32]}
BATCHED_SPACE = {
    "segment_length": [0, 8, 32],
    "sort_bucket_bits": [0, 10, 11, 12],
    "sort_items_per_lane": [0, 4, 8, 16],
    "pull_permille": [0, 500],
    "ms_tile": [0, 16, 32],
    **DENSITY_SPACE,
}


def coordinate_descent(measure, space, start=None, max_sweeps=3):
    """Minimise measure(config) over the product of space's value lists, one
    knob at a time. measure returns seconds (lower is better) or None for an
    invalid configuration. Returns (best config, best time, evaluations)."""
    best = dict(start or {k: v[0] for k, v in space.items()})
    cache = {}

    def cost(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in cache:
            cache[key] = measure(dict(cfg))
        return cache[key]

    best_t = cost(best)
    if best_t is None:
        raise RuntimeError("the starting configuration failed its check")
    for _ in range(max_sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                t = cost(cand)
                if t is not None and t < best_t:
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    return best, best_t, len(cache)


def _closed_form(G, win, N):
    guards = (np.arange(N, dtype=np.uint64) << np.uint64(32)) | np.uint64(0xFFFFFFFF)
    return np.sort(np.concatenate([win, guards, np.array([G.KEY_MAX, G.KEY_MAX], np.uint64)]))


def tune_updates(G, N, keys, slides=95, warmup=3, step=None, space=UPDATE_SPACE, log=print):
    """Time `slides` protocol slides (window = first half of keys, step =
    inserts of the next edges + deletes of the oldest) per candidate; returns
    (best knobs, seconds per slide, evaluations). The default covers the
    whole 100-slide protocol: the rebalance depth grows as the window slides
    (level 1-2 at first, 4-5 by the end on the power-law 1.63M-node graph),
    so a few early slides would favour too few device-driven levels."""
    half = len(keys) // 2
    step = step or max(1, half // 100)
    nsl = warmup + slides
    nb = 2 * step
    batches = np.empty(nsl * nb, np.uint64)
    for i in range(nsl):
        batches[i * nb: i * nb + step] = keys[half + i * step: half + (i + 1) * step]
        batches[i * nb + step: (i + 1) * nb] = keys[i * step: (i + 1) * step]
    vals = np.tile(np.concatenate([np.ones(step, np.uint8), np.zeros(step, np.uint8)]), nsl)
    dk = G.DeviceBuffer.from_numpy(batches)
    dv = G.DeviceBuffer.from_numpy(vals)
    expect = _closed_form(G, keys[nsl * step: half + nsl * step], N)

    def measure(cfg):
        try:
            g = G.GPMA(N, tuning=cfg)
        except G.GpmaError:
            return None
        try:
            G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
            for i in range(warmup):
                G.update_gpma_device(g, dk.ptr + i * nb * 8, dv.ptr + i * nb, nb)
            g.sync()
            a = time.perf_counter()
            for i in range(warmup, nsl):
                G.update_gpma_device(g, dk.ptr + i * nb * 8, dv.ptr + i * nb, nb)
            g.sync()
            t = (time.perf_counter() - a) / slides
            ok = np.array_equal(g.export_live(), expect)
        finally:
            g.close()
        log(f"tune: {cfg} -> {t * 1e3:.3f} ms/slide{'' if ok else ' (WRONG live set: rejected)'}")
        return t if ok else None

    try:
        return coordinate_descent(measure, space)
    finally:
        dk.free()
        dv.free()


def tune_bfs(G, N, keys, knobs, space=BFS_SPACE, reps=3, log=print):
    """BFS direction threshold on the loaded window; levels must equal the
    default configuration's."""
    half = len(keys) // 2
    g0 = G.GPMA(N, tuning=knobs)
    G.update_gpma(g0, keys[:half], np.ones(half, np.uint8))
    src = int(keys[0] >> np.uint64(32))
    ref, _ = G.gpma_bfs(g0, src)
    g0.close()

    def measure(cfg):
        try:
            g = G.GPMA(N, tuning=dict(knobs, **cfg))
        except G.GpmaError:
            return None
        try:
            G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
            lv, _ = G.gpma_bfs(g, src)
            if not np.array_equal(lv, ref):
                return None
            ts = []
            for _ in range(reps):
                a = time.perf_counter()
                G.gpma_bfs_count(g, src)
                ts.append(time.perf_counter() - a)
        finally:
            g.close()
        t = float(np.median(ts))
        log(f"tune: {cfg} -> BFS {t * 1e3:.3f} ms")
        return t

    return coordinate_descent(measure, space)


def batched_steps(N, keys, steps, fused=8, nsrc=64, seed=2):
    """Step inputs of the batched form: per step, `fused` batches of
    log-uniform 1e3..1e6 records (half inserts of the next window edges, half
    deletes of the oldest) and `nsrc` uniform sources. Returns (list of
    (keys, values, offsets, starts), inserted, deleted)."""
    half = len(keys) // 2
    rs = np.random.default_rng(seed)
    sizes = np.exp(rs.uniform(np.log(1e3), np.log(1e6), steps * fused)).astype(np.int64) // 2 * 2
    starts = rs.integers(0, N, (steps, nsrc)).astype(np.uint32)
    ins, dele, out = half, 0, []
    for i in range(steps):
        ks, vs, offs = [], [], [0]
        for b in range(fused):
            h = int(sizes[i * fused + b] // 2)
            ks += [keys[ins: ins + h], keys[dele: dele + h]]
            vs += [np.ones(h, np.uint8), np.zeros(h, np.uint8)]
            ins += h
            dele += h
            offs.append(offs[-1] + 2 * h)
        out.append((np.concatenate(ks), np.concatenate(vs), np.array(offs, np.uint64), starts[i]))
    if ins > len(keys):
        raise RuntimeError("not enough edges for the batched steps")
    return out, ins, dele


def tune_batched(G, N, keys, steps=12, warmup=2, space=BATCHED_SPACE, log=print):
    """Time `steps` captured batched steps per candidate (graphs captured
    before the timed steps); a candidate counts only if its live set equals
    the closed form and its last step's levels equal the default
    configuration's. Returns (best knobs, seconds per step, evaluations)."""
    half = len(keys) // 2
    plan, ins, dele = batched_steps(N, keys, warmup + steps)
    dk = [G.DeviceBuffer.from_numpy(p[0]) for p in plan]
    dv = [G.DeviceBuffer.from_numpy(p[1]) for p in plan]
    dlev = G.DeviceBuffer(64 * N * 4)
    expect = _closed_form(G, keys[dele: ins], N)
    max_records = max(int(p[2][-1]) for p in plan)
    ref = {}

    def measure(cfg):
        try:
            g = G.GPMA(N, tuning=cfg)
        except G.GpmaError:
            return None
        try:
            G.update_gpma(g, keys[:half], np.ones(half, np.uint8))
            G.gpma_step_prepare(g, max_records, 64, dlev.ptr)
            for i in range(warmup):
                G.gpma_step(g, dk[i].ptr, dv[i].ptr, plan[i][2], plan[i][3], levels_device_ptr=dlev.ptr)
            G.gpma_step_prepare(g, max_records, 64, dlev.ptr)  # at the BFS depth the warm-up found
            g.sync()
            a = time.perf_counter()
            for i in range(warmup, warmup + steps):
                G.gpma_step(g, dk[i].ptr, dv[i].ptr, plan[i][2], plan[i][3], levels_device_ptr=dlev.ptr)
            g.sync()
            t = (time.perf_counter() - a) / steps
            ok = np.array_equal(g.export_live(), expect)
            lv = dlev.to_numpy(np.uint32, 64 * N)
            if "levels" not in ref:
                ref["levels"] = lv
            ok = ok and np.array_equal(lv, ref["levels"])
        finally:
            g.close()
        log(f"tune: {cfg} -> {t * 1e3:.3f} ms/step{'' if ok else ' (WRONG result: rejected)'}")
        return t if ok else None

    try:
        return coordinate_descent(measure, space)
    finally:
        for b in dk + dv + [dlev]:
            b.free()


def load(path):
    """Knobs from a tuning file (for GPMA(..., tuning=...))."""
    with open(path) as f:
        d = json.load(f)
    return {k: int(v) for k, v in d["knobs"].items()}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--rmat-scale", type=int, default=0, help="R-MAT graph (2^scale nodes, 16x draws) instead")
    ap.add_argument("--slides", type=int, default=95)
    ap.add_argument("--batched", action="store_true", help="tune the captured batched step instead")
    args = ap.parse_args()
    import gpma_bfs_amd as G

    if args.rmat_scale:
        N = 1 << args.rmat_scale
        src, dst = G.gen_rmat(args.rmat_scale, 16 << args.rmat_scale, seed=1)
        workload = f"R-MAT scale {args.rmat_scale}, {16 << args.rmat_scale} draws, seed 1"
    else:
        N, E = 1632803, 30622564
        src, dst = G.gen_powerlaw(N, E, alpha=2.1, soft_cap=2800.0, p1=0.24, seed=0)
        workload = "power-law N=1,632,803 E=30,622,564 (alpha 2.1, soft cap 2800, p1 0.24, seed 0)"
    keys = G.make_keys(src, dst)
    t0 = time.time()
    if args.batched:
        knobs, t_step, n_eval = tune_batched(G, N, keys)
        out = {"workload": workload + "; batched steps: 8 fused log-uniform 1e3..1e6-record batches + "
                                      "64-source BFS (seed 2), one captured graph per step",
               "knobs": knobs, "step_ms": t_step * 1e3, "evaluations": n_eval, "seconds": time.time() - t0}
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out))
        return
    knobs, t_upd, n_upd = tune_updates(G, N, keys, slides=args.slides)
    bfs_knobs, t_bfs, n_bfs = tune_bfs(G, N, keys, knobs)
    knobs.update(bfs_knobs)
    out = {"workload": workload, "knobs": knobs, "slide_ms": t_upd * 1e3, "bfs_ms": t_bfs * 1e3,
           "evaluations": n_upd + n_bfs, "seconds": time.time() - t0}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
