"""Writes the golden fixtures in this directory from the reference's own
cpu_baseline (compiled from /root/reference/cpu_baseline into oracle/_ref by
oracle/build.py). Run in the build container only:

    python tests/golden/make_golden.py

Each fixture holds the generator parameters, a SHA-256 of the generated edge
list (so a generator change is caught), and the reference's outputs: BFS
levels before and after the slides, reachable counts, and the live edge set
after the slides (in full for small graphs, as count + SHA-256 for C1).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import gpma_bfs_amd as G  # noqa: E402
from oracle import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make(name, N, scale, ndraws, seed, slides, start, full_live):
    src, dst = G.gen_rmat(scale, ndraws, seed=seed)
    keep = (src < N) & (dst < N)
    src, dst = src[keep], dst[keep]
    t, lb, la, live = O.run_ref_cpu_baseline(N, src, dst, slides, start)
    out = dict(N=N, scale=scale, ndraws=ndraws, seed=seed, slides=slides, start=start,
               edges_sha256=sha(G.make_keys(src, dst)), n_edges=len(src),
               levels_before=lb, levels_after=la, reach_before=t["reach_before"], reach_after=t["reach_after"],
               live_count=len(live), live_sha256=sha(live))
    if full_live:
        out["live_after"] = live
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, t)


if __name__ == "__main__":
    make("c1_rmat16_seed1.npz", 1 << 16, 16, 1 << 20, 1, 100, 0, False)
    make("tiny_rmat8_seed4.npz", 256, 8, 3000, 4, 100, 0, True)
    make("small_rmat12_seed2.npz", 4096, 12, 60000, 2, 100, 3, True)
