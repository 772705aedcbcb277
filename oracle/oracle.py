"""TEST INFRASTRUCTURE ONLY: Python handle on the CPU oracle (liborcl.so).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as a
checker or a timed baseline; the product (gpma_bfs_amd) never imports it.
"""
import ctypes
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborcl.so")
REF_BIN = os.path.join(HERE, "_ref", "ref_cpu_baseline")

_P = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run oracle/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = _P
        L.orc_create.argtypes = [ctypes.c_uint32]
        L.orc_destroy.argtypes = [_P]
        L.orc_set_fix_root.argtypes = [ctypes.c_int]
        L.orc_update.argtypes = [_P, _P, _P, ctypes.c_uint64]
        L.orc_info.argtypes = [_P, _P]
        L.orc_level_stats.argtypes = [_P, _P, ctypes.c_uint32]
        L.orc_level_stats.restype = ctypes.c_uint32
        L.orc_thresholds.argtypes = [_P, _P, _P]
        L.orc_layout.argtypes = [_P, _P, _P, _P]
        L.orc_live.argtypes = [_P, _P]
        L.orc_live.restype = ctypes.c_uint64
        L.orc_bfs.argtypes = [_P, ctypes.c_uint32, _P]
        L.orc_bfs.restype = ctypes.c_uint32
        L.orc_bfs_edges.argtypes = [ctypes.c_uint32, _P, _P, ctypes.c_uint64, ctypes.c_uint32, _P]
        L.orc_bfs_edges.restype = ctypes.c_uint32
        L.orc_bfs_edges_multi.argtypes = [ctypes.c_uint32, _P, _P, ctypes.c_uint64, _P, ctypes.c_uint32, _P]
        L.orc_bfs_edges_multi.restype = ctypes.c_uint64
        L.orc_redispatch_slot.argtypes = [ctypes.c_uint32] * 4
        L.orc_redispatch_slot.restype = ctypes.c_uint32
        L.orc_fls.argtypes = [ctypes.c_uint32]
        L.orc_fls.restype = ctypes.c_uint32
        L.orc_density_table.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _P, _P]
        L.orc_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def threads():
    """Threads the restatement's parallel loops use (OMP_NUM_THREADS)."""
    return int(lib().orc_threads())


class OracleGPMA:
    """CPU restatement of the reference GPMA (layout-faithful).

    fix_root=True applies the product's one deliberate deviation: a root
    rebalance that would fail after a resize grows the array instead of
    dropping the pending inserts (the reference drops them,
    gpma.cuh:1034-1044). It changes nothing unless that failure occurs."""

    def __init__(self, row_num, fix_root=True):
        self.row_num = row_num
        lib().orc_set_fix_root(1 if fix_root else 0)
        self.h = lib().orc_create(row_num)
        self.fix_root = fix_root

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def update(self, keys, values):
        lib().orc_set_fix_root(1 if self.fix_root else 0)
        keys = np.ascontiguousarray(keys, np.uint64)
        values = np.ascontiguousarray(values, np.uint8)
        lib().orc_update(self.h, keys.ctypes.data, values.ctypes.data, len(keys))

    def info(self):
        a = (ctypes.c_uint64 * 7)()
        lib().orc_info(self.h, a)
        return dict(capacity=a[0], segment_length=a[1], tree_height=a[2], row_num=a[3], dropped=a[4],
                    resizes=a[5], bulk_loads=a[6])

    def thresholds(self):
        h = self.info()["tree_height"]
        lo = np.zeros(h + 1, np.uint32)
        hi = np.zeros(h + 1, np.uint32)
        lib().orc_thresholds(self.h, lo.ctypes.data, hi.ctypes.data)
        return lo, hi

    def layout(self):
        C = self.info()["capacity"]
        k = np.empty(C, np.uint64)
        v = np.empty(C, np.uint8)
        r = np.empty(self.row_num + 1, np.uint32)
        lib().orc_layout(self.h, k.ctypes.data, v.ctypes.data, r.ctypes.data)
        return k, v, r

    def live(self):
        n = lib().orc_live(self.h, None)
        out = np.empty(n, np.uint64)
        lib().orc_live(self.h, out.ctypes.data)
        return out

    def bfs(self, start):
        lv = np.empty(self.row_num, np.uint32)
        reach = lib().orc_bfs(self.h, start, lv.ctypes.data)
        return lv, reach

    def level_stats(self):
        out = (ctypes.c_uint64 * 300)()
        L = lib().orc_level_stats(self.h, out, 100)
        return np.array(out[: 3 * L], np.uint64).reshape(L, 3)


def bfs_edges(N, src, dst, start):
    """Textbook BFS over an explicit edge list (levels, reachable)."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    lv = np.empty(N, np.uint32)
    reach = lib().orc_bfs_edges(N, src.ctypes.data, dst.ctypes.data, len(src), start, lv.ctypes.data)
    return lv, reach


def bfs_edges_multi(N, src, dst, starts):
    """Textbook BFS from each start over one edge list: levels[len(starts)][N]
    and the reachable counts."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    st = np.ascontiguousarray(starts, np.uint32)
    lv = np.empty((len(st), N), np.uint32)
    lib().orc_bfs_edges_multi(N, src.ctypes.data, dst.ctypes.data, len(src), st.ctypes.data, len(st),
                              lv.ctypes.data)
    return lv, (lv != 0).sum(axis=1).astype(np.int64)


def redispatch_slot(width, S, m, i):
    return lib().orc_redispatch_slot(width, S, m, i)


def fls(x):
    return lib().orc_fls(x)


def density_table(S, h):
    lo = np.zeros(h + 1, np.uint32)
    hi = np.zeros(h + 1, np.uint32)
    lib().orc_density_table(S, h, lo.ctypes.data, hi.ctypes.data)
    return lo, hi


def normalize_layout(keys, values):
    """Slots holding no item read as KEY_NONE. The reference's bulk path
    leaves key 0 / value 0 in its gaps (gpma.cuh:942-943 value-initialises
    the new arrays); self-loops are excluded from every workload, so a key-0,
    value-0 slot is always such a gap."""
    k = keys.copy()
    k[(keys == 0) & (values == 0)] = np.uint64(0xFFFFFFFFFFFFFFFF)
    v = values.copy()
    v[k == np.uint64(0xFFFFFFFFFFFFFFFF)] = 0
    return k, v


def write_edge_file(path, N, src, dst):
    with open(path, "wb") as f:
        f.write(np.uint32(N).tobytes())
        f.write(np.uint64(len(src)).tobytes())
        f.write(np.ascontiguousarray(src, np.uint32).tobytes())
        f.write(np.ascontiguousarray(dst, np.uint32).tobytes())


def run_ref_cpu_baseline(N, src, dst, slides, start, timing_only=False, workdir=None, timeout=None):
    """Run the reference's cpu_baseline (compiled into oracle/_ref) on the
    protocol; returns (timings dict, levels_before, levels_after, live_after)."""
    if not os.path.exists(REF_BIN):
        raise FileNotFoundError(f"{REF_BIN} not built (needs the reference tree at build time)")
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        ef = os.path.join(d, "edges.bin")
        write_edge_file(ef, N, src, dst)
        p = subprocess.run([REF_BIN, ef, d, str(slides), str(start), "1" if timing_only else "0"],
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"ref_cpu_baseline failed: {p.stderr}")
        t = json.loads(p.stdout.strip().splitlines()[-1])
        if timing_only:
            return t, None, None, None
        lb = np.fromfile(os.path.join(d, "levels_before.u32"), np.uint32)
        la = np.fromfile(os.path.join(d, "levels_after.u32"), np.uint32)
        live = np.fromfile(os.path.join(d, "live_after.u64"), np.uint64)
        return t, lb, la, live
