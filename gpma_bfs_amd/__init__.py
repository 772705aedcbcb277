"""gpma_bfs_amd: an MI355X (gfx950) GPMA for dynamic graphs.

Host-side mirror of the reference's C++ interface (gpma.cuh, gpma_bfs.cuh):

    g = GPMA(row_num)                      # GPMA<GPU>(row_num), gpma.cuh:79-82
    update_gpma(g, keys, values)           # gpma.cuh:968-969
    levels, reachable = gpma_bfs(g, start)  # gpma_bfs.cuh:517-533 (+ demo's count)

keys are uint64 ``src << 32 | dst``; values are uint8, 0 = delete. Arrays
may be numpy arrays (host) or torch tensors on the GPU (passed by device
pointer). Everything runs through libgpma.so's C ABI; there is no CPU path.
"""
import ctypes
import os

import numpy as np

from ._lib import (  # noqa: F401
    COL_IDX_NONE,
    GPMA_ERR_ARG,
    GPMA_ERR_CAPACITY,
    GPMA_ERR_HIP,
    GPMA_ERR_KEY_RANGE,
    GPMA_ERR_NO_DEVICE,
    GPMA_ERR_OOM,
    GPMA_ERR_STOPPED,
    GPMA_ERR_TIMEOUT,
    GPMA_OK,
    KEY_MAX,
    KEY_NONE,
    GpmaConfig,
    GpmaCsr,
    GpmaError,
    GpmaRequestResult,
    GpmaServerConfig,
    GpmaServerStats,
    GpmaStats,
    GpmaStepStats,
    check,
    last_error,
    lib,
)


def _ptr(a, dtype):
    """(pointer, length, on_device) for a numpy array or a torch tensor."""
    if isinstance(a, np.ndarray):
        if a.dtype != dtype:
            raise TypeError(f"expected {np.dtype(dtype)}, got {a.dtype}")
        if not a.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data, a.shape[0], 0
    data_ptr = getattr(a, "data_ptr", None)
    if data_ptr is not None:  # torch tensor
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return data_ptr(), a.shape[0], 1 if a.is_cuda else 0
    raise TypeError(f"unsupported array type {type(a)}")


# gpma_config knobs a tuning dict may set (0 = default); the density
# thresholds are in thousandths (reference 80 / 420 / 920 / 840, gpma.cuh:40-43)
DENSITY_KNOBS = ("lower_leaf_permille", "lower_root_permille", "upper_leaf_permille", "upper_root_permille")
TUNING_KNOBS = ("pull_permille", "segment_length", "device_levels", "sort_bucket_bits") + DENSITY_KNOBS + (
    "sort_items_per_lane", "push_tile", "ms_tile")


class GPMA:
    """One device-resident GPMA (reference: class GPMA, gpma.cuh:31-102)."""

    def __init__(self, row_num, device=-1, bulk_threshold=0, pull_permille=0, segment_length=0, device_levels=0,
                 sort_bucket_bits=0, tuning=None):
        """tuning: a dict of the knobs below (as written by gpma_bfs_amd.tune); explicit
        arguments win over it. None of them changes the live keys or BFS results."""
        t = dict(tuning or {})
        kw = {"pull_permille": pull_permille, "segment_length": segment_length, "device_levels": device_levels,
              "sort_bucket_bits": sort_bucket_bits}
        for k, v in kw.items():
            if v:
                t[k] = v
        unknown = set(t) - set(TUNING_KNOBS)
        if unknown:
            raise GpmaError(GPMA_ERR_ARG, f"unknown tuning knobs {sorted(unknown)}")
        self.tuning = {k: int(t.get(k, 0)) for k in TUNING_KNOBS}
        cfg = GpmaConfig(ctypes.sizeof(GpmaConfig), device, bulk_threshold, self.tuning["pull_permille"],
                         self.tuning["segment_length"], self.tuning["device_levels"],
                         self.tuning["sort_bucket_bits"])
        for i, k in enumerate(DENSITY_KNOBS):
            cfg.density_permille[i] = self.tuning[k]
        cfg.sort_items_per_lane = self.tuning["sort_items_per_lane"]
        cfg.push_tile = self.tuning["push_tile"]
        cfg.ms_tile = self.tuning["ms_tile"]
        h = ctypes.c_void_p()
        check(lib.gpma_create(row_num, ctypes.byref(cfg), ctypes.byref(h)), "gpma_create")
        self._h = h
        self.row_num = row_num

    @property
    def handle(self):
        if self._h is None:
            raise GpmaError(GPMA_ERR_ARG, "GPMA already closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib.gpma_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self):
        s = GpmaStats()
        check(lib.gpma_get_stats(self.handle, ctypes.byref(s)), "gpma_get_stats")
        return {name: getattr(s, name) for name, _ in GpmaStats._fields_}

    @property
    def segment_length(self):
        return self.stats()["segment_length"]

    @property
    def tree_height(self):
        return self.stats()["tree_height"]

    def get_size(self):
        """Capacity in slots (reference get_size(), gpma.cuh:52)."""
        return self.stats()["capacity"]

    def csr(self):
        """Device pointers and shape (reference public members, gpma.cuh:34-50)."""
        c = GpmaCsr()
        check(lib.gpma_get_csr(self.handle, ctypes.byref(c)), "gpma_get_csr")
        return c

    def copy_layout(self):
        """Host copies of keys, values, row_offset (reference mirror(), gpma.cuh:90-101)."""
        C = self.get_size()
        k = np.empty(C, np.uint64)
        v = np.empty(C, np.uint8)
        r = np.empty(self.row_num + 1, np.uint32)
        check(lib.gpma_copy_layout(self.handle, k.ctypes.data, v.ctypes.data, r.ctypes.data), "gpma_copy_layout")
        return k, v, r

    def export_live(self):
        """Sorted live keys (guards and the two KEY_MAX sentinels included)."""
        n = ctypes.c_uint64()
        check(lib.gpma_export_live(self.handle, None, 0, ctypes.byref(n)), "gpma_export_live")
        out = np.empty(n.value, np.uint64)
        check(lib.gpma_export_live(self.handle, out.ctypes.data, n.value, ctypes.byref(n)), "gpma_export_live")
        return out[: n.value]

    def live_digest(self):
        """(count, sum mod 2^64, xor) of the live keys, computed on the device."""
        c, sm, x = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib.gpma_live_digest(self.handle, ctypes.byref(c), ctypes.byref(sm), ctypes.byref(x)),
              "gpma_live_digest")
        return c.value, sm.value, x.value

    def test_fail_allocs(self, skip, count):
        """Make device allocations fail (testing the out-of-memory paths)."""
        check(lib.gpma_test_fail_allocs(self.handle, skip, count), "gpma_test_fail_allocs")

    def sync(self):
        check(lib.gpma_sync(self.handle), "gpma_sync")

    def profile(self, enable=True, reset=False):
        """Per-phase HIP-event timing on the handle's stream."""
        if reset:
            check(lib.gpma_profile(self.handle, 2), "gpma_profile")
        check(lib.gpma_profile(self.handle, int(bool(enable))), "gpma_profile")

    def profile_phases(self, names=None):
        """Time only the named phases (None: all of them)."""
        mask = (1 << len(PHASES)) - 1 if names is None else sum(1 << PHASES.index(n) for n in names)
        check(lib.gpma_profile_phases(self.handle, mask), "gpma_profile_phases")

    def profile_report(self):
        ms = (ctypes.c_double * len(PHASES))()
        calls = (ctypes.c_uint64 * len(PHASES))()
        check(lib.gpma_get_profile(self.handle, ms, calls, len(PHASES)), "gpma_get_profile")
        return {name: {"ms": ms[i], "launches": calls[i]} for i, name in enumerate(PHASES)}


PHASES = ("stage", "sort", "locate", "compact", "leaf", "upper", "rebuild", "bfs_expand")


class DeviceBuffer:
    """Device allocation through the library's own HIP runtime (plumbing for
    benchmarks and tests: keeps batches resident in HBM)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib.gpma_device_malloc(ctypes.byref(p), nbytes), "gpma_device_malloc")
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib.gpma_memcpy(b.ptr, a.ctypes.data, a.nbytes, 0), "gpma_memcpy")
        return b

    def to_numpy(self, dtype, count, offset_bytes=0):
        out = np.empty(count, dtype)
        check(lib.gpma_memcpy(out.ctypes.data, self.ptr + offset_bytes, out.nbytes, 1), "gpma_memcpy")
        return out

    def free(self):
        if self.ptr:
            lib.gpma_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def update_gpma_device(gpma, keys_ptr, values_ptr, n, strict=True):
    """update_gpma on batches already resident in device memory (raw pointers)."""
    rc = lib.gpma_update(gpma.handle, keys_ptr, values_ptr, n, 1)
    if strict:
        check(rc, "gpma_update")
    return rc


def init_csr_gpma(row_num, **kw):
    """Older initialiser name the reference's drivers log (gpma_bfs_demo.cu:63)."""
    return GPMA(row_num, **kw)


def update_gpma(gpma, keys, values, strict=True):
    """Apply one update batch (reference update_gpma, gpma.cuh:968-1050).

    Unlike the reference, the inputs are not sorted or compacted in place.
    Returns the status code; with strict=True a non-zero status raises
    (GPMA_ERR_KEY_RANGE still applies the valid part of the batch)."""
    if isinstance(keys, np.ndarray):
        keys = np.ascontiguousarray(keys)
    if isinstance(values, np.ndarray):
        values = np.ascontiguousarray(values)
    kp, n, kdev = _ptr(keys, np.uint64)
    vp, nv, vdev = _ptr(values, np.uint8)
    if n != nv:
        raise GpmaError(GPMA_ERR_ARG, f"keys ({n}) and values ({nv}) differ in length")
    if kdev != vdev:
        raise GpmaError(GPMA_ERR_ARG, "keys and values must both be on the host or both on the device")
    rc = lib.gpma_update(gpma.handle, kp, vp, n, kdev)
    if strict:
        check(rc, "gpma_update")
    return rc


def update_gpma_batches(gpma, batches_keys, batches_values, strict=True):
    """Apply several batches in one fused call (same live keys as applying
    them one after another). Host numpy arrays, one pair per batch."""
    keys = np.ascontiguousarray(np.concatenate(batches_keys), np.uint64)
    values = np.ascontiguousarray(np.concatenate(batches_values), np.uint8)
    offsets = np.zeros(len(batches_keys) + 1, np.uint64)
    offsets[1:] = np.cumsum([len(k) for k in batches_keys])
    if int(offsets[-1]) != len(values):
        raise GpmaError(GPMA_ERR_ARG, "keys and values differ in length")
    rc = lib.gpma_update_batches(gpma.handle, keys.ctypes.data, values.ctypes.data, offsets.ctypes.data,
                                 len(batches_keys), 0)
    if strict:
        check(rc, "gpma_update_batches")
    return rc


def update_gpma_batches_device(gpma, keys_ptr, values_ptr, offsets, strict=True):
    """Fused batches already resident in device memory; offsets on the host."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    rc = lib.gpma_update_batches(gpma.handle, keys_ptr, values_ptr, offsets.ctypes.data, len(offsets) - 1, 1)
    if strict:
        check(rc, "gpma_update_batches")
    return rc


def gpma_bfs(gpma, start_node, results=None):
    """BFS levels from start_node (reference gpma_bfs.cuh:517-533).

    Returns (levels, reachable): levels[v] = 1 + hop distance, 0 if
    unreachable; reachable = number of non-zero levels (gpma_bfs_demo.cu:80).
    results may be a preallocated uint32 numpy array or a torch device tensor."""
    if results is None:
        results = np.empty(gpma.row_num, np.uint32)
    rp, n, dev = _ptr(results, np.uint32) if isinstance(results, np.ndarray) else _ptr(results, None)
    if n < gpma.row_num:
        raise GpmaError(GPMA_ERR_ARG, "results shorter than row_num")
    reach = ctypes.c_uint32()
    check(lib.gpma_bfs(gpma.handle, start_node, rp, dev, ctypes.byref(reach)), "gpma_bfs")
    return results, reach.value


def gpma_bfs_multi(gpma, starts, levels_device_ptr=None, return_levels=True):
    """BFS from up to 64 sources in one traversal. Returns (levels, reachable):
    levels[i] is the level array gpma_bfs(starts[i]) returns (None if
    return_levels is False or levels_device_ptr is given)."""
    starts = np.ascontiguousarray(starts, np.uint32)
    ns = len(starts)
    reach = np.empty(ns, np.uint32)
    out = None
    if levels_device_ptr is not None:
        check(lib.gpma_bfs_multi(gpma.handle, starts.ctypes.data, ns, levels_device_ptr, 1, reach.ctypes.data),
              "gpma_bfs_multi")
    elif return_levels:
        out = np.empty((ns, gpma.row_num), np.uint32)
        check(lib.gpma_bfs_multi(gpma.handle, starts.ctypes.data, ns, out.ctypes.data, 0, reach.ctypes.data),
              "gpma_bfs_multi")
    else:
        check(lib.gpma_bfs_multi(gpma.handle, starts.ctypes.data, ns, None, 0, reach.ctypes.data), "gpma_bfs_multi")
    return out, reach


def gpma_step(gpma, keys, values, offsets, starts=(), levels_device_ptr=None, return_levels=False, strict=True):
    """One batched step through a captured HIP graph: the fused update
    batches (records [offsets[b], offsets[b+1]) of keys/values, host numpy
    arrays or device pointers) then a BFS from each of `starts` (<= 64).
    Same results as update_gpma_batches followed by gpma_bfs_multi. Returns
    (status, levels, reachable); levels is None unless return_levels (host
    copy) is set; with levels_device_ptr the levels stay in that buffer."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    starts = np.ascontiguousarray(starts, np.uint32)
    ns = len(starts)
    if isinstance(keys, np.ndarray):
        keys = np.ascontiguousarray(keys, np.uint64)
        values = np.ascontiguousarray(values, np.uint8)
        if len(keys) != len(values) or len(keys) < int(offsets[-1]):
            raise GpmaError(GPMA_ERR_ARG, "keys and values must cover offsets[-1] records")
        kp, vp, dev = keys.ctypes.data, values.ctypes.data, 0
    else:
        kp, vp, dev = int(keys), int(values), 1
    reach = np.empty(max(ns, 1), np.uint32)
    out = None
    lp, ldev = None, 0
    if levels_device_ptr is not None:
        lp, ldev = levels_device_ptr, 1
    elif return_levels and ns:
        out = np.empty((ns, gpma.row_num), np.uint32)
        lp = out.ctypes.data
    rc = lib.gpma_step(gpma.handle, kp, vp, offsets.ctypes.data, len(offsets) - 1, dev,
                       starts.ctypes.data if ns else None, ns, lp, ldev, reach.ctypes.data if ns else None)
    if strict:
        check(rc, "gpma_step")
    return rc, out, reach[:ns]


def gpma_step_prepare(gpma, max_records, nsrc=0, levels_device_ptr=None, bfs_levels=0):
    """Capture gpma_step's graphs for every size class up to max_records
    now (see gpma_step_prepare in include/gpma.h)."""
    check(lib.gpma_step_prepare(gpma.handle, max_records, nsrc, levels_device_ptr, bfs_levels), "gpma_step_prepare")


def gpma_step_stats(gpma):
    """Counters of gpma_step: replays, captures, host-path steps and tails."""
    st = GpmaStepStats()
    check(lib.gpma_get_step_stats(gpma.handle, ctypes.byref(st)), "gpma_get_step_stats")
    return {k: getattr(st, k) for k, _ in st._fields_}


def gpma_bfs_count(gpma, start_node, levels_device_ptr=None):
    """BFS that keeps the levels on the device (optionally into a caller's
    device buffer of row_num uint32); returns only the reachable count."""
    reach = ctypes.c_uint32()
    check(lib.gpma_bfs(gpma.handle, start_node, levels_device_ptr, 1, ctypes.byref(reach)), "gpma_bfs")
    return reach.value


def gpma_csr_bfs(keys_ptr, values_ptr, row_offsets_ptr, node_size, edge_size, start_node):
    """Legacy overload over raw device pointers (reference gpma_bfs.cuh:511-515)."""
    out = np.empty(node_size, np.uint32)
    check(lib.gpma_csr_bfs(keys_ptr, values_ptr, row_offsets_ptr, node_size, edge_size, start_node, out.ctypes.data),
          "gpma_csr_bfs")
    return out


NOALLOC_MAX_NODES = 1 << 21


def gpma_csr_bfs_noalloc(keys_ptr, values_ptr, row_offsets_ptr, node_size, edge_size, start_node, levels_ptr,
                         stream=None):
    """gpma_csr_bfs that allocates nothing: levels go to the device buffer at
    levels_ptr (node_size uint32), scratch is the library's static per-device
    array (node_size <= NOALLOC_MAX_NODES). Returns the reachable count."""
    reach = ctypes.c_uint32()
    check(lib.gpma_csr_bfs_noalloc(keys_ptr, values_ptr, row_offsets_ptr, node_size, edge_size, start_node,
                                   levels_ptr, ctypes.byref(reach), stream), "gpma_csr_bfs_noalloc")
    return reach.value


def gen_rmat(scale, ndraws, a=0.57, b=0.19, c=0.19, seed=1):
    """Seeded R-MAT edge list: unique, no self-loops, shuffled (src, dst uint32)."""
    n = ctypes.c_uint64()
    src = np.empty(ndraws, np.uint32)  # the unique edges are at most the draws
    dst = np.empty(ndraws, np.uint32)
    check(lib.gpma_gen_rmat(scale, ndraws, a, b, c, seed, src.ctypes.data, dst.ctypes.data, ndraws,
                            ctypes.byref(n)), "gpma_gen_rmat")
    return src[:n.value].copy(), dst[:n.value].copy()


def gen_powerlaw(N, E, alpha=2.1, soft_cap=2800.0, p1=0.24, seed=0):
    """Seeded power-law out-degree graph with exactly E unique edges."""
    src = np.empty(E, np.uint32)
    dst = np.empty(E, np.uint32)
    check(lib.gpma_gen_powerlaw(N, E, alpha, soft_cap, p1, seed, src.ctypes.data, dst.ctypes.data),
          "gpma_gen_powerlaw")
    return src, dst


def key_digest(keys):
    """(count, sum mod 2^64, xor) of a u64 key array: what GPMA.live_digest()
    returns for a live set equal to it (order and duplicates of distinct keys
    do not matter; each key counts once per occurrence)."""
    k = np.asarray(keys, np.uint64)
    with np.errstate(over="ignore"):
        return int(len(k)), int(k.sum(dtype=np.uint64)), int(np.bitwise_xor.reduce(k)) if len(k) else 0


def combine_digests(*ds):
    """Digest of the concatenation of key sets with the given digests."""
    return (sum(d[0] for d in ds), sum(d[1] for d in ds) % (1 << 64),
            __import__("functools").reduce(lambda a, b: a ^ b, [d[2] for d in ds], 0))


def read_edge_text(path):
    """(N, src, dst) from an "N E" + "src dst" text file (the reference's
    dataset format)."""
    n, e = ctypes.c_uint32(), ctypes.c_uint64()
    p = os.fsencode(path)
    check(lib.gpma_read_edge_text(p, ctypes.byref(n), ctypes.byref(e), None, None, 0), "gpma_read_edge_text")
    src = np.empty(e.value, np.uint32)
    dst = np.empty(e.value, np.uint32)
    check(lib.gpma_read_edge_text(p, ctypes.byref(n), ctypes.byref(e), src.ctypes.data, dst.ctypes.data, e.value),
          "gpma_read_edge_text")
    return n.value, src, dst


def write_edge_text(path, N, src, dst):
    """Write an edge list in the reference's "N E" + "src dst" text format."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    if len(src) != len(dst):
        raise GpmaError(GPMA_ERR_ARG, "src and dst differ in length")
    check(lib.gpma_write_edge_text(os.fsencode(path), N, len(src), src.ctypes.data, dst.ctypes.data),
          "gpma_write_edge_text")


def make_keys(src, dst):
    return (src.astype(np.uint64) << np.uint64(32)) | dst.astype(np.uint64)


class GpmaServer:
    """Request server over one GPMA (include/gpma_serve.h): updates and BFS
    requests from any thread take effect in ticket order; runs of queued
    updates are fused into one launch, runs of BFS requests share one
    multi-source traversal. The GPMA must not be used directly meanwhile."""

    def __init__(self, gpma, max_bfs_group=0, hbm_budget_bytes=0, max_fused_records=0, max_queued_records=0,
                 order="strict"):
        """order: "strict" (every request sees exactly the requests ticketed
        before it) or "snapshot" (a BFS may also see updates ticketed after
        it, which lets queued BFS requests share one traversal)."""
        cfg = GpmaServerConfig(ctypes.sizeof(GpmaServerConfig), max_bfs_group, hbm_budget_bytes, max_fused_records,
                               max_queued_records, {"strict": 0, "snapshot": 1}[order])
        h = ctypes.c_void_p()
        check(lib.gpma_server_create(gpma.handle, ctypes.byref(cfg), ctypes.byref(h)), "gpma_server_create")
        self._h = h
        self.gpma = gpma
        self._keep = {}  # ticket -> arrays the request reads or writes

    def submit_update(self, keys, values, n=None):
        """keys/values: host numpy arrays (copied at submission), or raw device
        pointers (ints) with n records - those must stay valid until the
        request completes."""
        if isinstance(keys, int):
            if not isinstance(values, int) or n is None:
                raise GpmaError(GPMA_ERR_ARG, "device pointers need a values pointer and n")
            kp, vp, dev = keys, values, 1
        else:
            keys = np.ascontiguousarray(keys, np.uint64)
            values = np.ascontiguousarray(values, np.uint8)
            if len(keys) != len(values):
                raise GpmaError(GPMA_ERR_ARG, "keys and values differ in length")
            kp, vp, n, dev = keys.ctypes.data, values.ctypes.data, len(keys), 0
        t = ctypes.c_uint64()
        check(lib.gpma_server_submit_update(self._h, kp, vp, n, dev, ctypes.byref(t)), "gpma_server_submit_update")
        return t.value

    def submit_bfs(self, start, levels=True):
        """levels=True: a uint32[row_num] array is filled and returned by wait()."""
        out = np.empty(self.gpma.row_num, np.uint32) if levels else None
        t = ctypes.c_uint64()
        check(lib.gpma_server_submit_bfs(self._h, start, out.ctypes.data if out is not None else None,
                                         ctypes.byref(t)), "gpma_server_submit_bfs")
        if out is not None:
            self._keep[t.value] = out
        return t.value

    def wait(self, ticket, timeout_s=None):
        """Result dict of a request (status, kind, reachable, group, queue_us,
        service_us, latency_us, state_ticket, and levels for a BFS submitted
        with levels)."""
        r = GpmaRequestResult()
        us = -1 if timeout_s is None else int(timeout_s * 1e6)
        check(lib.gpma_server_wait(self._h, ticket, us, ctypes.byref(r)), "gpma_server_wait")
        out = {name: getattr(r, name) for name, _ in GpmaRequestResult._fields_}
        out["kind"] = "update" if r.kind == 0 else "bfs"
        lv = self._keep.pop(ticket, None)
        if lv is not None:
            out["levels"] = lv
        return out

    def stats(self):
        s = GpmaServerStats()
        check(lib.gpma_server_get_stats(self._h, ctypes.byref(s)), "gpma_server_get_stats")
        return {name: getattr(s, name) for name, _ in GpmaServerStats._fields_ if name != "reserved"}

    def close(self):
        """Serve everything submitted, then stop."""
        if self._h:
            check(lib.gpma_server_destroy(self._h), "gpma_server_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
