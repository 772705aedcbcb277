"""CPU tests of the C ABI library and host logic (no GPU compute)."""
import ctypes
import hashlib
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import gpma_bfs_amd as G
from gpma_bfs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("gpma.h", "gpma_serve.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|void\s*\*|const char\s*\*)\s*(gpma_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_every_c_header_is_checked():
    found = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert found == sorted(os.path.basename(h) for h in HEADERS)


def test_header_declares_expected_surface():
    names = declared_functions()
    for must in ("gpma_create", "gpma_destroy", "gpma_update", "gpma_bfs", "gpma_csr_bfs", "gpma_export_live",
                 "gpma_get_csr", "gpma_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_struct_sizes_match_header_layout():
    assert ctypes.sizeof(_lib.GpmaConfig) == 32 + 8 + 4 * 4
    assert ctypes.sizeof(_lib.GpmaCsr) == 3 * 8 + 8 + 4 * 4
    assert ctypes.sizeof(_lib.GpmaStats) == 8 + 4 * 4 + 6 * 8
    assert ctypes.sizeof(_lib.GpmaServerConfig) == 4 + 4 + 3 * 8 + 4 * 4
    assert ctypes.sizeof(_lib.GpmaRequestResult) == 4 * 4 + 4 * 8
    assert ctypes.sizeof(_lib.GpmaServerStats) == 8 * 8 + 2 * 4 + 4 * 8
    assert ctypes.sizeof(_lib.GpmaStepStats) == 6 * 8 + 2 * 4


def test_no_device_paths_fail_loudly():
    if _lib.lib.gpma_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = _lib.lib.gpma_create(16, None, ctypes.byref(h))
    assert rc == _lib.GPMA_ERR_NO_DEVICE
    assert "device" in _lib.last_error()
    with pytest.raises(G.GpmaError):
        G.GPMA(16)


def test_rmat_generator_deterministic_and_clean():
    a = G.gen_rmat(12, 50000, seed=3)
    b = G.gen_rmat(12, 50000, seed=3)
    c = G.gen_rmat(12, 50000, seed=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    k = G.make_keys(*a)
    assert len(np.unique(k)) == len(k)
    assert not np.any(a[0] == a[1])
    assert a[0].max() < 4096 and a[1].max() < 4096


def test_rmat_c1_matches_golden_edge_hash():
    fx = np.load(os.path.join(ROOT, "tests", "golden", "c1_rmat16_seed1.npz"))
    src, dst = G.gen_rmat(16, 1 << 20, seed=1)
    k = G.make_keys(src, dst)
    assert len(k) == int(fx["n_edges"])
    assert hashlib.sha256(k.tobytes()).hexdigest() == str(fx["edges_sha256"])


def test_powerlaw_generator_shape():
    N, E = 20000, 300000
    src, dst = G.gen_powerlaw(N, E, seed=5)
    assert len(src) == E
    k = G.make_keys(src, dst)
    assert len(np.unique(k)) == E
    assert not np.any(src == dst)
    assert src.max() < N and dst.max() < N
    deg = np.bincount(src, minlength=N)
    assert abs(deg.mean() - E / N) < 1e-9
    assert deg.max() > 20 * deg.mean()  # heavy tail
    # R-MAT-skewed destinations: low ids are far more popular than high ones
    indeg = np.bincount(dst, minlength=N)
    assert indeg[:N // 16].sum() > indeg[N // 2:].sum()
    src2, dst2 = G.gen_powerlaw(N, E, seed=5)
    assert np.array_equal(src, src2) and np.array_equal(dst, dst2)


def test_bench_rejects_multi_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr



def test_edge_text_format_round_trip_and_errors(tmp_path):
    """The reference's "N E" + "src dst" text format: a hand-written file
    (tabs, CRLF, blank lines) reads back as written; the writer's output
    equals numpy's own rendering and reads back; malformed files and short
    buffers are errors."""
    p = tmp_path / "g.txt"
    p.write_bytes(b"5 4\r\n0 1\r\n1\t2\n\n3 4\n  4 0\n")
    n, s, d = G.read_edge_text(p)
    assert n == 5 and s.tolist() == [0, 1, 3, 4] and d.tolist() == [1, 2, 4, 0]
    src, dst = G.gen_rmat(10, 3000, seed=11)
    q = tmp_path / "r.txt"
    G.write_edge_text(q, 1024, src, dst)
    lines = q.read_text().splitlines()
    assert lines[0] == f"1024 {len(src)}"
    assert lines[1:] == [f"{a} {b}" for a, b in zip(src.tolist(), dst.tolist())]
    n, s2, d2 = G.read_edge_text(q)
    assert n == 1024 and np.array_equal(s2, src) and np.array_equal(d2, dst)
    for bad in (b"", b"5", b"x 3\n", b"5 3\n0 1\n1 2\n", b"5 2\n0 1\n1 -2\n", b"5 1\n0 99999999999\n"):
        p.write_bytes(bad)
        with pytest.raises(G.GpmaError) as e:
            G.read_edge_text(p)
        assert e.value.code == G.GPMA_ERR_ARG
    with pytest.raises(G.GpmaError):
        G.read_edge_text(tmp_path / "missing.txt")
