"""ctypes bindings of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.

Parity of every function below is pinned in tests/test_oracle.py (closed
forms, scipy.sparse, brute force, cross-checks); see oracle.c's header.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CORE = 1
HUB = 2

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    u8p, i32p, u32p, i64p = (C.POINTER(C.c_uint8), C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                             C.POINTER(C.c_int64))
    L.oracle_cn_merge.argtypes = [C.c_uint32, C.c_uint64, u32p, i32p, i32p, C.c_int]
    L.oracle_cn_merge.restype = None
    L.oracle_cn_merge_range.argtypes = [C.c_uint32, u32p, i32p, i32p, C.c_uint32, C.c_uint32, C.c_int]
    L.oracle_cn_merge_range.restype = None
    L.oracle_cn_merge_slots.argtypes = [C.c_uint32, u32p, i32p, i32p, u32p, i32p, C.c_uint64, C.c_int]
    L.oracle_cn_merge_slots.restype = None
    L.oracle_cn_bitmap.argtypes = [C.c_uint32, C.c_uint64, u32p, i32p, i32p, C.c_int]
    L.oracle_cn_bitmap.restype = None
    L.oracle_triangle_count.argtypes = [C.c_uint32, u32p, i32p, C.c_int]
    L.oracle_triangle_count.restype = C.c_uint64
    L.oracle_triangle_count_bitmap.argtypes = [C.c_uint32, u32p, i32p, C.c_int]
    L.oracle_triangle_count_bitmap.restype = C.c_uint64
    L.oracle_scan_core_fp.argtypes = [C.c_uint32, u32p, i32p, i32p, C.c_double, C.c_int, u8p, i32p, i32p]
    L.oracle_scan_core_fp.restype = None
    L.oracle_scan_core_int.argtypes = [C.c_uint32, u32p, i32p, i32p, C.c_int64, C.c_int64, C.c_int,
                                       u8p, i32p, i32p]
    L.oracle_scan_core_int.restype = None
    L.oracle_scan_cluster.argtypes = [C.c_uint32, u32p, i32p, u8p, i32p, i32p, i32p, i32p,
                                      i64p, i64p, i64p, i64p]
    L.oracle_scan_cluster.restype = C.c_int
    L.oracle_scan_write_result.argtypes = [C.c_char_p, C.c_uint32, u32p, i32p, u8p, i32p, i32p, i32p]
    L.oracle_scan_write_result.restype = C.c_int
    L.oracle_threads.argtypes = []
    L.oracle_threads.restype = C.c_int
    L.oracle_eps_string.argtypes = [C.c_double, C.c_char_p]
    L.oracle_eps_string.restype = None
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _csr(g):
    return (np.ascontiguousarray(g.node_off, dtype=np.uint32),
            np.ascontiguousarray(g.edge_dst, dtype=np.int32))


def cn_merge(g, nthreads: int = 0) -> np.ndarray:
    """Per-slot counts 2 + |N(u) ∩ N(v)| (CheckCoreSortedArray)."""
    off, adj = _csr(g)
    cnt = np.empty(g.m, dtype=np.int32)
    lib().oracle_cn_merge(g.n, g.m, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(cnt, C.c_int32), nthreads)
    return cnt


def cn_merge_range(g, cnt: np.ndarray, u_lo: int, u_hi: int, nthreads: int = 0) -> None:
    off, adj = _csr(g)
    lib().oracle_cn_merge_range(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(cnt, C.c_int32),
                                u_lo, u_hi, nthreads)


def cn_merge_slots(g, cnt: np.ndarray, slots, srcs, nthreads: int = 0) -> None:
    """Merge-count the given u<v slots (with their source vertices) into cnt."""
    off, adj = _csr(g)
    sl = np.ascontiguousarray(slots, dtype=np.uint32)
    sr = np.ascontiguousarray(srcs, dtype=np.int32)
    lib().oracle_cn_merge_slots(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(cnt, C.c_int32),
                                _p(sl, C.c_uint32), _p(sr, C.c_int32), sl.shape[0], nthreads)


def cn_bitmap(g, nthreads: int = 0) -> np.ndarray:
    """Per-slot counts by the CPU bitmap path (CheckCoreBitmap)."""
    off, adj = _csr(g)
    cnt = np.empty(g.m, dtype=np.int32)
    lib().oracle_cn_bitmap(g.n, g.m, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(cnt, C.c_int32), nthreads)
    return cnt


def triangle_count(g, nthreads: int = 0) -> int:
    off, adj = _csr(g)
    return int(lib().oracle_triangle_count(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), nthreads))


def triangle_count_bitmap(g, nthreads: int = 0) -> int:
    """T by the bit-vector triangle count (TriCntBitmap)."""
    off, adj = _csr(g)
    return int(lib().oracle_triangle_count_bitmap(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), nthreads))


def eps_fraction(eps) -> Fraction:
    """Exact rational value of a decimal eps string (or the shortest decimal of a float)."""
    return Fraction(str(eps)) if not isinstance(eps, str) else Fraction(eps)


def scan_core(g, cnt, eps, mu: int, exact: bool = True):
    """Similarity + core step.  exact=True uses the integer decision,
    exact=False the reference's long-double decision.  Returns (sim, core_count, label)."""
    off, adj = _csr(g)
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    sim = np.empty(g.m, dtype=np.uint8)
    cc = np.empty(g.n, dtype=np.int32)
    label = np.empty(g.n, dtype=np.int32)
    if exact:
        fr = eps_fraction(eps)
        lib().oracle_scan_core_int(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(cnt, C.c_int32),
                                   fr.numerator, fr.denominator, mu, _p(sim, C.c_uint8),
                                   _p(cc, C.c_int32), _p(label, C.c_int32))
    else:
        lib().oracle_scan_core_fp(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(cnt, C.c_int32),
                                  float(eps), mu, _p(sim, C.c_uint8), _p(cc, C.c_int32),
                                  _p(label, C.c_int32))
    return sim, cc, label


def scan_cluster(g, sim, label, old_id=None):
    """Union-find, hubs, cluster ids.  Returns dict with label, root, cluster_id and counters."""
    off, adj = _csr(g)
    label = np.array(label, dtype=np.int32, copy=True)
    sim = np.ascontiguousarray(sim, dtype=np.uint8)
    root = np.empty(g.n, dtype=np.int32)
    cid = np.empty(g.n, dtype=np.int32)
    cnts = [C.c_int64() for _ in range(4)]
    oid = None if old_id is None else np.ascontiguousarray(old_id, dtype=np.int32)
    lib().oracle_scan_cluster(g.n, _p(off, C.c_uint32), _p(adj, C.c_int32), _p(sim, C.c_uint8),
                              _p(label, C.c_int32), None if oid is None else _p(oid, C.c_int32),
                              _p(root, C.c_int32), _p(cid, C.c_int32), *[C.byref(c) for c in cnts])
    return {"label": label, "root": root, "cluster_id": cid, "core_num": cnts[0].value,
            "cluster_num": cnts[1].value, "hub_num": cnts[2].value, "outlier_num": cnts[3].value}


def scan(g, cnt, eps, mu: int, exact: bool = True, old_id=None):
    sim, cc, label = scan_core(g, cnt, eps, mu, exact)
    res = scan_cluster(g, sim, label, old_id)
    res["sim"] = sim
    res["core_count"] = cc
    return res


def write_result(path: str, g, res, old_id=None) -> None:
    off, adj = _csr(g)
    oid = None if old_id is None else np.ascontiguousarray(old_id, dtype=np.int32)
    rc = lib().oracle_scan_write_result(path.encode(), g.n, _p(off, C.c_uint32), _p(adj, C.c_int32),
                                        _p(res["sim"], C.c_uint8), _p(res["label"], C.c_int32),
                                        _p(res["cluster_id"], C.c_int32),
                                        None if oid is None else _p(oid, C.c_int32))
    if rc != 0:
        raise OSError(f"cannot write {path}")


def relabel_degree(g):
    """Degree relabel as (node_off, edge_dst, old_of_new): new ids by degree,
    descending, ties by ascending old id, every list mapped and sorted.

    Follows SCAN_XP::Reorder (scan-query/scan_xp.cpp:1145-1222: old_vid_dict,
    then each list mapped through the new ids and re-sorted) with the order of
    ReorderDegDescending (:1279-1301), whose unstable parallel sort leaves
    equal degrees in an unspecified order; here they keep ascending old id.
    Written in numpy (a stable argsort and one sort of (source, neighbor)
    keys); pinned in tests/test_oracle.py by the properties that determine
    the result uniquely."""
    n = g.n
    deg = np.diff(np.asarray(g.node_off, dtype=np.int64))
    old_of_new = np.argsort(-deg, kind="stable").astype(np.int32)
    new_of_old = np.empty(n, dtype=np.int64)
    new_of_old[old_of_new] = np.arange(n, dtype=np.int64)
    off = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(deg[old_of_new], out=off[1:])
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    keys = (new_of_old[src] << 32) | new_of_old[np.asarray(g.edge_dst, dtype=np.int64)]
    keys.sort()
    return off, (keys & 0xFFFFFFFF).astype(np.int32), old_of_new


def threads() -> int:
    return int(lib().oracle_threads())


def eps_string(eps: float) -> str:
    buf = C.create_string_buffer(64)
    lib().oracle_eps_string(float(eps), buf)
    return buf.value.decode()
