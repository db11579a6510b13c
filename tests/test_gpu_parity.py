"""Parity of the HIP counting paths with the oracle, through the C ABI.

Bit-exact comparison of every slot count and the triangle total on seeded
graphs, including the full-size R-MAT scale-22 and Chung-Lu 2^23 graphs
(every slot against the oracle's bit-vector counts), plus size-independent
properties (reverse-slot symmetry, sum of counts = 6T).
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import bruteforce as BF
import oracle
from acctricnt_amd import graph as G

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
PATHS = ["forward", "bitmap", "merge"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx(built):
    from acctricnt_amd import cn

    assert cn.device_count() > 0, "no HIP device visible"
    c = cn.Context(0)
    yield c
    c.close()


def gpu_counts(ctx, g, path="forward", raw=False):
    with ctx.upload(g) as rg:
        return rg.count_all_edges(path, raw=raw)


def check_against_oracle(ctx, g, path):
    print(f"[parity] {g.name} n={g.n} m={g.m} path={path}", flush=True)
    cnt, t = gpu_counts(ctx, g, path)
    ref = oracle.cn_merge(g)
    bad = np.nonzero(cnt != ref)[0]
    assert bad.size == 0, f"{bad.size} slots differ, first {bad[:5]} gpu {cnt[bad[:5]]} ref {ref[bad[:5]]}"
    assert t == oracle.triangle_count(g)


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden(ctx, name, path):
    from test_oracle import golden_graph

    ent = GOLDEN[name]
    g = golden_graph(name)
    print(f"[golden] {name} path={path}", flush=True)
    cnt, t = gpu_counts(ctx, g, path)
    assert sha(cnt) == ent["cnt_sha256"]
    assert t == ent["triangles"]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("maker", [
    lambda: G.gen_er(16384, 16, seed=20190101),
    lambda: G.gen_rmat(16, 16, seed=1),
    lambda: G.gen_rmat(18, 16, seed=2),
    lambda: G.gen_rmat(14, 16, seed=3, relabel=False),
    lambda: G.gen_chung_lu(1 << 16, 2.1, 32, 1 << 13, seed=4),
    lambda: G.gen_chung_lu(1 << 15, 2.1, 24, 1 << 14, seed=5, relabel=True),
], ids=["er16384", "rmat16", "rmat18", "rmat14_unrelabelled", "chunglu16", "chunglu15_relabel"])
def test_seeded_graphs(ctx, maker, path):
    check_against_oracle(ctx, maker(), path)


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("seed", range(8))
def test_random_small(ctx, seed, path):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 300))
    k = int(rng.integers(0, n * 8 + 1))
    g = G.from_edges(n, rng.integers(0, n, k), rng.integers(0, n, k), name=f"rand{seed}")
    check_against_oracle(ctx, g, path)


@pytest.mark.parametrize("path", PATHS)
def test_closed_forms(ctx, path):
    from math import comb

    for k in (2, 3, 5, 16, 70):
        cnt, t = gpu_counts(ctx, BF.complete(k), path)
        assert (cnt == k).all() and t == comb(k, 3)
    cnt, t = gpu_counts(ctx, BF.wheel(9), path)
    assert t == 9
    cnt, t = gpu_counts(ctx, BF.windmill(6), path)
    assert (cnt == 3).all() and t == 6
    for g in (BF.complete_bipartite(4, 9), BF.cycle(10), BF.star(300)):
        cnt, t = gpu_counts(ctx, g, path)
        assert (cnt == 2).all() and t == 0


@pytest.mark.parametrize("path", PATHS)
def test_edge_cases(ctx, path):
    # one vertex, no edges
    g = G.from_csr(np.array([0, 0], np.uint32), np.zeros(0, np.int32))
    cnt, t = gpu_counts(ctx, g, path)
    assert cnt.size == 0 and t == 0
    # isolated vertices around a triangle
    g = G.from_edges(10, [2, 3, 2], [3, 7, 7])
    cnt, t = gpu_counts(ctx, g, path)
    assert (cnt == 3).all() and t == 1
    # forward lists longer than the LDS counter array: K_2600 (LDS bitmap, counters
    # spilling to atomics), also placed at ids above 2^17 (bitmap over [min N(b), b))
    for k, base in ((2600, 0), (2200, 140_000)):
        s, d = np.triu_indices(k, 1)
        g = G.from_edges(base + k, s + base, d + base)
        cnt, t = gpu_counts(ctx, g, path)
        assert (cnt == k).all()
        assert t == k * (k - 1) * (k - 2) // 6


def wide_span_graph(base, n_far, n_near, n_up, seed):
    """A source b = base whose lower neighbors span more than 2^17 ids (n_far
    ids near 0 and n_near ids just below b) with n_up upper neighbors, each
    joined to three random lower neighbors and to the next upper neighbor."""
    rng = np.random.default_rng(seed)
    low = np.concatenate([np.arange(n_far), np.arange(base - n_near, base)])
    up = np.arange(base + 1, base + 1 + n_up)
    src = [np.full(low.size + up.size, base), np.repeat(up, 3), up[:-1], low[:50]]
    dst = [np.concatenate([low, up]), rng.choice(low, 3 * up.size), up[1:], low[-50:]]
    return G.from_edges(base + n_up + 1, np.concatenate(src), np.concatenate(dst),
                        name=f"wide{n_far}+{n_near}x{n_up}")


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("shape", [(100, 900, 700), (1000, 2000, 300), (10, 40, 40)])
def test_wide_span_sources(ctx, path, shape):
    # |N<(b)| = 1000 with a span > 2^17: LDS hash table; |N<(b)| = 3000: global-memory
    # lookups; a small wide source: wave unit
    check_against_oracle(ctx, wide_span_graph(150_000, *shape, seed=sum(shape)), path)


@pytest.mark.parametrize("unroll", [1, 2, 4])
def test_unroll_variants(ctx, unroll):
    for g in (G.gen_rmat(15, 16, seed=21), G.gen_chung_lu(1 << 14, 2.1, 24, 4096, seed=22)):
        with ctx.upload(g) as rg:
            rg.set_param("fwd_unroll", unroll)
            assert rg.get_param("fwd_unroll") == unroll
            cnt, t = rg.count_all_edges("forward")
            # the triangle-only pass has its own loop per unroll (pipelined loads)
            t_only = rg.triangle_total("forward")
        t_ref = oracle.triangle_count(g)
        assert np.array_equal(cnt, oracle.cn_merge(g))
        assert t == t_ref and t_only == t_ref


def test_pinned_and_pageable_host_buffers(ctx):
    """Uploads and downloads go through pinned staging lanes from pageable
    memory and straight from pinned memory; both give the same CSR and counts.
    The graph is large enough (m > 2M slots) for the staged path."""
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    hip.hipHostFree.argtypes = [C.c_void_p]
    held = []

    def pinned(count, dtype):
        nbytes = count * np.dtype(dtype).itemsize
        p = C.c_void_p()
        assert hip.hipHostMalloc(C.byref(p), nbytes, 0) == 0
        held.append(p)
        return np.frombuffer((C.c_byte * nbytes).from_address(p.value), dtype=dtype, count=count)

    g = G.gen_rmat(17, 16, seed=13)
    assert g.m * 4 > (8 << 20)
    ref = oracle.cn_merge(g)
    try:
        off_p, adj_p = pinned(g.n + 1, np.uint32), pinned(g.m, np.int32)
        off_p[:] = g.node_off
        adj_p[:] = g.edge_dst
        for src in (g, G.CSRGraph(off_p, adj_p)):
            with ctx.upload(src) as rg:
                r = rg.csr()
                assert np.array_equal(r.node_off, g.node_off) and np.array_equal(r.edge_dst, g.edge_dst)
                cnt, t = rg.count_all_edges("forward")
                assert np.array_equal(cnt, ref) and t == oracle.triangle_count(g)
                out = pinned(g.m, np.int32)
                assert rg._lib.acc_cn_download(rg.handle, out.ctypes.data_as(C.POINTER(C.c_int32)), None) == 0
                assert np.array_equal(out, ref)
    finally:
        for p in held:
            hip.hipHostFree(p)


def test_raw_flag(ctx):
    g = G.gen_rmat(12, 16, seed=9)
    c2, t2 = gpu_counts(ctx, g, "forward")
    c0, t0 = gpu_counts(ctx, g, "forward", raw=True)
    assert np.array_equal(c2, c0 + 2) and t2 == t0


def test_repeat_is_deterministic(ctx):
    g = G.gen_rmat(15, 16, seed=10)
    with ctx.upload(g) as rg:
        a, ta = rg.count_all_edges("forward")
        b, tb = rg.count_all_edges("forward")
    assert np.array_equal(a, b) and ta == tb


MALFORMED = {
    "selfloop": ([0, 2, 4, 6], [0, 1, 0, 2, 1, 2]),
    "unsorted": ([0, 2, 4, 6], [2, 1, 0, 2, 0, 1]),
    "duplicate": ([0, 2, 4, 6], [1, 2, 0, 2, 0, 0]),
    "range": ([0, 2, 4, 6], [1, 5, 0, 2, 0, 1]),
    "asym_upper": ([0, 2, 3, 3], [1, 2, 0]),  # (0,2) has no reverse
    "asym_lower": ([0, 1, 2, 3], [1, 0, 0]),  # (2,0) has no reverse; every upper slot has one
    # as many upper as lower slots, but (2,3) and (3,1) have no reverse
    "asym_balanced": ([0, 1, 2, 3, 4], [1, 0, 3, 1]),
    # the same beside matched edges: (0,1), (0,2) are symmetric, (2,3) / (3,1) are not
    "asym_balanced_long": ([0, 2, 3, 5, 6], [1, 2, 0, 0, 3, 1]),
}


@pytest.mark.parametrize("case", sorted(MALFORMED))
@pytest.mark.parametrize("relabel", [False, True])
def test_malformed_upload_rejected(ctx, case, relabel):
    from acctricnt_amd import _ffi

    off, adj = MALFORMED[case]
    g = G.CSRGraph(np.array(off, np.uint32), np.array(adj, np.int32))
    with pytest.raises(_ffi.AccError) as e:
        ctx.upload(g, relabel=relabel)
    assert e.value.code == -2
    msg = str(e.value)
    want = {"selfloop": "Self loop", "unsorted": "not sorted", "duplicate": "not sorted",
            "range": "out of range", "asym_upper": "not symmetric", "asym_lower": "not symmetric",
            "asym_balanced": "not symmetric", "asym_balanced_long": "not symmetric"}[case]
    assert want in msg, msg


@pytest.mark.parametrize("case", ["star_out", "asym_balanced"])
def test_trusted_asymmetric_upload(ctx, case):
    """An asymmetric CSR uploaded without validation (the caller vouches for it)
    still uploads and counts within its arrays: with more upper than lower slots
    the reverse slots come from the per-slot binary search, otherwise from the
    sort.  Neither graph has a triangle: an edge listed both ways gets 2 in
    both slots, a slot whose reverse is absent keeps its raw count, 0."""
    if case == "star_out":  # vertex 0 lists 1..10, no list lists 0
        off = np.array([0] + [10] * 11, np.uint32)
        adj = np.arange(1, 11, dtype=np.int32)
        want = np.zeros(10, np.int32)
    else:  # (0,1) listed both ways, (2,3) and (3,1) one way only
        off, adj = (np.array(x, t) for x, t in zip(MALFORMED[case], (np.uint32, np.int32)))
        want = np.array([2, 2, 0, 0], np.int32)
    g = G.CSRGraph(off, adj)
    with ctx.upload(g, validate=False) as rg:
        cnt, t = rg.count_all_edges("forward")
    assert np.array_equal(cnt, want) and t == 0


def test_unknown_upload_flags_rejected(ctx):
    """Upload flags other than NO_VALIDATE and RELABEL are refused (ACC_EINVAL)
    before anything is allocated."""
    from acctricnt_amd import _ffi

    g = BF.complete(5)
    off = np.ascontiguousarray(g.node_off, np.uint32)
    adj = np.ascontiguousarray(g.edge_dst, np.int32)
    h = C.c_void_p()
    rc = ctx._lib.acc_graph_upload(ctx.handle, g.n, g.m, _ffi.ptr(off, C.c_uint32), _ffi.ptr(adj, C.c_int32),
                                   0x4, C.byref(h))
    assert rc == -1 and not h.value
    assert "unknown upload flags" in ctx._lib.acc_last_error().decode()


def test_paths_agree_rmat20(ctx):
    """R-MAT scale 20: the forward, bitmap and merge paths give identical counts,
    with T equal to the oracle's triangle count."""
    g = G.gen_rmat(20, 16, seed=3)
    print(f"[paths rmat20] n={g.n} m={g.m}", flush=True)
    res = {}
    with ctx.upload(g) as rg:
        for p in PATHS:
            res[p] = rg.count_all_edges(p)
    t = oracle.triangle_count(g)
    for p in PATHS:
        assert np.array_equal(res[p][0], res["forward"][0]), p
        assert res[p][1] == t, p


@pytest.fixture(scope="module")
def rmat22():
    """R-MAT scale 22 / ef 16 (the benchmark's graph) with the oracle's
    bit-vector counts of every slot and its merge-based triangle total."""
    g = G.gen_rmat(22, 16, seed=1)
    ref = oracle.cn_bitmap(g)
    t = oracle.triangle_count(g)
    print(f"[rmat22] n={g.n} m={g.m} T={t}", flush=True)
    return g, ref, t


@pytest.mark.parametrize("path,params", [("forward", {}), ("forward", {"chunk": 128}),
                                         ("forward", {"chunk": 256, "fwd_unroll": 2}), ("bitmap", {}), ("merge", {})],
                         ids=["forward", "forward_chunk128", "forward_chunk256_u2", "bitmap", "merge"])
def test_full_size_rmat22(ctx, rmat22, path, params):
    """R-MAT scale 22 / ef 16, every one of the 128.3M slots bit-exact against
    the oracle's counts on each path, T equal to the oracle's triangle count,
    plus symmetry and sum = 6T."""
    g, ref, t_ref = rmat22
    print(f"[rmat22] path={path} params={params}", flush=True)
    with ctx.upload(g) as rg:
        for k, v in params.items():
            rg.set_param(k, v)
        cnt, t = rg.count_all_edges(path)
    bad = np.nonzero(cnt != ref)[0]
    assert bad.size == 0, f"{bad.size} slots differ, first {bad[:5]} gpu {cnt[bad[:5]]} ref {ref[bad[:5]]}"
    assert t == t_ref
    assert int((cnt.astype(np.int64) - 2).sum()) == 6 * t


@pytest.mark.parametrize("path", PATHS + ["auto"])
def test_triangle_total(ctx, path):
    """T from acc_triangle_total (the triangle-only pass on the forward path)
    against the oracle on seeded graphs, cliques and sources of every bin."""
    cases = [G.gen_rmat(14, 16, seed=71), G.gen_er(4000, 20, seed=72), G.gen_chung_lu(1 << 14, 2.1, 24, 4096, seed=73),
             wide_span_graph(150_000, 100, 900, 700, seed=74), wide_span_graph(150_000, 1000, 2000, 300, seed=75)]
    s, d = np.triu_indices(2600, 1)
    cases.append(G.from_edges(2600, s, d))
    cases.append(G.from_csr(np.array([0, 0], np.uint32), np.zeros(0, np.int32)))
    for g in cases:
        with ctx.upload(g) as rg:
            assert rg.triangle_total(path) == oracle.triangle_count(g), g.name


def test_triangle_total_leaves_counts(ctx):
    """The triangle-only pass writes no per-slot counts: counts and their total
    from an earlier full pass stay as they were."""
    g = G.gen_rmat(15, 16, seed=76)
    with ctx.upload(g) as rg:
        cnt, t = rg.count_all_edges("forward")
        for name, value in (("chunk", 128), ("wave_slots", 0), ("persist", 2), ("fwd_unroll", 1)):
            rg.set_param(name, value)
            assert rg.triangle_total("forward") == t, (name, value)
        cnt2, t2 = rg.download()
    assert np.array_equal(cnt, cnt2) and t2 == t


def test_triangle_total_rmat22(ctx, rmat22):
    g, _, t_ref = rmat22
    with ctx.upload(g) as rg:
        assert rg.triangle_total("forward") == t_ref


def test_full_size_chunglu23(ctx):
    """Chung-Lu, exponent 2.1, n = 2^23, average degree 32, expected degrees
    capped at 2^20: every slot bit-exact against the oracle's counts, T equal
    to the oracle's, sum = 6T, and SCAN at eps 0.3 and 0.6, mu 5, equal to
    the oracle's SCAN over the oracle's counts."""
    g = G.gen_chung_lu(1 << 23, 2.1, 32, 1 << 20, seed=1)
    print(f"[chunglu23] n={g.n} m={g.m} max_deg={int(g.degrees().max())}", flush=True)
    with ctx.upload(g) as rg:
        cnt, t = rg.count_all_edges("forward")
        assert rg.triangle_total("forward") == t
        scans = {}
        with rg.scan_workspace() as ws:
            for eps in ("0.3", "0.6"):
                ws.run(eps, 5)
                ws.synchronize()
                scans[eps] = ws.results()
    ref = oracle.cn_bitmap(g)
    bad = np.nonzero(cnt != ref)[0]
    assert bad.size == 0, f"{bad.size} slots differ, first {bad[:5]} gpu {cnt[bad[:5]]} ref {ref[bad[:5]]}"
    assert int((cnt.astype(np.int64) - 2).sum()) == 6 * t
    assert t == oracle.triangle_count(g)
    for eps, res in scans.items():
        want = oracle.scan(g, ref, eps, 5)
        assert np.array_equal(res["label"], want["label"]), eps
        core = want["label"] == 1
        assert np.array_equal(res["cluster_id"][core], want["cluster_id"][core]), eps
        for k in ("core_num", "cluster_num", "hub_num", "outlier_num"):
            assert res[k] == want[k], (eps, k)


@pytest.mark.parametrize("hash_max", [0, 100, 256, 2048, 8192, 16384])
def test_bitmap_hash_and_hub_classes(ctx, hash_max):
    """The bitmap path with every split between LDS hash tables (2^12, 2^14,
    2^15 keys) and HBM hub bitmaps: hash_max 0 sends every source to the hub
    bitmaps (ER-16384 then needs four batches of 4096 hubs), 100 splits small
    sources across both; counts bit-exact against the oracle, and the merge
    path's windows of 2^11 and 2^14 ids alongside."""
    graphs = [G.gen_er(16384, 16, seed=20190101), G.gen_rmat(17, 16, seed=7),
              G.gen_chung_lu(1 << 16, 2.1, 32, 40000, seed=8), wide_span_graph(150_000, 1000, 2000, 300, seed=75)]
    for g in graphs:
        print(f"[bitmap classes] {g.name} n={g.n} m={g.m} max_deg={int(g.degrees().max())} hash_max={hash_max}",
              flush=True)
        ref = oracle.cn_bitmap(g)
        t_ref = int((ref.astype(np.int64) - 2).sum()) // 6
        with ctx.upload(g) as rg:
            rg.set_param("bmp_hash_max", hash_max)
            for path in ("bitmap", "merge"):
                cnt, t = rg.count_all_edges(path)
                bad = np.nonzero(cnt != ref)[0]
                assert bad.size == 0, f"{path}: {bad.size} slots differ, first {bad[:5]}"
                assert t == t_ref
