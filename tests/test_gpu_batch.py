"""Packed multi-graph batches, captured replay, tuning parameters and the
autotuner, through the C ABI.

Every graph of a packed batch must get exactly the per-slot counts and the
triangle total the oracle gives for that graph alone; a captured counting
pass must reproduce the eager pass on every replay; every tuning parameter
setting, and whatever the autotuner chooses, must leave the counts bit-exact.
"""
import numpy as np
import pytest

import bruteforce as BF
import oracle
from acctricnt_amd import graph as G

pytestmark = pytest.mark.gpu

PATHS = ["forward", "bitmap", "merge"]


@pytest.fixture(scope="module")
def ctx(built):
    from acctricnt_amd import cn

    assert cn.device_count() > 0, "no HIP device visible"
    c = cn.Context(0)
    yield c
    c.close()


def empty_graph(n):
    return G.CSRGraph(np.zeros(n + 1, np.uint32), np.zeros(0, np.int32), name=f"empty{n}")


def clique(k):
    s, d = np.triu_indices(k, 1)
    return G.from_edges(k, s, d, name=f"K{k}")


@pytest.fixture(scope="module")
def small_batch():
    graphs = G.batch_mix(num_graphs=40, seed=7, n_lo_log2=5, n_hi_log2=11, deg_lo=2, deg_hi=40)
    graphs[3:3] = [empty_graph(0), empty_graph(5), clique(300), BF.star(200), G.gen_rmat(13, 16, seed=5)]
    refs = [(oracle.cn_merge(g), oracle.triangle_count(g)) for g in graphs]
    return graphs, refs


def check_batch(graphs, refs, packed, cnt, totals):
    so = packed.meta["slot_off"]
    for k, (g, (rc, rt)) in enumerate(zip(graphs, refs)):
        got = cnt[so[k]:so[k + 1]]
        bad = np.nonzero(got != rc)[0]
        assert bad.size == 0, f"graph {k} ({g.name} n={g.n}): {bad.size} slots differ"
        assert int(totals[k]) == rt, f"graph {k}: total {int(totals[k])} != {rt}"


@pytest.mark.parametrize("path", PATHS)
def test_packed_batch_parity(ctx, small_batch, path):
    graphs, refs = small_batch
    packed, vo = G.pack(graphs)
    print(f"[batch] {len(graphs)} graphs n={packed.n} m={packed.m} path={path}", flush=True)
    with ctx.upload(packed) as rg:
        rg.set_partition(vo)
        cnt, t = rg.count_all_edges(path)
        totals = rg.graph_totals()
    check_batch(graphs, refs, packed, cnt, totals)
    assert t == sum(r[1] for r in refs)


def test_capture_replay(ctx, small_batch):
    graphs, refs = small_batch
    packed, vo = G.pack(graphs)
    with ctx.upload(packed) as rg:
        rg.set_partition(vo)
        eager, t_eager = rg.count_all_edges("forward")
        rg.capture("forward")
        for r in range(3):
            print(f"[replay] {r}", flush=True)
            rg.replay()
            ctx.synchronize()
            cnt, t = rg.download()
            totals = rg.graph_totals()
            assert np.array_equal(cnt, eager) and t == t_eager, f"replay {r}"
            check_batch(graphs, refs, packed, cnt, totals)
            # eager work between replays overwrites the counts in another format
            raw, _ = rg.count_all_edges("merge", raw=True)
            assert np.array_equal(raw + 2, eager)


def test_capture_replay_bitmap_hubs(ctx):
    """Captured bitmap path with hub sources (persistent HBM-bitmap kernel)."""
    g = G.gen_chung_lu(1 << 14, 2.1, 24, 8192, seed=31)
    ref = oracle.cn_merge(g)
    with ctx.upload(g) as rg:
        rg.capture("bitmap")
        for r in range(3):
            rg.replay()
            ctx.synchronize()
            cnt, t = rg.download()
            assert np.array_equal(cnt, ref), f"replay {r}"
            rg.count_all_edges("forward")


PARAM_CASES = [("persist", 0), ("persist", 1), ("persist", 4), ("persist", 24), ("persist", 64), ("wave_slots", 0), ("wave_slots", 16),
               ("wave_low", 0), ("wave_low", 64), ("chunk", 64), ("chunk", 128), ("chunk", 300)]


@pytest.mark.parametrize("name,value", PARAM_CASES)
def test_params_keep_counts(ctx, name, value):
    for g in (G.gen_rmat(15, 16, seed=41), G.gen_chung_lu(1 << 14, 2.1, 24, 4096, seed=42)):
        with ctx.upload(g) as rg:
            rg.set_param(name, value)
            assert rg.get_param(name) == value
            cnt, t = rg.count_all_edges("forward")
            t_only = rg.triangle_total("forward")
        t_ref = oracle.triangle_count(g)
        assert np.array_equal(cnt, oracle.cn_merge(g)), (name, value, g.name)
        assert t == t_ref and t_only == t_ref, (name, value, g.name)


def test_params_rejected(ctx):
    from acctricnt_amd import _ffi

    g = G.gen_rmat(10, 8, seed=1)
    with ctx.upload(g) as rg:
        for name, value in (("fwd_unroll", 3), ("wave_slots", 65), ("wave_low", 257), ("chunk", 63),
                            ("chunk", 513), ("persist", 257), ("auto_path", 0)):
            with pytest.raises(_ffi.AccError):
                rg.set_param(name, value)
        with pytest.raises(_ffi.AccError):
            rg.set_param(99, 1)


@pytest.mark.parametrize("quick", [True, False])
def test_tune_keeps_counts(ctx, quick):
    g = G.gen_rmat(16, 16, seed=43)
    ref = oracle.cn_merge(g)
    with ctx.upload(g) as rg:
        rep = rg.tune(quick=quick)
        print(f"[tune quick={quick}] {rep['path']} ms={rep['ms']:.3f} trials={len(rep['trials'])}", flush=True)
        assert rep["path"] in PATHS
        for key in ("fwd_unroll", "wave_slots", "wave_low", "chunk", "persist"):
            assert rg.get_param(key) == rep[key]
        assert len(rep["trials"]) >= 3
        cnt, t = rg.count_all_edges("auto")
    assert np.array_equal(cnt, ref)


def test_tuned_batch_capture(ctx, small_batch):
    graphs, refs = small_batch
    packed, vo = G.pack(graphs)
    with ctx.upload(packed) as rg:
        rg.set_partition(vo)
        rg.tune(quick=True)
        rg.capture("auto")
        rg.replay()
        ctx.synchronize()
        cnt, _ = rg.download()
        check_batch(graphs, refs, packed, cnt, rg.graph_totals())


def test_partition_rejected(ctx):
    from acctricnt_amd import _ffi

    a, b = clique(5), clique(4)
    packed, vo = G.pack([a, b])
    with ctx.upload(packed) as rg:
        for bad in ([0, 5, 8], [0, 6, 9], [1, 5, 9], [0, 5, 4, 9]):
            with pytest.raises(_ffi.AccError):
                rg.set_partition(np.array(bad, np.uint32))
        rg.set_partition(vo)
        with pytest.raises(_ffi.AccError):
            rg.graph_totals()  # nothing computed since the partition was set
        rg.compute("forward")
        assert list(rg.graph_totals()) == [10, 4]
        rg.set_partition(np.array([0, 9], np.uint32))  # one graph holding both cliques
        rg.compute("merge")
        assert list(rg.graph_totals()) == [14]
    # an edge joining the two ranges
    g = G.from_edges(9, [0, 1, 0, 4], [1, 2, 2, 5])
    with ctx.upload(g) as rg:
        with pytest.raises(_ffi.AccError):
            rg.set_partition(np.array([0, 5, 9], np.uint32))


def test_changes_drop_capture(ctx):
    """A captured pass holds plan and partition addresses: changing a parameter
    or the partition drops it, and replay then reports an error."""
    from acctricnt_amd import _ffi

    a, b = clique(6), clique(5)
    packed, vo = G.pack([a, b])
    with ctx.upload(packed) as rg:
        rg.set_partition(vo)
        for change in (lambda: rg.set_param("chunk", 128), lambda: rg.set_param("persist", 2),
                       lambda: rg.set_partition(vo)):
            rg.capture("forward")
            rg.replay()
            ctx.synchronize()
            assert list(rg.graph_totals()) == [20, 10]
            change()
            with pytest.raises(_ffi.AccError):
                rg.replay()
        rg.capture("forward")
        rg.replay()
        ctx.synchronize()
        assert list(rg.graph_totals()) == [20, 10]


def test_full_batch_mix(ctx):
    """The 512-graph mix (vertex counts log-uniform in [2^8, 2^16], average degree in
    [4, 64], ER / R-MAT / Chung-Lu): per-graph totals equal the oracle's triangle
    counts, sum of counts = 6 T per graph, and bit-exact counts on 48 sampled graphs,
    through the captured pass."""
    graphs = G.batch_mix()
    packed, vo = G.pack(graphs)
    print(f"[batch512] n={packed.n} m={packed.m}", flush=True)
    with ctx.upload(packed) as rg:
        rg.set_partition(vo)
        rg.capture("forward")
        rg.replay()
        ctx.synchronize()
        cnt, t = rg.download()
        totals = rg.graph_totals()
    so = packed.meta["slot_off"]
    c64 = cnt.astype(np.int64) - 2
    ref_t = np.array([oracle.triangle_count(g) for g in graphs], np.uint64)
    assert np.array_equal(totals, ref_t)
    assert t == int(ref_t.sum())
    sums = np.add.reduceat(c64, so[:-1]) if packed.m else np.zeros(len(graphs))
    sums[np.diff(so) == 0] = 0
    assert np.array_equal(sums, 6 * ref_t.astype(np.int64))
    for k in np.random.default_rng(5).choice(len(graphs), 48, replace=False):
        assert np.array_equal(cnt[so[k]:so[k + 1]], oracle.cn_merge(graphs[k])), k
