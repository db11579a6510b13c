"""SCAN serving over resident graphs (acc_server_*), through the C ABI.

Every served query must return exactly the labels, cluster ids and counters
of the oracle's SCAN for that graph and (eps, mu), and of the same query run
alone through a SCAN workspace, whatever the other queries in its batch and
whatever was admitted or evicted meanwhile.
"""
import threading

import numpy as np
import pytest

import bruteforce as BF
import oracle
from acctricnt_amd import _ffi, cn, graph as G, serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(built):
    assert cn.device_count() > 0, "no HIP device visible"
    c = cn.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def graphs():
    gs = [G.gen_rmat(12, 16, seed=61), G.gen_er(3000, 12, seed=62), G.gen_chung_lu(5000, 2.1, 16, 800, seed=63),
          G.gen_rmat(14, 16, seed=64), G.from_edges(5, [0, 0, 0, 1, 1, 2, 3], [1, 2, 3, 2, 3, 3, 4])]
    cnts = [oracle.cn_merge(g) for g in gs]
    return gs, cnts


QUERIES = [(e, mu) for e in ("0.1", "0.3", "0.5", "0.7", "0.9") for mu in (2, 5, 10)]


def check(res, g, cnt, eps, mu):
    ref = oracle.scan(g, cnt, eps, mu)
    assert np.array_equal(res["label"], ref["label"]), f"labels differ eps={eps} mu={mu}"
    core = ref["label"] == 1
    assert np.array_equal(res["cluster_id"][core], ref["cluster_id"][core])
    assert (res["cluster_id"][~core] == -1).all()
    for k in ("core_num", "cluster_num", "hub_num", "outlier_num"):
        assert res[k] == ref[k], k


@pytest.mark.parametrize("index", [True, False])
def test_serve_matches_oracle_and_alone(ctx, graphs, index):
    gs, cnts = graphs
    rng = np.random.default_rng(7)
    with serve.Server(ctx, max_batch=8, streams=3, slots=48, index=index) as sv:
        gids = [sv.add_graph(g) for g in gs]
        # the mix: all queries submitted at once, batched across graphs
        mix = [(int(rng.integers(len(gs))), *QUERIES[int(rng.integers(len(QUERIES)))]) for _ in range(120)]
        tickets = [sv.submit(gids[k], eps, mu) for k, eps, mu in mix]
        results = []
        for t in tickets:
            sv.wait(t)
            results.append(sv.result(t))
            sv.release(t)
        st = sv.stats()
        print(f"[serve] {st}", flush=True)
        assert st["completed"] == len(mix) and st["failed"] == 0
        assert st["max_batch_seen"] > 1, "queries were never batched"
        assert st["index_queries"] == (len(mix) if index else 0)
        # the same queries alone, through a workspace
        alone = {}
        for k, eps, mu in set(mix):
            with ctx.upload(gs[k]) as rg, rg.scan_workspace() as ws:
                ws.run(eps, mu)
                ws.synchronize()
                alone[(k, eps, mu)] = ws.results()
        for (k, eps, mu), res in zip(mix, results):
            a = alone[(k, eps, mu)]
            assert np.array_equal(res["label"], a["label"]) and np.array_equal(res["cluster_id"], a["cluster_id"])
        for k, eps, mu in sorted(set(mix))[:40]:
            i = mix.index((k, eps, mu))
            check(results[i], gs[k], cnts[k], eps, mu)


def test_serve_threads(ctx, graphs):
    gs, cnts = graphs
    with serve.Server(ctx, max_batch=4, streams=2, slots=16) as sv:
        gids = [sv.add_graph(g) for g in gs]
        out = {}
        errors = []

        def client(c):
            try:
                for i in range(12):
                    k = (c + i) % len(gs)
                    eps, mu = QUERIES[(3 * c + i) % len(QUERIES)]
                    out[(c, i)] = (k, eps, mu, sv.query(gids[k], eps, mu))
            except Exception as e:  # reported below
                errors.append(e)

        th = [threading.Thread(target=client, args=(c,)) for c in range(6)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        for k, eps, mu, res in out.values():
            check(res, gs[k], cnts[k], eps, mu)


def test_serve_eviction(ctx, graphs):
    """A budget that holds one of two equal-sized graphs at a time (plus one
    result slot): alternating queries evict and re-admit them from the
    server's host copies, and every answer stays exact."""
    gs, cnts = graphs
    pair = [gs[3], G.gen_rmat(14, 16, seed=65)]
    pair_cnt = [cnts[3], oracle.cn_merge(pair[1])]
    with serve.Server(ctx, slots=1) as probe:
        gid = probe.add_graph(pair[1])
        probe.query(gid, "0.5", 5)
        foot = max(probe.graph_info(gid)["device_bytes"], 1)
        slot = probe.stats()["slot_bytes"]
    budget = int(1.6 * foot) + slot + (1 << 16)
    with serve.Server(ctx, hbm_budget=budget, max_batch=2, streams=1, slots=1) as sv:
        gids = [sv.add_graph(g) for g in pair]
        for i in range(6):
            k = i % 2
            eps, mu = QUERIES[i]
            check(sv.query(gids[k], eps, mu), pair[k], pair_cnt[k], eps, mu)
        st = sv.stats()
        print(f"[serve eviction] budget={budget} {st}", flush=True)
        assert st["evictions"] >= 1 and st["admissions"] >= 3
        assert st["graph_bytes"] + st["slot_bytes"] <= budget


def test_serve_errors(ctx, graphs):
    gs, _ = graphs
    with serve.Server(ctx, hbm_budget=4 << 20, slots=2) as sv:
        with pytest.raises(_ffi.AccError) as e:
            sv.add_graph(gs[3])  # larger than the whole budget
        assert e.value.code == -6
        gid = sv.add_graph(gs[4])
        with pytest.raises(_ffi.AccError) as e:
            sv.submit(gid + 100, "0.5", 2)
        assert e.value.code == -7
        with pytest.raises(_ffi.AccError):
            sv.submit(gid, (-1, 2), 2)
        with pytest.raises(_ffi.AccError) as e:
            sv.wait(12345, 0)
        assert e.value.code == -7
        t = sv.submit(gid, "0.6", 2)
        sv.wait(t)
        r = sv.result(t)
        assert list(r["label"]) == [1, 1, 1, 1, 0] and r["core_num"] == 4
        sv.release(t)
        with pytest.raises(_ffi.AccError):
            sv.result(t)
        # served graphs take no upload flag but ACC_UPLOAD_NO_VALIDATE (relabel before adding)
        import ctypes as C

        g4 = gs[4]
        off = np.ascontiguousarray(g4.node_off, np.uint32)
        adj = np.ascontiguousarray(g4.edge_dst, np.int32)
        gid_out = C.c_int32()
        rc = sv._lib.acc_server_add_graph(sv._h, g4.n, g4.m, _ffi.ptr(off, C.c_uint32), _ffi.ptr(adj, C.c_int32),
                                          cn.UPLOAD_RELABEL, C.byref(gid_out))
        assert rc == -1
        # a malformed graph is rejected at registration
        bad = G.CSRGraph(np.array([0, 2, 4, 6], np.uint32), np.array([2, 1, 0, 2, 0, 1], np.int32))
        with pytest.raises(_ffi.AccError) as e:
            sv.add_graph(bad)
        assert e.value.code == -2
        sv.remove_graph(gid)
        with pytest.raises(_ffi.AccError):
            sv.submit(gid, "0.5", 2)


def test_serve_cancel_pending(ctx, graphs):
    gs, cnts = graphs
    with serve.Server(ctx, max_batch=1, streams=1, slots=1) as sv:
        gid = sv.add_graph(gs[3])
        ts = [sv.submit(gid, "0.3", 5) for _ in range(8)]
        # the last ones are still pending behind the single slot: cancel one
        sv.release(ts[-1])
        for t in ts[:-1]:
            sv.wait(t)
            check(sv.result(t), gs[3], cnts[3], "0.3", 5)
            sv.release(t)
        st = sv.stats()
        assert st["cancelled"] + st["completed"] == 8


def cube():
    s = [u for u in range(8) for b in range(3) if u < u ^ (1 << b)]
    d = [u ^ (1 << b) for u in range(8) for b in range(3) if u < u ^ (1 << b)]
    return G.from_edges(8, s, d, name="Q3")


def test_serve_index_ties_and_fallback(ctx):
    """Similarities equal to eps take the exact test: the cube Q3 has
    cn^2 / D = 4/16 on every edge, so eps = 0.5 makes every edge similar and
    eps = 0.51 none; K_7 has similarity 1 everywhere (eps = 1).  An eps whose
    numerator and denominator exceed 2^26 is outside the index and takes the
    per-slot kernels in the same batch."""
    from fractions import Fraction

    k7 = G.from_edges(7, *np.triu_indices(7, 1))
    rm = G.gen_rmat(12, 8, seed=66)
    cases = [(cube(), "0.5", 3), (cube(), "0.51", 1), (cube(), "0.5", 4), (k7, "1", 6), (k7, "1", 7),
             (rm, Fraction(123456789, 987654321), 2), (rm, "0.3", 2), (rm, Fraction(67108865, 134217729), 3)]
    with serve.Server(ctx, max_batch=8, streams=1, slots=8) as sv:
        ids = {}
        for g, _, _ in cases:
            if id(g) not in ids:
                ids[id(g)] = sv.add_graph(g)
        ts = [sv.submit(ids[id(g)], eps, mu) for g, eps, mu in cases]
        for (g, eps, mu), t in zip(cases, ts):
            sv.wait(t)
            check(sv.result(t), g, oracle.cn_merge(g), eps, mu)
            sv.release(t)
        st = sv.stats()
        assert st["index_queries"] == len(cases) - 2
    assert oracle.scan(cube(), oracle.cn_merge(cube()), "0.5", 3)["core_num"] == 8
    assert oracle.scan(cube(), oracle.cn_merge(cube()), "0.51", 1)["core_num"] == 0


def test_serve_remove_graph_with_unreleased_result(ctx):
    """A finished query whose graph is removed before its ticket is released:
    its result stays readable with the graph's own size, the release resets
    every cluster id its cores held, so the next query on the same (single)
    result slot reports -1 for all its non-cores, and the removed id stays
    unknown."""
    a = BF.complete(40)  # every vertex a core of one cluster at eps 0.5, mu 2
    b = G.gen_er(3000, 4, seed=67)  # few cores at eps 0.9
    with serve.Server(ctx, max_batch=1, streams=1, slots=1) as sv:
        ga = sv.add_graph(a)
        gb = sv.add_graph(b)
        t = sv.submit(ga, "0.5", 2)
        sv.wait(t)
        sv.remove_graph(ga)
        r = sv.result(t)
        assert r["label"].shape == (a.n,) and r["core_num"] == a.n
        assert (r["cluster_id"] == 0).all()
        sv.release(t)
        with pytest.raises(_ffi.AccError) as e:
            sv.submit(ga, "0.5", 2)
        assert e.value.code == -7
        res = sv.query(gb, "0.9", 2)
        check(res, b, oracle.cn_merge(b), "0.9", 2)


def test_serve_wait_timeout(ctx, graphs):
    """acc_server_wait returns ACC_ETIMEDOUT while a query cannot run: with one
    result slot held by a finished, unreleased query, the next query stays
    pending; once the slot is released it runs and its result is exact."""
    gs, cnts = graphs
    with serve.Server(ctx, max_batch=1, streams=1, slots=1) as sv:
        gid = sv.add_graph(gs[0])
        t1 = sv.submit(gid, "0.3", 5)
        sv.wait(t1)
        t2 = sv.submit(gid, "0.5", 2)
        with pytest.raises(_ffi.AccError) as e:
            sv.wait(t2, 200)
        assert e.value.code == -8
        assert sv.poll(t2) == 0  # ACC_QUERY_PENDING
        check(sv.result(t1), gs[0], cnts[0], "0.3", 5)
        sv.release(t1)
        sv.wait(t2, 60_000)
        check(sv.result(t2), gs[0], cnts[0], "0.5", 2)
        sv.release(t2)


def test_serve_configured_workload(ctx):
    """The serving configuration: 16 resident R-MAT graphs (scales drawn
    uniformly from 16..22 with the benchmark's seed, edge factor 16,
    degree-relabelled), every (eps, mu) of eps in {0.1, ..., 0.9} and mu in
    {2, 5, 10, 20} submitted at once in a shuffled order, so that many
    batches across graphs are in flight together, all through the similarity
    index.  Each served result equals the oracle's SCAN over the oracle's
    counts: every (eps, mu) on the graphs of scale <= 20, a fixed third of
    them on the larger ones; a sample also equals the same query run alone."""
    rng = np.random.default_rng(5)
    scales = [int(x) for x in rng.integers(16, 23, size=16)]
    gs = [G.gen_rmat(sc, 16, seed=1000 + i) for i, sc in enumerate(scales)]
    print(f"[serve configured] scales {scales}", flush=True)
    epss = [f"0.{k}" for k in range(1, 10)]
    mus = [2, 5, 10, 20]
    combos = [(k, e, mu) for k in range(len(gs)) for e in epss for mu in mus
              if scales[k] <= 20 or (epss.index(e) + mus.index(mu)) % 3 == 0]
    order = np.random.default_rng(11).permutation(len(combos))
    with serve.Server(ctx, max_batch=14, streams=4, slots=64) as sv:
        gids = [sv.add_graph(g) for g in gs]
        assert sv.stats()["resident_graphs"] == len(gs)
        tickets = [(combos[i], sv.submit(gids[combos[i][0]], combos[i][1], combos[i][2])) for i in order]
        results = {}
        for (k, eps, mu), t in tickets:
            sv.wait(t)
            results[(k, eps, mu)] = sv.result(t)
            sv.release(t)
        st = sv.stats()
        print(f"[serve configured] {st}", flush=True)
        assert st["completed"] == len(combos) and st["failed"] == 0 and st["evictions"] == 0
        assert st["index_queries"] == len(combos)
        assert st["max_batch_seen"] > 1
    for k in range(len(gs)):
        cnt = oracle.cn_bitmap(gs[k])
        mine = [c for c in combos if c[0] == k]
        print(f"[serve configured] graph {k} scale {scales[k]}: {len(mine)} queries", flush=True)
        for c in mine:
            check(results[c], gs[k], cnt, c[1], c[2])
            del results[c]
    sample = [combos[i] for i in order[:6]]
    for k, eps, mu in sample:
        with ctx.upload(gs[k]) as rg, rg.scan_workspace() as ws:
            ws.run(eps, mu)
            ws.synchronize()
            a = ws.results()
        ref = oracle.scan(gs[k], oracle.cn_bitmap(gs[k]), eps, mu)
        assert np.array_equal(a["label"], ref["label"]), (k, eps, mu)
