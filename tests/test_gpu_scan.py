"""Parity of the GPU SCAN engine (acc_scan_*) with the oracle's SCAN pipeline
and the golden fixtures: similarity bits, labels (core / hub / outlier),
cluster ids, the four counters and the result file, byte for byte."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from acctricnt_amd import cn, graph as G

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx(built):
    assert cn.device_count() > 0, "no HIP device visible"
    c = cn.Context(0)
    yield c
    c.close()


def gpu_scan(rg, ws, eps, mu):
    ws.run(eps, mu)
    ws.synchronize()
    res = ws.results(sim=True, core_count=True)
    res["pairs"] = ws.pairs()
    return res


def check_vs_oracle(g, res, eps, mu, old=None):
    ref = oracle.scan(g, oracle.cn_merge(g), eps, mu, old_id=old)
    assert np.array_equal(res["sim"], ref["sim"]), "similarity bits differ"
    assert np.array_equal(res["core_count"], ref["core_count"])
    assert np.array_equal(res["label"], ref["label"]), "labels differ"
    core = ref["label"] == 1
    assert np.array_equal(res["cluster_id"][core], ref["cluster_id"][core])
    for k in ("core_num", "cluster_num", "hub_num", "outlier_num"):
        assert res[k] == ref[k], k
    return ref


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_scan_golden(ctx, name, tmp_path):
    from test_oracle import golden_graph

    g = golden_graph(name)
    with ctx.upload(g) as rg, rg.scan_workspace() as ws:
        for q in GOLDEN[name]["scan"]:
            print(f"[scan golden] {name} eps={q['eps']} mu={q['mu']}", flush=True)
            res = gpu_scan(rg, ws, q["eps"], q["mu"])
            assert sha(res["sim"]) == q["sim_sha256"]
            assert sha(res["label"]) == q["label_sha256"]
            for k in ("core_num", "cluster_num", "hub_num", "outlier_num"):
                assert res[k] == q[k], (name, q, k)
            path = tmp_path / "r.txt"
            cn.write_scan_result(str(path), g.n, res["label"], res["cluster_id"], res["pairs"])
            assert hashlib.sha256(path.read_bytes()).hexdigest() == q["result_sha256"]


@pytest.mark.parametrize("maker,queries", [
    (lambda: G.gen_rmat(16, 16, seed=41), [("0.1", 2), ("0.3", 5), ("0.5", 10), ("0.9", 20)]),
    (lambda: G.gen_chung_lu(1 << 15, 2.1, 32, 1 << 12, seed=42), [("0.3", 5), ("0.6", 5)]),
    (lambda: G.gen_er(16384, 16, seed=20190101), [("0.2", 5)]),
], ids=["rmat16", "chunglu15", "er16384"])
def test_scan_vs_oracle(ctx, maker, queries):
    g = maker()
    with ctx.upload(g) as rg, rg.scan_workspace() as ws:
        for eps, mu in queries:
            print(f"[scan] {g.name} eps={eps} mu={mu}", flush=True)
            check_vs_oracle(g, gpu_scan(rg, ws, eps, mu), eps, mu)


def test_scan_relabelled_old_ids(ctx, tmp_path):
    """Degree-relabelled graph: cluster ids and result lines in original ids."""
    g = G.gen_rmat(14, 16, seed=43, relabel=True)
    old = g.old_of_new
    with ctx.upload(g) as rg, rg.scan_workspace() as ws:
        rg.set_vertex_ids(old)
        res = gpu_scan(rg, ws, "0.4", 4)
    ref = check_vs_oracle(g, res, "0.4", 4, old=old)
    a, b = tmp_path / "gpu.txt", tmp_path / "oracle.txt"
    cn.write_scan_result(str(a), g.n, res["label"], res["cluster_id"], res["pairs"], old_id=old)
    oracle.write_result(str(b), g, ref, old_id=old)
    assert a.read_bytes() == b.read_bytes()


def test_scan_repeated_queries_independent(ctx):
    """Queries on one workspace leave no state behind: a query after others equals a fresh one."""
    g = G.gen_rmat(14, 16, seed=44)
    with ctx.upload(g) as rg:
        with rg.scan_workspace() as ws:
            for eps, mu in (("0.9", 2), ("0.1", 20), ("0.5", 5)):
                ws.run(eps, mu)
            ws.synchronize()
            mixed = ws.results(sim=True)
        with rg.scan_workspace() as ws2:
            alone = gpu_scan(rg, ws2, "0.5", 5)
    for k in ("label", "cluster_id", "sim"):
        assert np.array_equal(mixed[k], alone[k]), k


def test_scan_full_size_rmat22(ctx):
    """R-MAT scale 22: GPU SCAN equals the oracle's SCAN for eps 0.2 / mu 5."""
    g = G.gen_rmat(22, 16, seed=1)
    print(f"[scan rmat22] n={g.n} m={g.m}", flush=True)
    with ctx.upload(g) as rg, rg.scan_workspace() as ws:
        cnt, _ = rg.count_all_edges()
        res = gpu_scan(rg, ws, "0.2", 5)
    ref = oracle.scan(g, cnt, "0.2", 5)
    assert np.array_equal(res["sim"], ref["sim"])
    assert np.array_equal(res["label"], ref["label"])
    core = ref["label"] == 1
    assert np.array_equal(res["cluster_id"][core], ref["cluster_id"][core])
    for k in ("core_num", "cluster_num", "hub_num", "outlier_num"):
        assert res[k] == ref[k], k


def test_scan_repeated_low_eps_rmat18(ctx):
    """Many unions (eps 0.1, mu 2 on R-MAT scale 18), repeated: every run must give
    the oracle's labels and cluster ids (a root pointer overwritten by a racing
    path-halving store used to leave a core outside its root now and then)."""
    g = G.gen_rmat(18, 16, seed=1018)
    ref = oracle.scan(g, oracle.cn_merge(g), "0.1", 2)
    core = ref["label"] == 1
    want_cid = np.where(core, ref["cluster_id"], -1)
    with ctx.upload(g) as rg, rg.scan_workspace() as ws:
        for rep in range(8):
            ws.run("0.1", 2)
            ws.synchronize()
            res = ws.results()
            assert np.array_equal(res["label"], ref["label"]), rep
            assert np.array_equal(res["cluster_id"], want_cid), rep
