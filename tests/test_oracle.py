"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Every oracle function is checked against something it does not compute
itself: closed forms, scipy.sparse products, brute-force set intersection,
a SCAN written from the definitions (tests/bruteforce.py), the golden
fixtures of tests/golden/golden.json, and one known answer from a run of the
reference (K4 plus a pendant vertex).  A planted mistake in the oracle's
output must be caught by the same comparisons.
"""
import hashlib
import json
import os
from math import comb

import numpy as np
import pytest

import bruteforce as BF
import oracle
from acctricnt_amd import graph as G

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden_graph(name):
    spec = GOLDEN[name]["spec"]
    k = spec["kind"]
    if k == "er":
        return G.gen_er(spec["n"], spec["avg_deg"], seed=spec["seed"])
    if k == "rmat":
        return G.gen_rmat(spec["scale"], spec["edge_factor"], seed=spec["seed"], relabel=spec["relabel"])
    if k == "chung_lu":
        return G.gen_chung_lu(spec["n"], spec["gamma"], spec["avg_deg"], spec["max_deg"], seed=spec["seed"])
    if k == "k4_pendant":
        return BF.k4_pendant()
    if k == "wheel":
        return BF.wheel(spec["k"])
    if k == "windmill":
        return BF.windmill(spec["k"])
    if k == "complete_bipartite":
        return BF.complete_bipartite(spec["a"], spec["b"])
    if k == "cliques":
        return BF.disjoint_cliques(spec["sizes"])
    raise KeyError(k)


@pytest.fixture(scope="module", autouse=True)
def _built(built):
    return built


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("n", [2, 3, 4, 7, 12])
def test_complete_graph(n):
    g = BF.complete(n)
    for cnt in (oracle.cn_merge(g), oracle.cn_bitmap(g)):
        assert (cnt == n).all()  # CN = n-2, stored +2
    assert oracle.triangle_count(g) == comb(n, 3)
    assert oracle.triangle_count_bitmap(g) == comb(n, 3)


@pytest.mark.parametrize("g", [BF.complete_bipartite(3, 5), BF.cycle(4), BF.cycle(9), BF.star(7)],
                         ids=lambda g: g.name)
def test_triangle_free(g):
    assert (oracle.cn_merge(g) == 2).all()
    assert (oracle.cn_bitmap(g) == 2).all()
    assert oracle.triangle_count(g) == 0
    assert oracle.triangle_count_bitmap(g) == 0


@pytest.mark.parametrize("k", [5, 8, 13])
def test_wheel(k):
    g = BF.wheel(k)
    cnt = oracle.cn_merge(g) - 2
    src = g.sources()
    hub_edge = (src == 0) | (g.edge_dst == 0)
    assert (cnt[hub_edge] == 2).all() and (cnt[~hub_edge] == 1).all()
    assert oracle.triangle_count(g) == k
    assert oracle.triangle_count_bitmap(g) == k


@pytest.mark.parametrize("k", [1, 4, 9])
def test_windmill(k):
    g = BF.windmill(k)
    assert (oracle.cn_merge(g) == 3).all()
    assert oracle.triangle_count(g) == k
    assert oracle.triangle_count_bitmap(g) == k


# ------------------------------------------------------------ golden fixtures
@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_counts(name):
    ent = GOLDEN[name]
    g = golden_graph(name)
    assert (g.n, g.m) == (ent["n"], ent["m"])
    cnt = oracle.cn_merge(g)
    assert sha(cnt) == ent["cnt_sha256"]
    assert sha(oracle.cn_bitmap(g)) == ent["cnt_sha256"]
    assert oracle.triangle_count(g) == ent["triangles"]
    assert oracle.triangle_count_bitmap(g) == ent["triangles"]


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_scan(name, tmp_path):
    ent = GOLDEN[name]
    g = golden_graph(name)
    cnt = oracle.cn_merge(g)
    for q in ent["scan"]:
        for exact in (True, False):
            res = oracle.scan(g, cnt, q["eps"], q["mu"], exact=exact)
            for key in ("core_num", "cluster_num", "hub_num", "outlier_num"):
                assert res[key] == q[key], (name, q, key, exact)
            assert sha(res["sim"]) == q["sim_sha256"]
            assert sha(res["label"]) == q["label_sha256"]
            path = tmp_path / "res.txt"
            oracle.write_result(str(path), g, res)
            text = path.read_bytes()
            assert hashlib.sha256(text).hexdigest() == q["result_sha256"], (name, q)
            if "result_lines" in q:
                assert text.decode().splitlines() == q["result_lines"]


def test_k4_pendant_reference_run(tmp_path):
    """Known answer from a run of the reference binaries: Triangle Cnt 4;
    eps 0.6 mu 2 -> Core 4, Cluster 1, Hub 0, Outlier 1, result file with
    cores 0..3 in cluster 0 and the line 'n 4 0'."""
    g = BF.k4_pendant()
    assert oracle.triangle_count(g) == 4
    res = oracle.scan(g, oracle.cn_merge(g), "0.6", 2)
    assert (res["core_num"], res["cluster_num"], res["hub_num"], res["outlier_num"]) == (4, 1, 0, 1)
    p = tmp_path / f"scanxp-result-{oracle.eps_string(0.6)}-2.txt"
    oracle.write_result(str(p), g, res)
    assert p.name == "scanxp-result-0.6-2.txt"
    assert p.read_text().splitlines() == ["c/n vertex_id cluster_id", "c 0 0", "c 1 0", "c 2 0", "c 3 0",
                                          "n 4 0"]


# ------------------------------------------------------------ scipy / brute force
@pytest.mark.parametrize("maker", [
    lambda: G.gen_er(3000, 12, seed=3),
    lambda: G.gen_rmat(11, 8, seed=5),
    lambda: G.gen_rmat(10, 16, seed=6, relabel=False),
    lambda: G.gen_chung_lu(4000, 2.1, 10, 500, seed=9),
])
def test_counts_vs_scipy(maker):
    g = maker()
    ref = BF.cn_scipy(g)
    assert np.array_equal(oracle.cn_merge(g), ref)
    assert np.array_equal(oracle.cn_bitmap(g), ref)
    t = BF.triangles_scipy(g)
    assert oracle.triangle_count(g) == t
    assert oracle.triangle_count_bitmap(g) == t


@pytest.mark.parametrize("maker", [
    lambda: G.gen_rmat(12, 16, seed=1),
    lambda: G.gen_chung_lu(5000, 2.1, 16, 1000, seed=2),
    lambda: G.gen_rmat(11, 8, seed=3, relabel=False),
], ids=["rmat12", "chunglu5000", "rmat11_unrelabelled"])
def test_merge_slots_vs_scipy(maker):
    """oracle_cn_merge_slots against scipy's ((A·A)∘A)[u,v] + 2: given every
    u<v slot, it writes every slot of the graph; given a random subset, it
    writes exactly those slots and their reverse slots, each equal to scipy's
    count, and leaves every other slot untouched."""
    g = maker()
    ref = BF.cn_scipy(g)
    src = g.sources()
    up = np.nonzero(src < g.edge_dst)[0]
    full = np.full(g.m, -1, np.int32)
    oracle.cn_merge_slots(g, full, up, src[up])
    assert np.array_equal(full, ref)
    rng = np.random.default_rng(7)
    sel = rng.choice(up, up.size // 5, replace=False)
    part = np.full(g.m, -1, np.int32)
    oracle.cn_merge_slots(g, part, sel, src[sel])
    want = np.zeros(g.m, bool)
    want[sel] = True
    want[g.reverse_slots()[sel]] = True
    assert np.array_equal(part >= 0, want)
    assert np.array_equal(part[want], ref[want])
    # a planted off-by-one in one sampled slot is caught by the same comparison
    bad = part.copy()
    bad[sel[0]] += 1
    assert not np.array_equal(bad[want], ref[want])


@pytest.mark.parametrize("seed", range(6))
def test_counts_vs_sets_tiny(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    k = int(rng.integers(0, n * 3))
    g = G.from_edges(n, rng.integers(0, n, k), rng.integers(0, n, k))
    assert np.array_equal(oracle.cn_merge(g), BF.cn_bruteforce(g))


def test_invariants_rmat():
    g = G.gen_rmat(14, 16, seed=1)
    cnt = oracle.cn_merge(g).astype(np.int64)
    rev = g.reverse_slots()
    assert np.array_equal(cnt, cnt[rev])
    d = g.degrees()
    src = g.sources()
    assert ((cnt - 2) <= np.minimum(d[src], d[g.edge_dst]) - 1).all()
    t = oracle.triangle_count(g)
    assert (cnt - 2).sum() == 6 * t


def test_thread_count_invariance():
    g = G.gen_rmat(13, 16, seed=2)
    a = oracle.cn_merge(g, nthreads=1)
    b = oracle.cn_merge(g, nthreads=8)
    assert np.array_equal(a, b)
    assert oracle.triangle_count(g, 1) == oracle.triangle_count(g, 8)


def test_merge_range_partition():
    """Counting source ranges separately and merging equals the full count."""
    g = G.gen_rmat(12, 16, seed=4)
    full = oracle.cn_merge(g)
    part = np.full(g.m, 2, np.int32)
    cuts = [0, 7, 100, 1000, g.n]
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        oracle.cn_merge_range(g, part, lo, hi)
    assert np.array_equal(part, full)


def test_planted_mistake_is_caught():
    """An off-by-one planted in one slot of the oracle's output breaks every check."""
    g = G.gen_rmat(10, 16, seed=8)
    cnt = oracle.cn_merge(g)
    bad = cnt.copy()
    bad[g.m // 2] += 1
    assert not np.array_equal(bad, BF.cn_scipy(g))
    rev = g.reverse_slots()
    assert not np.array_equal(bad, bad[rev])


# ------------------------------------------------------------ degree relabel
def relabel_properties_hold(g, off, adj, old) -> bool:
    """The properties that determine a degree relabel uniquely: old is a
    permutation, new degrees are old degrees through it, non-increasing, equal
    degrees in ascending old id, every list strictly increasing, and the edge
    set maps back onto the original one."""
    n = g.n
    if sorted(old.tolist()) != list(range(n)) or off.shape[0] != n + 1 or adj.shape[0] != g.m:
        return False
    d_old, d_new = g.degrees(), np.diff(off.astype(np.int64))
    if not np.array_equal(d_old[old], d_new) or (np.diff(d_new) > 0).any():
        return False
    if ((np.diff(d_new) == 0) & (np.diff(old.astype(np.int64)) < 0)).any():
        return False
    src = np.repeat(np.arange(n), d_new)
    same_src = src[1:] == src[:-1]
    if (np.diff(adj.astype(np.int64))[same_src] <= 0).any():
        return False
    mapped = set(zip(old[src].tolist(), old[adj].tolist()))
    return mapped == set(zip(g.sources().tolist(), g.edge_dst.tolist()))


RELABEL_GRAPHS = [
    lambda: G.gen_rmat(10, 16, seed=71),
    lambda: G.gen_er(3000, 6, seed=72),
    lambda: G.gen_chung_lu(4000, 2.1, 10, 600, seed=73),
    lambda: G.from_edges(9, [0, 0, 0, 0, 5, 6], [1, 2, 3, 4, 6, 7]),  # star, path, isolated 8
    lambda: G.from_edges(6, [0, 1, 2, 3, 4, 5], [1, 2, 3, 4, 5, 0]),  # cycle: every degree equal
    lambda: G.from_edges(4, [], []),
]


@pytest.mark.parametrize("maker", RELABEL_GRAPHS)
def test_relabel_restatement_pinned(maker):
    g = maker()
    off, adj, old = oracle.relabel_degree(g)
    assert relabel_properties_hold(g, off, adj, old)
    # brute force from the definition on the small ones
    if g.n <= 16:
        d = g.degrees()
        order = sorted(range(g.n), key=lambda v: (-int(d[v]), v))
        new = {v: i for i, v in enumerate(order)}
        lists = [sorted(new[int(w)] for w in g.edge_dst[g.node_off[v]:g.node_off[v + 1]]) for v in order]
        assert old.tolist() == order
        assert adj.tolist() == [w for lst in lists for w in lst]
    # the product's host relabel (libacc_graph) agrees
    h = G.relabel_degree(g)
    assert np.array_equal(h.node_off, off) and np.array_equal(h.edge_dst, adj)
    assert np.array_equal(h.old_of_new, old)


def test_relabel_planted_mistakes_are_caught():
    """Swapping two equal-degree vertices, or leaving one list unsorted, breaks the properties."""
    g = G.gen_rmat(9, 8, seed=74)
    off, adj, old = oracle.relabel_degree(g)
    d = np.diff(off.astype(np.int64))
    i = int(np.nonzero(np.diff(d) == 0)[0][0])
    bad_old = old.copy()
    bad_old[[i, i + 1]] = bad_old[[i + 1, i]]
    assert not relabel_properties_hold(g, off, adj, bad_old)
    v = int(np.nonzero(d >= 2)[0][0])
    bad_adj = adj.copy()
    a, b = int(off[v]), int(off[v])
    bad_adj[[a, b + 1]] = bad_adj[[b + 1, a]]
    assert not relabel_properties_hold(g, off, bad_adj, old)


# ------------------------------------------------------------ SCAN semantics
@pytest.mark.parametrize("eps,mu", [("0.2", 5), ("0.5", 3), ("0.7", 2), ("0.35", 4)])
def test_scan_vs_bruteforce(eps, mu):
    g = G.gen_rmat(10, 12, seed=12)
    cnt = oracle.cn_merge(g)
    label, cid, counts, lines, sim = BF.scan_bruteforce(g, cnt, eps, mu)
    res = oracle.scan(g, cnt, eps, mu)
    assert np.array_equal(res["sim"].astype(bool), sim)
    assert np.array_equal(res["label"], label)
    core = label == 1
    assert np.array_equal(res["cluster_id"][core], cid[core])
    assert {k: res[k] for k in counts} == counts


def test_scan_with_relabel_maps_old_ids(tmp_path):
    """Cluster ids and result lines use the original ids after a degree relabel."""
    g = G.gen_rmat(9, 12, seed=13, relabel=True)
    cnt = oracle.cn_merge(g)
    label, cid, counts, lines, sim = BF.scan_bruteforce(g, cnt, "0.4", 3, old_id=g.old_of_new)
    res = oracle.scan(g, cnt, "0.4", 3, old_id=g.old_of_new)
    p = tmp_path / "r.txt"
    oracle.write_result(str(p), g, res, old_id=g.old_of_new)
    assert p.read_text().splitlines() == lines


def test_integer_similarity_equals_long_double():
    """The exact integer test equals the reference's long-double test on every
    slot of the ER and R-MAT configurations for eps = 0.1 .. 0.9."""
    for g in (G.gen_er(16384, 16, seed=20190101), G.gen_rmat(13, 16, seed=3)):
        cnt = oracle.cn_merge(g)
        for k in range(1, 10):
            eps = f"0.{k}"
            si, _, li = oracle.scan_core(g, cnt, eps, 5, exact=True)
            sf, _, lf = oracle.scan_core(g, cnt, eps, 5, exact=False)
            assert np.array_equal(si, sf), eps
            assert np.array_equal(li, lf)


def test_similarity_boundary_cases():
    """cn/sqrt(D) exactly equal to eps is similar in both decisions."""
    # cn = 3, D = 25 -> 3/5 = 0.6 exactly; cn = 4, D = 100 -> 0.4; cn = 2, D = 4 -> 1.0
    for c, du1, dv1, eps in [(3, 5, 5, "0.6"), (4, 10, 10, "0.4"), (2, 2, 2, "1.0"), (3, 25, 1, "0.6")]:
        assert BF.sim_reference_fp(c, du1, dv1, float(eps))
        fr = oracle.eps_fraction(eps)
        assert fr.denominator ** 2 * c * c >= fr.numerator ** 2 * du1 * dv1


def test_complete_graph_scan():
    g = BF.complete(6)
    cnt = oracle.cn_merge(g)
    res = oracle.scan(g, cnt, "0.9", 5)
    assert res["core_num"] == 6 and res["cluster_num"] == 1
    assert (res["cluster_id"] == 0).all()
    res = oracle.scan(g, cnt, "0.9", 6)
    assert res["core_num"] == 0 and res["outlier_num"] == 6


def test_disjoint_cliques_scan():
    g = BF.disjoint_cliques([3, 4, 5])
    res = oracle.scan(g, oracle.cn_merge(g), "0.9", 2)
    assert res["cluster_num"] == 3
    assert sorted(set(res["cluster_id"].tolist())) == [0, 3, 7]


def test_hub_detection():
    """Vertex 8 joins two separate K4 clusters without being similar enough to be core."""
    s, d = [], []
    for base in (0, 4):
        a, b = np.triu_indices(4, 1)
        s += (a + base).tolist()
        d += (b + base).tolist()
    s += [8, 8]
    d += [0, 4]
    g = G.from_edges(9, s, d)
    res = oracle.scan(g, oracle.cn_merge(g), "0.6", 3)
    assert res["label"][8] == oracle.HUB
    assert res["hub_num"] == 1 and res["cluster_num"] == 2


@pytest.mark.parametrize("eps,expect", [(0.2, "0.2"), (0.25, "0.25"), (0.5, "0.5"), (1.0, "1."), (0.123456, "0.123456")])
def test_eps_string(eps, expect):
    assert oracle.eps_string(eps) == expect == ("%f" % eps).rstrip("0u")
