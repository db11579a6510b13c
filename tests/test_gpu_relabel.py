"""Degree relabel on the device (ACC_UPLOAD_RELABEL), through the C ABI.

The resident CSR and old ids after the relabel must equal the oracle's
relabel (new ids by degree, descending, ties by ascending old id; every list
mapped and sorted), and counts, triangle totals and SCAN on the relabelled
resident graph must equal the oracle's on the relabelled CSR, with cluster ids
in original ids.
"""
import numpy as np
import pytest

import oracle
from acctricnt_amd import _ffi, cn, graph as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(built):
    assert cn.device_count() > 0, "no HIP device visible"
    c = cn.Context(0)
    yield c
    c.close()


GRAPHS = {
    "rmat14_raw": lambda: G.gen_rmat(14, 16, seed=81),
    "chunglu15_raw": lambda: G.gen_chung_lu(1 << 15, 2.1, 24, 1 << 14, seed=82),
    "er16384": lambda: G.gen_er(16384, 16, seed=83),
    "star_path_isolated": lambda: G.from_edges(9, [0, 0, 0, 0, 5, 6], [1, 2, 3, 4, 6, 7]),
    "cycle": lambda: G.from_edges(6, [0, 1, 2, 3, 4, 5], [1, 2, 3, 4, 5, 0]),
    "no_edges": lambda: G.from_edges(4, [], []),
    "one_vertex": lambda: G.from_edges(1, [], []),
    "already_relabelled": lambda: G.gen_rmat(13, 8, seed=84, relabel=True),
}


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_relabel_matches_oracle(ctx, name):
    g = GRAPHS[name]()
    print(f"[relabel] {name} n={g.n} m={g.m}", flush=True)
    off, adj, old = oracle.relabel_degree(g)
    with ctx.upload(g, relabel=True) as rg:
        steps = rg.timings()  # device time of the relabel's steps
        assert {"rl_degree_order", "rl_offsets", "rl_copy_offsets"} <= set(steps)
        assert all(t >= 0 for t in steps.values())
        r = rg.csr()
        assert np.array_equal(r.node_off, off)
        assert np.array_equal(r.edge_dst, adj)
        assert np.array_equal(r.old_of_new, old)
        h = G.CSRGraph(off, adj)
        if g.m:
            cnt, tri = rg.count_all_edges()
            assert np.array_equal(cnt, oracle.cn_merge(h))
            assert tri == oracle.triangle_count(h) == oracle.triangle_count(g)
        assert rg.triangle_total() == oracle.triangle_count(g)


def test_relabel_scan_reports_original_ids(ctx, tmp_path):
    """SCAN on a graph relabelled on the device: labels per new id, cluster ids
    and the result file in original ids, as SCAN_XP with the deg reorder."""
    g = G.gen_rmat(14, 16, seed=85)
    off, adj, old = oracle.relabel_degree(g)
    h = G.CSRGraph(off, adj)
    ref = oracle.scan(h, oracle.cn_merge(h), "0.4", 4, old_id=old)
    with ctx.upload(g, relabel=True) as rg, rg.scan_workspace() as ws:
        ws.run("0.4", 4)
        ws.synchronize()
        res = ws.results(sim=True)
        pairs = ws.pairs()
    assert np.array_equal(res["sim"], ref["sim"])
    assert np.array_equal(res["label"], ref["label"])
    core = ref["label"] == 1
    assert np.array_equal(res["cluster_id"][core], ref["cluster_id"][core])
    a, b = tmp_path / "gpu.txt", tmp_path / "oracle.txt"
    cn.write_scan_result(str(a), g.n, res["label"], res["cluster_id"], pairs, old_id=old)
    oracle.write_result(str(b), h, ref, old_id=old)
    assert a.read_bytes() == b.read_bytes()


def test_relabel_rejects_bad_csr(ctx):
    """The relabel indexes by neighbor id, so it validates even when asked not to."""
    bad = G.CSRGraph(np.array([0, 1, 2], np.uint32), np.array([1, 5], np.int32))
    with pytest.raises(_ffi.AccError) as e:
        ctx.upload(bad, validate=False, relabel=True)
    assert e.value.code == -2


def test_relabel_full_size_rmat22(ctx):
    """R-MAT scale 22 as generated (ids not relabelled): the device relabel
    equals the host relabel of the product's graph library, array for array,
    and the triangle total is unchanged."""
    g = G.gen_rmat(22, 16, seed=1, relabel=False)
    print(f"[relabel rmat22] n={g.n} m={g.m}", flush=True)
    h = G.relabel_degree(g)
    with ctx.upload(g, relabel=True) as rg:
        r = rg.csr()
        assert np.array_equal(r.node_off, h.node_off)
        assert np.array_equal(r.edge_dst, h.edge_dst)
        assert np.array_equal(r.old_of_new, h.old_of_new)
        t_relabel = rg.triangle_total()
    with ctx.upload(g) as rg:
        assert rg.triangle_total() == t_relabel


def test_relabel_full_size_chunglu23(ctx):
    """Chung-Lu 2^23 as generated: the device relabel equals the host relabel."""
    g = G.gen_chung_lu(1 << 23, 2.1, 32, 1 << 20, seed=1, relabel=False)
    print(f"[relabel chunglu23] n={g.n} m={g.m}", flush=True)
    h = G.relabel_degree(g)
    with ctx.upload(g, relabel=True) as rg:
        r = rg.csr()
    assert np.array_equal(r.node_off, h.node_off)
    assert np.array_equal(r.edge_dst, h.edge_dst)
    assert np.array_equal(r.old_of_new, h.old_of_new)
