"""Shards of the forward path (acc_graph_set_shard), through the C ABI.

A share is the part of a pass one GPU of a multi-GPU run would do: the
sources in a contiguous vertex range of equal probe work.  Run one after the
other on one GPU, the shares' raw counts must add up to the whole graph's,
slot by slot, and their triangle totals to its total; the ranges must tile
[0, n) in order.
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
    "rmat16": lambda: G.gen_rmat(16, 16, seed=91),
    "chunglu15": lambda: G.gen_chung_lu(1 << 15, 2.1, 24, 1 << 14, seed=92),
    "er16384": lambda: G.gen_er(16384, 16, seed=93),
    "rmat14_unrelabelled": lambda: G.gen_rmat(14, 16, seed=94, relabel=False),
}


@pytest.mark.parametrize("parts", [2, 3, 8])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_shares_add_up_to_the_whole(ctx, name, parts):
    g = GRAPHS[name]()
    print(f"[shard] {name} n={g.n} m={g.m} parts={parts}", flush=True)
    ref = oracle.cn_merge(g).astype(np.int64) - 2
    t_ref = oracle.triangle_count(g)
    with ctx.upload(g) as rg:
        acc = np.zeros(g.m, np.int64)
        t_sum = 0
        ranges = []
        for s in range(parts):
            ranges.append(rg.set_shard(parts, s))
            c, t = rg.count_all_edges("forward", raw=True)
            assert rg.triangle_total("forward") == t  # the triangle-only pass over the same share
            acc += c
            t_sum += t
        assert rg.set_shard(1, 0) == (0, g.n)
        whole, t_whole = rg.count_all_edges("forward", raw=True)
    assert ranges[0][0] == 0 and ranges[-1][1] == g.n
    assert all(a[1] == b[0] for a, b in zip(ranges, ranges[1:]))
    assert np.array_equal(acc, ref) and t_sum == t_ref
    assert np.array_equal(whole.astype(np.int64), ref) and t_whole == t_ref


def test_shares_balance_probe_work(ctx):
    """Eight shares of an R-MAT graph: no share's estimated work (per upper
    slot (b, c) its probe length + 4 + twice the triangles with middle b and
    top c, |N<(b)| + 256 per 512-slot unit) exceeds an equal split by more
    than its heaviest single source."""
    import scipy.sparse as sp

    g = G.gen_rmat(17, 16, seed=95)
    src, dst = g.sources(), g.edge_dst
    up = src < dst
    off = g.node_off.astype(np.int64)
    # probe length of upper slot (b, c): |N(c) ∩ [0, b)| = reverse slot - start of N(c)
    lens = g.reverse_slots()[up].astype(np.int64) - off[dst[up]]
    # hits of probe (b, c): #{a < b : a in N(b) ∩ N(c)} = (tril(A) A)[b, c]
    A = sp.csr_matrix((np.ones(g.m, np.int64), g.edge_dst, g.node_off.astype(np.int64)), shape=(g.n, g.n))
    hits = np.asarray((sp.tril(A, -1).tocsr() @ A)[src[up], dst[up]]).ravel()
    work = np.zeros(g.n, np.int64)
    np.add.at(work, src[up], lens + 4 + 2 * hits)
    hs = off[:-1] + np.bincount(src[~up], minlength=g.n)  # lower slots come first in each list
    nup = off[1:] - hs
    nl = hs - off[:-1]
    work += (nup + 511) // 512 * (nl + 256)
    total = int(work.sum())
    with ctx.upload(g) as rg:
        for s in range(8):
            lo, hi = rg.set_shard(8, s)
            share = int(work[lo:hi].sum())
            assert share <= total / 8 + int(work.max()) + 1, (s, lo, hi, share, total)


def test_shard_arguments(ctx):
    g = G.gen_rmat(10, 8, seed=96)
    with ctx.upload(g) as rg:
        with pytest.raises(_ffi.AccError):
            rg.set_shard(0, 0)
        with pytest.raises(_ffi.AccError):
            rg.set_shard(4, 4)
        assert rg.set_shard(1, 0) == (0, g.n)


def test_scan_refuses_share_counts(ctx):
    """SCAN never clusters on one share's counts: while the forward path is
    restricted to a share, a query is refused; with the whole graph restored
    it recounts and equals the oracle's SCAN."""
    g = G.gen_rmat(12, 16, seed=97)
    with ctx.upload(g) as rg:
        rg.set_shard(4, 1)
        rg.count_all_edges("forward")
        with rg.scan_workspace() as ws:
            with pytest.raises(_ffi.AccError) as e:
                ws.run("0.3", 5)
            assert e.value.code == -1
        assert rg.set_shard(1, 0) == (0, g.n)
        with rg.scan_workspace() as ws:
            ws.run("0.3", 5)
            ws.synchronize()
            res = ws.results()
    want = oracle.scan(g, oracle.cn_merge(g), "0.3", 5)
    assert np.array_equal(res["label"], want["label"])
    assert np.array_equal(res["cluster_id"], want["cluster_id"])
