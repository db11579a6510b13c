"""Host side of multi-graph batches: packing a disjoint union keeps every
graph's counts, and the batch mix has the stated shape."""
import numpy as np

import oracle
from acctricnt_amd import graph as G


def test_pack_is_disjoint_union():
    graphs = G.batch_mix(num_graphs=12, seed=3, n_lo_log2=4, n_hi_log2=9, deg_lo=2, deg_hi=20)
    graphs.insert(2, G.CSRGraph(np.zeros(1, np.uint32), np.zeros(0, np.int32)))
    packed, vo = G.pack(graphs)
    G.validate(packed)
    so = packed.meta["slot_off"]
    assert vo[0] == 0 and vo[-1] == packed.n and so[-1] == packed.m
    assert np.array_equal(np.diff(vo), [g.n for g in graphs])
    cnt = oracle.cn_merge(packed)
    for k, g in enumerate(graphs):
        assert np.array_equal(packed.edge_dst[so[k]:so[k + 1]] - vo[k], g.edge_dst)
        assert np.array_equal(cnt[so[k]:so[k + 1]], oracle.cn_merge(g))
    assert oracle.triangle_count(packed) == sum(oracle.triangle_count(g) for g in graphs)


def test_batch_mix_shape():
    graphs = G.batch_mix(num_graphs=64, seed=11)
    kinds = {g.meta["kind"] for g in graphs}
    assert kinds == {"er", "rmat", "chung_lu"}
    for g in graphs:
        assert 2 ** 8 - 1 <= g.meta["n"] <= 2 ** 16 + 1 and g.n == g.meta["n"]
        assert 4 <= g.meta["avg_deg"] <= 64
        G.validate(g)
    again = G.batch_mix(num_graphs=64, seed=11)
    assert all(np.array_equal(a.edge_dst, b.edge_dst) for a, b in zip(graphs, again))


def test_wide_span_graph_shape():
    from test_gpu_parity import wide_span_graph

    for (nf, nn, nu), want_low in (((100, 900, 700), 1000), ((1000, 2000, 300), 3000)):
        g = wide_span_graph(150_000, nf, nn, nu, seed=1)
        G.validate(g)
        b = 150_000
        nb = g.edge_dst[g.node_off[b]:g.node_off[b + 1]]
        assert (nb < b).sum() == want_low and (nb > b).sum() == nu
        assert b - nb.min() > 2 ** 17
