"""Host CSR graphs: seeded generators, degree relabelling and the reference's
``b_degree.bin``/``b_adj.bin`` format, backed by ``libacc_graph.so``.

A :class:`CSRGraph` mirrors the CSR part of the reference's ``struct Graph``
(scan-query/util/graph.h:36-77): ``node_off`` uint32[n+1] and ``edge_dst``
int32[m], sorted and symmetric, without self-loops.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _ffi
from ._ffi import check_graph, graph_lib, i32p, u32p


@dataclass
class CSRGraph:
    node_off: np.ndarray  # uint32[n+1]
    edge_dst: np.ndarray  # int32[m]
    old_of_new: np.ndarray | None = None  # int32[n] after a degree relabel
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.node_off.shape[0] - 1)

    @property
    def m(self) -> int:
        return int(self.edge_dst.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.node_off.astype(np.int64))

    def sources(self) -> np.ndarray:
        """Source vertex of every slot."""
        return np.repeat(np.arange(self.n, dtype=np.int32), self.degrees())

    def reverse_slots(self) -> np.ndarray:
        rev = np.empty(self.m, dtype=np.uint32)
        h = _to_handle(self)
        try:
            check_graph(graph_lib().acc_hgraph_reverse_slots(h, _ffi.ptr(rev, C.c_uint32)))
        finally:
            graph_lib().acc_hgraph_free(h)
        return rev

    def algorithmic_bytes(self) -> int:
        """B_alg = 4(n+1) + 8m + 4 * sum over undirected edges of min(d_u, d_v)."""
        d = self.degrees()
        src = self.sources()
        up = src < self.edge_dst
        smin = np.minimum(d[src[up]], d[self.edge_dst[up]]).sum(dtype=np.int64)
        return int(4 * (self.n + 1) + 8 * self.m + 4 * int(smin))

    def forward_bytes(self, triangles: int, rev: np.ndarray | None = None) -> dict:
        """Algorithmic bytes of one forward counting launch (every triangle
        {a < b < c} found once at its middle vertex b by probing N(c) ∩ [0, b)
        against N<(b)), term by term:
          offsets      4 (n + 1)               read once;
          lower_lists  4 * m/2                 N<(b) of every source, read once;
          upper_slots  16 * m/2                per upper slot (b, c): c, its reverse
                                               slot and the start of N(c) read, the
                                               slot's count written;
          probes       4 * P                   P = sum over upper slots (b, c) of
                                               |N(c) ∩ [0, b)| = rev(b, c) - start of N(c);
          updates      4 * T + 4 * m/2         one 4-byte update of the lower slot
                                               (c, a) per triangle, one flush per
                                               lower slot (b, a).
        These are the bytes the kernel cannot avoid moving; HBM traffic above
        them is re-reading, below them is cache reuse."""
        if rev is None:
            rev = self.reverse_slots()
        src = self.sources()
        up = src < self.edge_dst
        off = self.node_off.astype(np.int64)
        probes = int((rev[up].astype(np.int64) - off[self.edge_dst[up]]).sum())
        half = self.m // 2
        terms = {"offsets": 4 * (self.n + 1), "lower_lists": 4 * half, "upper_slots": 16 * half,
                 "probes": 4 * probes, "updates": 4 * int(triangles) + 4 * half}
        return {"bytes": int(sum(terms.values())), "terms": terms, "probe_elements": probes}


def _from_handle(h, name="", meta=None) -> CSRGraph:
    lib = graph_lib()
    n = lib.acc_hgraph_n(h)
    m = lib.acc_hgraph_m(h)
    off = np.ctypeslib.as_array(lib.acc_hgraph_offsets(h), shape=(n + 1,)).copy()
    adj = (np.ctypeslib.as_array(lib.acc_hgraph_adj(h), shape=(m,)).copy() if m
           else np.zeros(0, dtype=np.int32))
    return CSRGraph(off.astype(np.uint32, copy=False), adj.astype(np.int32, copy=False),
                    name=name, meta=meta or {})


def _to_handle(g: CSRGraph):
    h = C.c_void_p()
    off = np.ascontiguousarray(g.node_off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.edge_dst, dtype=np.int32)
    check_graph(graph_lib().acc_hgraph_from_csr(g.n, g.m, _ffi.ptr(off, C.c_uint32),
                                                _ffi.ptr(adj, C.c_int32), C.byref(h)))
    return h


def _finish(h, relabel: bool, name: str, meta: dict) -> CSRGraph:
    lib = graph_lib()
    try:
        old = None
        if relabel:
            n = lib.acc_hgraph_n(h)
            old = np.empty(n, dtype=np.int32)
            check_graph(lib.acc_hgraph_relabel_degree(h, _ffi.ptr(old, C.c_int32)))
        g = _from_handle(h, name, meta)
        g.old_of_new = old
        return g
    finally:
        lib.acc_hgraph_free(h)


def from_edges(n: int, src, dst, relabel: bool = False, name: str = "edges") -> CSRGraph:
    """Symmetrise, deduplicate and drop self-loops of an undirected edge list."""
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    h = C.c_void_p()
    check_graph(graph_lib().acc_hgraph_from_edges(n, s.shape[0], _ffi.ptr(s, C.c_int32),
                                                  _ffi.ptr(d, C.c_int32), C.byref(h)))
    return _finish(h, relabel, name, {"kind": "edges"})


def relabel_degree(g: CSRGraph) -> CSRGraph:
    """A copy of g relabelled on the host by degree, descending, ties by
    ascending id (``acc_hgraph_relabel_degree``); ``old_of_new`` holds the
    original id of every new vertex."""
    return _finish(_to_handle(g), True, g.name, dict(g.meta))


def from_csr(node_off, edge_dst, name: str = "csr") -> CSRGraph:
    """Validate caller CSR arrays (sorted, symmetric, no self-loops) and wrap them."""
    g = CSRGraph(np.ascontiguousarray(node_off, dtype=np.uint32),
                 np.ascontiguousarray(edge_dst, dtype=np.int32), name=name)
    validate(g)
    return g


def validate(g: CSRGraph) -> None:
    check_graph(graph_lib().acc_csr_validate(g.n, g.m, _ffi.ptr(g.node_off, C.c_uint32),
                                             _ffi.ptr(g.edge_dst, C.c_int32)))


def verify_counts(g: CSRGraph, cnt, base: int = 2, nthreads: int = 0) -> tuple[int, int]:
    """CPU bit-vector recount of the u<v slots of a count array (the drop-in's
    Verify step): (number of mismatching u<v slots, first such slot or m)."""
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    if cnt.shape[0] != g.m:
        raise ValueError("count array length differs from m")
    bad, first = C.c_uint64(), C.c_uint64()
    check_graph(graph_lib().acc_csr_verify_counts(g.n, g.m, _ffi.ptr(g.node_off, C.c_uint32),
                                                  _ffi.ptr(g.edge_dst, C.c_int32), _ffi.ptr(cnt, C.c_int32),
                                                  int(base), int(nthreads), C.byref(bad), C.byref(first)))
    return bad.value, first.value


def gen_er(n: int, avg_deg: float, seed: int, relabel: bool = False) -> CSRGraph:
    """Erdos-Renyi: n*avg_deg/2 uniform pairs (u != v), then deduplicated."""
    samples = int(round(n * avg_deg / 2))
    h = C.c_void_p()
    check_graph(graph_lib().acc_gen_er(n, samples, seed, C.byref(h)))
    return _finish(h, relabel, f"er-n{n}-d{avg_deg:g}-s{seed}",
                   {"kind": "er", "n": n, "avg_deg": avg_deg, "seed": seed})


def gen_rmat(scale: int, edge_factor: float, seed: int, a=0.57, b=0.19, c=0.19,
             n: int | None = None, relabel: bool = True) -> CSRGraph:
    """R-MAT over 2^scale ids with edge_factor*n samples (Graph500 convention);
    with n < 2^scale, endpoints >= n are redrawn."""
    nv = (1 << scale) if n is None else int(n)
    samples = int(round(edge_factor * nv))
    h = C.c_void_p()
    check_graph(graph_lib().acc_gen_rmat(nv, scale, samples, a, b, c, seed, C.byref(h)))
    return _finish(h, relabel, f"rmat-s{scale}-ef{edge_factor:g}-s{seed}",
                   {"kind": "rmat", "scale": scale, "n": nv, "edge_factor": edge_factor,
                    "abc": (a, b, c), "seed": seed, "relabel": relabel})


def gen_chung_lu(n: int, gamma: float, avg_deg: float, max_deg: float, seed: int,
                 relabel: bool = False) -> CSRGraph:
    h = C.c_void_p()
    check_graph(graph_lib().acc_gen_chung_lu(n, gamma, avg_deg, max_deg, seed, C.byref(h)))
    return _finish(h, relabel, f"cl-n{n}-g{gamma:g}-d{avg_deg:g}-s{seed}",
                   {"kind": "chung_lu", "n": n, "gamma": gamma, "avg_deg": avg_deg,
                    "max_deg": max_deg, "seed": seed})


def read_bins(directory: str) -> CSRGraph:
    h = C.c_void_p()
    check_graph(graph_lib().acc_hgraph_read_bins(directory.encode(), C.byref(h)))
    try:
        return _from_handle(h, name=directory)
    finally:
        graph_lib().acc_hgraph_free(h)


def write_bins(g: CSRGraph, directory: str) -> None:
    h = _to_handle(g)
    try:
        check_graph(graph_lib().acc_hgraph_write_bins(h, directory.encode()))
    finally:
        graph_lib().acc_hgraph_free(h)


def convert_edge_list(txt_path: str, out_dir: str) -> None:
    check_graph(graph_lib().acc_convert_edge_list(txt_path.encode(), out_dir.encode()))


def batch_mix(num_graphs: int = 512, seed: int = 2024, n_lo_log2: float = 8.0,
              n_hi_log2: float = 16.0, deg_lo: float = 4.0, deg_hi: float = 64.0):
    """The mixed batch: vertex counts log-uniform in [2^n_lo_log2, 2^n_hi_log2],
    average degree uniform in [deg_lo, deg_hi], type ER / R-MAT(0.57,0.19,0.19)
    / Chung-Lu(2.1) with probability 1/3 each; graph i is seeded by (seed, i)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(num_graphs):
        n = int(round(2.0 ** rng.uniform(n_lo_log2, n_hi_log2)))
        deg = float(rng.uniform(deg_lo, deg_hi))
        kind = ("er", "rmat", "chung_lu")[int(rng.integers(0, 3))]
        specs.append((kind, n, deg, seed * 1000003 + i))
    graphs = []
    for kind, n, deg, s in specs:
        if kind == "er":
            g = gen_er(n, deg, s)
        elif kind == "rmat":
            scale = max(1, math.ceil(math.log2(n)))
            g = gen_rmat(scale, deg / 2.0, s, n=n, relabel=False)
        else:
            g = gen_chung_lu(n, 2.1, deg, float(n), s)
        g.meta.update({"kind": kind, "n": n, "avg_deg": deg, "seed": s})
        graphs.append(g)
    return graphs


def pack(graphs) -> tuple[CSRGraph, np.ndarray]:
    """Packed multi-graph CSR: the disjoint union of ``graphs``, graph k owning
    vertices ``[vertex_off[k], vertex_off[k+1])`` (its ids shifted by
    ``vertex_off[k]``).  Returns the packed graph and the int32 graph offsets
    (uint32[len(graphs)+1])."""
    ns = np.array([g.n for g in graphs], dtype=np.int64)
    ms = np.array([g.m for g in graphs], dtype=np.int64)
    vertex_off = np.zeros(len(graphs) + 1, dtype=np.int64)
    np.cumsum(ns, out=vertex_off[1:])
    slot_off = np.zeros(len(graphs) + 1, dtype=np.int64)
    np.cumsum(ms, out=slot_off[1:])
    if vertex_off[-1] >= 2**31 or slot_off[-1] >= 2**32:
        raise ValueError("packed graph exceeds int32 ids / uint32 offsets")
    node_off = np.empty(int(vertex_off[-1]) + 1, dtype=np.uint32)
    edge_dst = np.empty(int(slot_off[-1]), dtype=np.int32)
    for k, g in enumerate(graphs):
        v0, s0 = int(vertex_off[k]), int(slot_off[k])
        node_off[v0:v0 + g.n] = g.node_off[:-1].astype(np.int64) + s0
        edge_dst[s0:s0 + g.m] = g.edge_dst + v0
    node_off[-1] = slot_off[-1]
    packed = CSRGraph(node_off, edge_dst, name=f"pack{len(graphs)}",
                      meta={"graphs": len(graphs), "slot_off": slot_off})
    return packed, vertex_off.astype(np.uint32)
