"""Independent references used to pin the oracle: small closed-form graphs,
brute-force set intersection, scipy.sparse products, and a SCAN written
straight from its definitions (components by BFS rather than union-find,
similarity by exact rational arithmetic)."""
from __future__ import annotations

from collections import deque
from fractions import Fraction

import numpy as np
import scipy.sparse as sp

from acctricnt_amd import graph as G


# ------------------------------------------------------------------ graphs
def complete(n):
    s, d = np.triu_indices(n, 1)
    return G.from_edges(n, s, d, name=f"K{n}")


def complete_bipartite(a, b):
    s = np.repeat(np.arange(a), b)
    d = a + np.tile(np.arange(b), a)
    return G.from_edges(a + b, s, d, name=f"K{a},{b}")


def cycle(n):
    s = np.arange(n)
    return G.from_edges(n, s, (s + 1) % n, name=f"C{n}")


def star(k):
    return G.from_edges(k + 1, np.zeros(k, np.int32), np.arange(1, k + 1), name=f"S{k}")


def wheel(k):
    """Hub 0 joined to a rim cycle 1..k."""
    rim = np.arange(1, k + 1)
    s = np.concatenate([np.zeros(k, np.int64), rim])
    d = np.concatenate([rim, np.roll(rim, -1)])
    return G.from_edges(k + 1, s, d, name=f"W{k}")


def windmill(k):
    """k triangles sharing vertex 0."""
    s, d = [], []
    for t in range(k):
        a, b = 1 + 2 * t, 2 + 2 * t
        s += [0, 0, a]
        d += [a, b, b]
    return G.from_edges(2 * k + 1, s, d, name=f"F{k}")


def k4_pendant():
    """K4 on {0,1,2,3} plus the pendant edge (3,4)."""
    s, d = np.triu_indices(4, 1)
    return G.from_edges(5, np.append(s, 3), np.append(d, 4), name="K4+pendant")


def disjoint_cliques(sizes):
    s, d, base = [], [], 0
    for k in sizes:
        a, b = np.triu_indices(k, 1)
        s.append(a + base)
        d.append(b + base)
        base += k
    return G.from_edges(base, np.concatenate(s), np.concatenate(d), name=f"cliques{sizes}")


# ------------------------------------------------------------------ counts
def neighbor_sets(g):
    off, adj = g.node_off, g.edge_dst
    return [set(adj[off[u]:off[u + 1]].tolist()) for u in range(g.n)]


def cn_bruteforce(g):
    """2 + |N(u) ∩ N(v)| per slot by Python set intersection."""
    ns = neighbor_sets(g)
    out = np.empty(g.m, np.int32)
    off, adj = g.node_off, g.edge_dst
    for u in range(g.n):
        for j in range(off[u], off[u + 1]):
            out[j] = 2 + len(ns[u] & ns[adj[j]])
    return out


def adjacency(g):
    A = sp.csr_matrix((np.ones(g.m, np.int64), g.edge_dst.astype(np.int64), g.node_off.astype(np.int64)),
                      shape=(g.n, g.n))
    return A


def cn_scipy(g):
    """Per-slot ((A·A) ∘ A)[u, v] + 2, read in CSR slot order."""
    A = adjacency(g)
    P = (A @ A).multiply(A).tocsr()
    P.sort_indices()
    # P has exactly the sparsity of A when every product entry on an edge is > 0;
    # read it through a dense lookup per row to cover zero entries.
    out = np.empty(g.m, np.int64)
    off, adj = g.node_off, g.edge_dst
    for u in range(g.n):
        row = dict(zip(P.indices[P.indptr[u]:P.indptr[u + 1]].tolist(),
                       P.data[P.indptr[u]:P.indptr[u + 1]].tolist()))
        for j in range(off[u], off[u + 1]):
            out[j] = 2 + row.get(int(adj[j]), 0)
    return out


def triangles_scipy(g):
    A = adjacency(g).astype(np.float64)
    return int(round((A @ A).multiply(A).sum() / 6))


# ------------------------------------------------------------------ SCAN
def scan_bruteforce(g, cnt, eps, mu, old_id=None):
    """SCAN from the definitions.  Returns (label, cluster_id, counts, lines)."""
    eps = Fraction(eps) if isinstance(eps, str) else Fraction(str(eps))
    n, off, adj = g.n, g.node_off, g.edge_dst
    deg = np.diff(off.astype(np.int64))
    oid = np.arange(n) if old_id is None else np.asarray(old_id)
    sim = np.zeros(g.m, bool)
    for u in range(n):
        for j in range(off[u], off[u + 1]):
            v = adj[j]
            c = int(cnt[j])
            sim[j] = Fraction(c * c) >= eps * eps * int((deg[u] + 1) * (deg[v] + 1))
    core = np.array([sim[off[u]:off[u + 1]].sum() >= mu for u in range(n)])
    comp = -np.ones(n, np.int64)
    for s in range(n):
        if not core[s] or comp[s] >= 0:
            continue
        comp[s] = s
        q = deque([s])
        while q:
            x = q.popleft()
            for j in range(off[x], off[x + 1]):
                y = adj[j]
                if core[y] and sim[j] and comp[y] < 0:
                    comp[y] = s
                    q.append(y)
    cid = -np.ones(n, np.int64)
    for c in set(comp[core].tolist()):
        members = np.nonzero(comp == c)[0]
        cid[members] = oid[members].min()
    label = np.where(core, 1, 0)
    for a in range(n):
        if core[a]:
            continue
        comps = {comp[adj[j]] for j in range(off[a], off[a + 1]) if core[adj[j]]}
        if len(comps) >= 2:
            label[a] = 2
    counts = {"core_num": int(core.sum()), "cluster_num": len(set(comp[core].tolist())),
              "hub_num": int((label == 2).sum()), "outlier_num": int((label == 0).sum())}
    lines = ["c/n vertex_id cluster_id"]
    for u in sorted(np.nonzero(core)[0].tolist(), key=lambda x: oid[x]):
        lines.append(f"c {oid[u]} {cid[u]}")
    pairs = set()
    for u in np.nonzero(core)[0]:
        for j in range(off[u], off[u + 1]):
            v = adj[j]
            if not core[v] and sim[j]:
                pairs.add((int(cid[u]), int(oid[v])))
    for c, v in sorted(pairs):
        lines.append(f"n {v} {c}")
    return label, cid, counts, lines, sim


def sim_reference_fp(cnt, du1, dv1, eps):
    """The reference's decision: (double)((long double)cnt / sqrtl(du1*dv1)) >= eps."""
    c = np.longdouble(cnt)
    d = np.longdouble(du1) * np.longdouble(dv1)
    return np.float64(c / np.sqrt(d)) >= np.float64(eps)
