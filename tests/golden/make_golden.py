"""Writes tests/golden/golden.json: known answers for the oracle and the HIP path.

The expected values come from the independent references in tests/bruteforce.py
(scipy.sparse per-edge products, trace(A^3)/6, Python set intersection and a
SCAN written from its definitions with exact rational similarity), never from
the oracle under test.  Each entry records the generator parameters and the
SHA-256 of the generated CSR, so a change of the generators is detected.

The reference build is not available in this container (its set-intersection
and SCAN sources include the un-vendored sparsepp header), so no fixture here
comes from running the reference; the one known answer taken from a reference
run is the K4+pendant case below (Triangle Cnt 4; eps 0.6, mu 2: Core 4,
Cluster 1, Hub 0, Outlier 1, result "c 0..3 -> 0" and "n 4 0").

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from acctricnt_amd import graph as G  # noqa: E402
import bruteforce as BF  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def csr_sha(g) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.node_off, dtype=np.uint32).tobytes())
    h.update(np.ascontiguousarray(g.edge_dst, dtype=np.int32).tobytes())
    return h.hexdigest()


CASES = [
    # name, builder, generator spec (for regeneration), SCAN queries
    ("er16384", lambda: G.gen_er(16384, 16, seed=20190101),
     {"kind": "er", "n": 16384, "avg_deg": 16, "seed": 20190101}, [("0.2", 5), ("0.5", 2), ("0.1", 10)]),
    ("rmat12", lambda: G.gen_rmat(12, 16, seed=7),
     {"kind": "rmat", "scale": 12, "edge_factor": 16, "seed": 7, "relabel": True},
     [("0.3", 5), ("0.6", 5), ("0.9", 2)]),
    ("chunglu8192", lambda: G.gen_chung_lu(8192, 2.1, 16, 2048, seed=11),
     {"kind": "chung_lu", "n": 8192, "gamma": 2.1, "avg_deg": 16, "max_deg": 2048, "seed": 11},
     [("0.3", 5), ("0.6", 5)]),
    ("k4_pendant", BF.k4_pendant, {"kind": "k4_pendant"}, [("0.6", 2)]),
    ("wheel7", lambda: BF.wheel(7), {"kind": "wheel", "k": 7}, [("0.5", 3)]),
    ("windmill5", lambda: BF.windmill(5), {"kind": "windmill", "k": 5}, [("0.7", 2)]),
    ("k3_4", lambda: BF.complete_bipartite(3, 4), {"kind": "complete_bipartite", "a": 3, "b": 4}, [("0.4", 2)]),
    ("cliques_4_5_6", lambda: BF.disjoint_cliques([4, 5, 6]), {"kind": "cliques", "sizes": [4, 5, 6]},
     [("0.8", 3)]),
]


def main():
    out = {}
    for name, build, spec, queries in CASES:
        g = build()
        cnt = BF.cn_scipy(g).astype(np.int32)
        if g.m <= 300_000:
            assert np.array_equal(cnt, BF.cn_bruteforce(g)), name
        tri = BF.triangles_scipy(g)
        entry = {"spec": spec, "n": g.n, "m": g.m, "csr_sha256": csr_sha(g), "cnt_sha256": sha(cnt),
                 "cnt_sum": int(cnt.astype(np.int64).sum()), "triangles": tri, "scan": []}
        if g.m <= 64:
            entry["cnt"] = cnt.tolist()
            entry["node_off"] = g.node_off.tolist()
            entry["edge_dst"] = g.edge_dst.tolist()
        for eps, mu in queries:
            label, cid, counts, lines, sim = BF.scan_bruteforce(g, cnt, eps, mu)
            text = "\n".join(lines) + "\n"
            q = {"eps": eps, "mu": mu, **counts, "result_sha256": hashlib.sha256(text.encode()).hexdigest(),
                 "label_sha256": sha(label.astype(np.int32)), "sim_sha256": sha(sim.astype(np.uint8))}
            if len(lines) <= 40:
                q["result_lines"] = lines
            entry["scan"].append(q)
        out[name] = entry
        print(name, g.n, g.m, tri, flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
