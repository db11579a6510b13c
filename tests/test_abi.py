"""The C-ABI libraries load and export every declared symbol; host-side
graph utilities behave as specified.  No GPU compute is issued here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bruteforce as BF
from acctricnt_amd import _ffi, graph as G
from acctricnt_amd._cn_sigs import SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(acc_\w+)\s*\(", text))


@pytest.fixture(scope="module", autouse=True)
def _built(built):
    return built


def test_cn_library_exports_every_declared_symbol():
    lib = C.CDLL(_ffi.CN_LIB_PATH)
    names = declared("acc_cn.h")
    assert names, "no declarations parsed"
    for name in sorted(names):
        assert hasattr(lib, name), name
    assert names == set(SIGNATURES), "ctypes signatures out of step with include/acc_cn.h"


def test_graph_library_exports_every_declared_symbol():
    lib = C.CDLL(_ffi.GRAPH_LIB_PATH)
    for name in sorted(declared("acc_graph.h")):
        assert hasattr(lib, name), name


def test_cn_library_declares_without_gpu():
    lib = _ffi.cn_lib()
    c = C.c_int(-1)
    assert lib.acc_device_count(C.byref(c)) == 0
    assert c.value >= 0


def test_context_error_without_device():
    lib = _ffi.cn_lib()
    c = C.c_int(0)
    lib.acc_device_count(C.byref(c))
    if c.value:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    rc = lib.acc_ctx_create(0, C.byref(h))
    assert rc == -5  # ACC_EDEVICE
    assert b"device" in lib.acc_last_error()


def test_generators_are_deterministic_and_valid():
    for mk in (lambda: G.gen_er(5000, 10, seed=1), lambda: G.gen_rmat(12, 16, seed=2),
               lambda: G.gen_chung_lu(6000, 2.1, 12, 800, seed=3), lambda: G.gen_rmat(11, 4, seed=4, n=1500)):
        a, b = mk(), mk()
        assert np.array_equal(a.node_off, b.node_off) and np.array_equal(a.edge_dst, b.edge_dst)
        G.validate(a)
        assert a.m % 2 == 0


def test_rmat_relabel_is_degree_descending_and_stable():
    raw = G.gen_rmat(12, 16, seed=5, relabel=False)
    g = G.gen_rmat(12, 16, seed=5, relabel=True)
    d = g.degrees()
    assert (np.diff(d) <= 0).all()
    old = g.old_of_new
    assert sorted(old.tolist()) == list(range(g.n))
    dr = raw.degrees()
    assert np.array_equal(dr[old], d)
    # ties keep old-id order
    for k in np.unique(d):
        ids = old[d == k]
        assert (np.diff(ids) > 0).all()
    # same edge set under the mapping
    s_new = old[g.sources()]
    t_new = old[g.edge_dst]
    e1 = set(zip(s_new.tolist(), t_new.tolist()))
    e2 = set(zip(raw.sources().tolist(), raw.edge_dst.tolist()))
    assert e1 == e2


def test_bins_roundtrip(tmp_path):
    g = G.gen_rmat(10, 8, seed=9)
    G.write_bins(g, str(tmp_path))
    raw = open(tmp_path / "b_degree.bin", "rb").read()
    hdr = np.frombuffer(raw[:12], dtype=np.int32)
    assert hdr[0] == 4 and hdr[1] == g.n and hdr[2] == g.m
    h = G.read_bins(str(tmp_path))
    assert np.array_equal(h.node_off, g.node_off) and np.array_equal(h.edge_dst, g.edge_dst)


def test_edge_list_converter(tmp_path):
    txt = tmp_path / "e.txt"
    # first < second lines kept, others (incl. self-loops and reversed) dropped, duplicates merged
    txt.write_text("# comment\n0 1\n1 2\n2 0\n3 3\n1 2\n0 3\n")
    G.convert_edge_list(str(txt), str(tmp_path))
    g = G.read_bins(str(tmp_path))
    assert g.n == 4
    edges = set(zip(g.sources().tolist(), g.edge_dst.tolist()))
    assert edges == {(0, 1), (1, 0), (1, 2), (2, 1), (0, 3), (3, 0)}


@pytest.mark.parametrize("case,msg", [
    ("selfloop", b"Self loop"), ("unsorted", b"not sorted"), ("asym", b"not symmetric"), ("range", b"range")])
def test_validation_errors(case, msg):
    off = np.array([0, 2, 4, 6], np.uint32)
    adj = np.array([1, 2, 0, 2, 0, 1], np.int32)  # triangle
    if case == "selfloop":
        adj = np.array([0, 1, 0, 2, 1, 2], np.int32)
    elif case == "unsorted":
        adj = np.array([2, 1, 0, 2, 0, 1], np.int32)
    elif case == "asym":
        off = np.array([0, 2, 3, 5], np.uint32)
        adj = np.array([1, 2, 0, 0, 1], np.int32)
    elif case == "range":
        adj = np.array([1, 5, 0, 2, 0, 1], np.int32)
    with pytest.raises(_ffi.AccError) as e:
        G.from_csr(off, adj)
    assert e.value.code == -2
    assert msg in str(e.value).encode()


def test_algorithmic_bytes_formula():
    g = BF.complete(5)
    # n=5, m=20, every undirected edge has min degree 4, 10 edges
    assert g.algorithmic_bytes() == 4 * 6 + 8 * 20 + 4 * 10 * 4


def test_result_writer_matches_oracle(tmp_path):
    """The library's result-file writer reproduces the oracle's file byte for byte,
    with and without an old-id mapping."""
    import oracle
    from acctricnt_amd import cn

    for relabel in (False, True):
        g = G.gen_rmat(10, 12, seed=31, relabel=relabel)
        old = g.old_of_new if relabel else None
        res = oracle.scan(g, oracle.cn_merge(g), "0.4", 3, old_id=old)
        a = tmp_path / f"a{relabel}.txt"
        oracle.write_result(str(a), g, res, old_id=old)
        src = g.sources()
        core = res["label"] == 1
        sel = res["sim"].astype(bool) & core[src] & ~core[g.edge_dst]
        ids = g.edge_dst[sel] if old is None else old[g.edge_dst[sel]]
        pairs = np.stack([res["cluster_id"][src[sel]], ids], axis=1)
        pairs = np.concatenate([pairs, pairs[:3]])  # duplicates are removed by the writer
        b = tmp_path / f"b{relabel}.txt"
        cn.write_scan_result(str(b), g.n, res["label"], res["cluster_id"], pairs, old_id=old)
        assert a.read_bytes() == b.read_bytes()


def test_result_path_format(tmp_path):
    from acctricnt_amd import cn

    assert cn.scan_result_path("/d", 0.2, 5) == "/d/scanxp-result-0.2-5.txt"
    assert cn.scan_result_path("/d", 1.0, 2) == "/d/scanxp-result-1.-2.txt"


@pytest.mark.parametrize("maker", [lambda: G.gen_rmat(11, 16, seed=21), lambda: BF.wheel(9),
                                   lambda: G.gen_chung_lu(3000, 2.1, 12, 600, seed=22)])
def test_verify_counts_recount(maker):
    """The drop-in's Verify recount (acc_csr_verify_counts) accepts scipy's
    per-slot counts, and reports exactly the u<v slots where a count was
    changed (a change in a lower slot is outside the check, as in the
    reference's Verify, which recounts u<v slots only)."""
    g = maker()
    ref = BF.cn_scipy(g).astype(np.int32)
    assert G.verify_counts(g, ref) == (0, g.m)
    assert G.verify_counts(g, ref - 2, base=0) == (0, g.m)
    src = g.sources()
    up = np.nonzero(src < g.edge_dst)[0]
    low = np.nonzero(src > g.edge_dst)[0]
    bad = ref.copy()
    pick = np.sort(np.random.default_rng(3).choice(up, min(5, up.size), replace=False))
    bad[pick] += 1
    bad[low[0]] -= 1
    assert G.verify_counts(g, bad, nthreads=3) == (pick.size, int(pick[0]))
