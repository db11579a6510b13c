"""SCAN serving over resident graphs (``acc_server_*`` in include/acc_cn.h).

A :class:`Server` keeps graphs resident in HBM with their common-neighbor
counts computed once at admission, admits and evicts them by device footprint
against a budget, and answers SCAN queries ``(graph, eps, mu)`` submitted from
any thread.  A native worker thread batches pending queries across graphs into
one set of SCAN launches per batch and keeps several batches in flight
(continuous batching).  Results are the labels and cluster ids that
:meth:`acctricnt_amd.cn.ScanWorkspace.run` gives for the same graph and
parameters alone (the reference runs one such query per process:
scan-query/scan_xp.cpp:2083-2110).
"""
from __future__ import annotations

import ctypes as C
from fractions import Fraction

import numpy as np

from . import _ffi
from ._cn_sigs import ServerConfig, ServerStats
from ._ffi import check_cn
from .cn import Context, UPLOAD_NO_VALIDATE

QUERY_PENDING, QUERY_RUNNING, QUERY_DONE, QUERY_FAILED = 0, 1, 2, 3


def eps_fraction(eps) -> Fraction:
    """Exact eps: a decimal string or float is read as the decimal it prints as."""
    if isinstance(eps, Fraction):
        return eps
    if isinstance(eps, tuple):
        return Fraction(int(eps[0]), int(eps[1]))
    return Fraction(str(eps))


class Server:
    def __init__(self, ctx: Context, hbm_budget: int | None = None, max_batch: int | None = None,
                 streams: int | None = None, slots: int | None = None, index: bool = True):
        self.ctx = ctx
        self._lib = ctx._lib
        cfg = ServerConfig()
        self._lib.acc_server_config_default(C.byref(cfg))
        if hbm_budget is not None:
            cfg.hbm_budget = int(hbm_budget)
        if max_batch is not None:
            cfg.max_batch = int(max_batch)
        if streams is not None:
            cfg.streams = int(streams)
        if slots is not None:
            cfg.slots = int(slots)
        cfg.no_index = 0 if index else 1
        self.config = {k: getattr(cfg, k) for k in ("hbm_budget", "max_batch", "streams", "slots", "no_index")}
        h = C.c_void_p()
        check_cn(self._lib.acc_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self._h = h
        self.graph_n = {}
        ctx._servers.add(self)  # closed before the context

    def add_graph(self, graph, validate: bool = True) -> int:
        off = np.ascontiguousarray(graph.node_off, dtype=np.uint32)
        adj = np.ascontiguousarray(graph.edge_dst, dtype=np.int32)
        gid = C.c_int32()
        check_cn(self._lib.acc_server_add_graph(self._h, graph.n, graph.m, _ffi.ptr(off, C.c_uint32),
                                                _ffi.ptr(adj, C.c_int32), 0 if validate else UPLOAD_NO_VALIDATE,
                                                C.byref(gid)))
        self.graph_n[gid.value] = graph.n
        return gid.value

    def remove_graph(self, gid: int) -> None:
        check_cn(self._lib.acc_server_remove_graph(self._h, int(gid)))
        self.graph_n.pop(gid, None)

    def graph_info(self, gid: int) -> dict:
        n, m, res, b = C.c_uint32(), C.c_uint64(), C.c_int(), C.c_uint64()
        check_cn(self._lib.acc_server_graph_info(self._h, int(gid), C.byref(n), C.byref(m), C.byref(res),
                                                 C.byref(b)))
        return {"n": n.value, "m": m.value, "resident": bool(res.value), "device_bytes": b.value}

    def submit(self, gid: int, eps, mu: int) -> int:
        f = eps_fraction(eps)
        t = C.c_uint64()
        check_cn(self._lib.acc_server_submit(self._h, int(gid), f.numerator, f.denominator, int(mu), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> None:
        check_cn(self._lib.acc_server_wait(self._h, ticket, int(timeout_ms)))

    def poll(self, ticket: int) -> int:
        st = C.c_int()
        check_cn(self._lib.acc_server_poll(self._h, ticket, C.byref(st)))
        return st.value

    def result(self, ticket: int, copy: bool = True) -> dict:
        """labels (int8), cluster ids (numpy views into pinned memory unless
        copy), the four counters and the latency of a completed query."""
        lab, cid = C.POINTER(C.c_int8)(), C.POINTER(C.c_int32)()
        n = C.c_uint32()
        counts = (C.c_int64 * 4)()
        lat = C.c_double()
        check_cn(self._lib.acc_server_result(self._h, ticket, C.byref(lab), C.byref(cid), C.byref(n), counts,
                                             C.byref(lat)))
        k = n.value
        label = np.ctypeslib.as_array(lab, shape=(k,)) if k else np.zeros(0, np.int8)
        cluster = np.ctypeslib.as_array(cid, shape=(k,)) if k else np.zeros(0, np.int32)
        if copy:
            label, cluster = label.copy(), cluster.copy()
        return {"label": label, "cluster_id": cluster, "core_num": counts[0], "cluster_num": counts[1],
                "hub_num": counts[2], "outlier_num": counts[3], "latency_ms": lat.value}

    def release(self, ticket: int) -> None:
        check_cn(self._lib.acc_server_release(self._h, ticket))

    def query(self, gid: int, eps, mu: int) -> dict:
        """Submit, wait, read and release one query."""
        t = self.submit(gid, eps, mu)
        try:
            self.wait(t)
            return self.result(t)
        finally:
            self.release(t)

    def stats(self) -> dict:
        st = ServerStats()
        check_cn(self._lib.acc_server_get_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.acc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
