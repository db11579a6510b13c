"""The allocation-free triangle total (acc_triangle_total_device), through
the C ABI: on device-resident CSR arrays it equals acc_triangle_total of the
same graph uploaded as usual, and the oracle's triangle count, on two calls
in a row (the second on a smaller graph, so stale partial sums from the
first would show), and it reports a malformed CSR as ~0 without touching
memory outside the arrays."""
import ctypes as C

import numpy as np
import pytest

import bruteforce as BF
import oracle
from acctricnt_amd import _ffi, cn, graph as G

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def ctx(built):
    assert cn.device_count() > 0, "no HIP device visible"
    c = cn.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


class DeviceCSR:
    """The CSR and a result word in device memory, allocated by the test."""

    def __init__(self, hip, off, adj):
        self.hip = hip
        self.n, self.m = off.shape[0] - 1, adj.shape[0]
        self.ptrs = [C.c_void_p() for _ in range(3)]
        for p, size in zip(self.ptrs, (4 * (self.n + 1), 4 * max(self.m, 1), 8)):
            assert hip.hipMalloc(C.byref(p), size) == 0
        assert hip.hipMemcpy(self.ptrs[0], off.ctypes.data, 4 * (self.n + 1), H2D) == 0
        if self.m:
            assert hip.hipMemcpy(self.ptrs[1], adj.ctypes.data, 4 * self.m, H2D) == 0

    def total(self, ctx):
        lib = ctx._lib
        rc = lib.acc_triangle_total_device(ctx.handle, self.n, self.m, self.ptrs[0], self.ptrs[1], self.ptrs[2])
        assert rc == 0, lib.acc_last_error()
        ctx.synchronize()
        out = np.zeros(1, np.uint64)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptrs[2], 8, D2H) == 0
        return int(out[0])

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def device_total(ctx, hip, g):
    d = DeviceCSR(hip, np.ascontiguousarray(g.node_off, np.uint32), np.ascontiguousarray(g.edge_dst, np.int32))
    try:
        return d.total(ctx)
    finally:
        d.free()


def test_two_calls_larger_then_smaller(ctx, hip):
    big, small = G.gen_rmat(16, 16, seed=81), G.gen_rmat(11, 16, seed=82)
    for g in (big, small):
        print(f"[no-alloc] {g.name} n={g.n} m={g.m}", flush=True)
        with ctx.upload(g) as rg:
            want = rg.triangle_total("forward")
        assert want == oracle.triangle_count(g)
        assert device_total(ctx, hip, g) == want


@pytest.mark.parametrize("maker", [
    lambda: G.gen_chung_lu(1 << 14, 2.1, 24, 4096, seed=83),
    lambda: G.gen_rmat(13, 16, seed=84, relabel=False),
    lambda: G.gen_er(20000, 12, seed=85),
    lambda: BF.complete(1100),
    lambda: BF.wheel(9),
    lambda: G.from_csr(np.array([0, 0], np.uint32), np.zeros(0, np.int32)),
], ids=["chunglu14", "rmat13_raw", "er20000", "K1100", "wheel9", "single_vertex"])
def test_matches_triangle_total(ctx, hip, maker):
    """Staged (|N<(b)| <= 1024) and global-memory membership (K_1100 has
    sources with up to 1099 lower neighbors), raw ids, and empty graphs."""
    g = maker()
    print(f"[no-alloc] {g.name} n={g.n} m={g.m}", flush=True)
    with ctx.upload(g) as rg:
        want = rg.triangle_total("forward")
    assert want == oracle.triangle_count(g)
    assert device_total(ctx, hip, g) == want


def test_malformed_reported(ctx, hip):
    """An id outside [0, n) and an offset beyond m give ~0; a good graph
    afterwards is counted again (the error word is cleared per call)."""
    k4 = BF.complete(4)
    off = np.ascontiguousarray(k4.node_off, np.uint32)
    adj = np.ascontiguousarray(k4.edge_dst, np.int32).copy()
    adj[-1] = 4  # vertex 3's last neighbor: out of range
    d = DeviceCSR(hip, off, adj)
    try:
        assert d.total(ctx) == (1 << 64) - 1
    finally:
        d.free()
    off2 = off.copy()
    off2[2] = 40  # beyond m
    d = DeviceCSR(hip, off2, np.ascontiguousarray(k4.edge_dst, np.int32))
    try:
        assert d.total(ctx) == (1 << 64) - 1
    finally:
        d.free()
    assert device_total(ctx, hip, k4) == 4
    d = DeviceCSR(hip, np.zeros(1, np.uint32), np.zeros(0, np.int32))  # no vertices
    try:
        assert d.total(ctx) == 0
    finally:
        d.free()


def test_arguments_rejected(ctx):
    lib = ctx._lib
    assert lib.acc_triangle_total_device(ctx.handle, 3, 0, None, None, None) == -1
    assert lib.acc_triangle_total_device(None, 3, 0, None, None, None) == -1
