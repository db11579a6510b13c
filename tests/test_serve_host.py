"""The serving loop's failure handling on the CPU: the policy code behind
acc_server_* (acctricnt_amd/csrc/serve_core.h) driven over a fake device
(tests/serve/serve_driver.cpp, built here with g++): admission against a
small budget and a fake allocator, LRU eviction with queries queued,
re-admission, injected allocation, launch and completion failures,
cancellation, slot exhaustion, result-slot reuse, and timed waits driven by
an injected clock.  No GPU is used."""
import ctypes as C
import os
import subprocess
import threading
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "acctricnt_amd", "csrc")

OK, EINVAL, ENOMEM, EDEVICE, EBUDGET, ENOTFOUND, ETIMEDOUT = 0, -1, -3, -5, -6, -7, -8
PENDING, RUNNING, DONE, FAILED = 0, 1, 2, 3


class Stats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in (
        "submitted", "completed", "failed", "cancelled", "batches", "batched_queries", "max_batch_seen",
        "admissions", "evictions", "graph_bytes", "slot_bytes", "resident_graphs", "pending", "index_queries")]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("serve") / "libserve_fake.so"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-shared", "-fPIC", "-pthread", "-Wall", "-I" + CSRC,
                    "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "serve", "serve_driver.cpp"),
                    "-o", str(out)], check=True, capture_output=True, text=True)
    L = C.CDLL(str(out))
    vp, u64, i32, i64 = C.c_void_p, C.c_uint64, C.c_int32, C.c_int64
    L.fake_create.restype = vp
    L.fake_create.argtypes = [u64, u64, C.c_int, C.c_int, C.c_int, C.c_int]
    L.fake_shutdown.argtypes = [vp]
    L.fake_destroy.argtypes = [vp]
    L.fake_add_graph.argtypes = [vp, C.c_uint32, u64, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(i32)]
    L.fake_remove_graph.argtypes = [vp, i32]
    L.fake_graph_info.argtypes = [vp, i32, C.POINTER(C.c_int), C.POINTER(u64)]
    L.fake_submit.argtypes = [vp, i32, i64, i64, i32, C.POINTER(u64)]
    L.fake_wait.argtypes = [vp, u64, i32]
    L.fake_poll.argtypes = [vp, u64, C.POINTER(C.c_int)]
    L.fake_result.argtypes = [vp, u64, C.POINTER(i64), C.POINTER(C.c_int8), C.POINTER(C.c_int32), C.c_uint32,
                              C.POINTER(C.c_double)]
    L.fake_release.argtypes = [vp, u64]
    L.fake_stats.argtypes = [vp, C.POINTER(Stats)]
    L.fake_complete.argtypes = [vp, C.c_int, C.c_int]
    L.fake_inflight.argtypes = [vp, C.c_int]
    L.fake_fail.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.fake_advance.argtypes = [vp, i64]
    L.fake_allocated.restype = u64
    L.fake_allocated.argtypes = [vp]
    L.fake_order_violations.argtypes = [vp]
    L.fake_uploads.argtypes = [vp]
    L.fake_releases.argtypes = [vp]
    L.fake_last_error.restype = C.c_char_p
    return L


class Server:
    """One core over a fake device; estimate(n, m) = 1000 + 8 m + 4 n bytes,
    actual footprint 100 below it; a slot for (n, m) holds 100 + n + 4 m."""

    def __init__(self, L, budget, phys=1 << 40, streams=1, slots=8, max_batch=4, gated=True):
        self.L = L
        self.h = L.fake_create(budget, phys, streams, slots, max_batch, 1 if gated else 0)

    def close(self):
        self.L.fake_destroy(self.h)

    def add(self, n=10, m=0):
        off = np.zeros(n + 1, np.uint32)
        off[-1] = m
        adj = np.zeros(max(m, 1), np.int32)
        gid = C.c_int32(-1)
        rc = self.L.fake_add_graph(self.h, n, m, off.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   adj.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(gid))
        return rc, gid.value

    def resident(self, gid):
        r, b = C.c_int(), C.c_uint64()
        assert self.L.fake_graph_info(self.h, gid, C.byref(r), C.byref(b)) == OK
        return bool(r.value), b.value

    def submit(self, gid, p=1, q=2, mu=3):
        t = C.c_uint64()
        rc = self.L.fake_submit(self.h, gid, p, q, mu, C.byref(t))
        return rc, t.value

    def wait(self, t, timeout_ms=-1):
        return self.L.fake_wait(self.h, t, timeout_ms)

    def state(self, t):
        s = C.c_int(-1)
        rc = self.L.fake_poll(self.h, t, C.byref(s))
        return s.value if rc == OK else rc

    def result(self, t, cap=64):
        counts = (C.c_int64 * 4)()
        lab = np.zeros(cap, np.int8)
        cid = np.zeros(cap, np.int32)
        lat = C.c_double()
        rc = self.L.fake_result(self.h, t, counts, lab.ctypes.data_as(C.POINTER(C.c_int8)),
                                cid.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(lat))
        return rc, list(counts), lab, cid, lat.value

    def release(self, t):
        return self.L.fake_release(self.h, t)

    def stats(self):
        s = Stats()
        self.L.fake_stats(self.h, C.byref(s))
        return s

    def complete(self, stream=0, k=1):
        self.L.fake_complete(self.h, stream, k)

    def inflight(self, stream=0):
        return self.L.fake_inflight(self.h, stream)

    def until(self, cond, what, limit=5.0):
        """Waits (real time) for the core's threads to reach a state."""
        t0 = time.time()
        while not cond():
            assert time.time() - t0 < limit, f"timed out waiting for {what}"
            time.sleep(0.002)


def est(n=10, m=0):
    return 1000 + 8 * m + 4 * n


SLOT = 110  # result slot bytes for a graph of 10 vertices and no edges


@pytest.fixture
def mk(lib):
    made = []

    def make(*a, **k):
        s = Server(lib, *a, **k)
        made.append(s)
        return s
    yield make
    for s in made:
        s.close()


def test_idle_slots_grow_without_evicting(mk):
    """Registering a graph grows the idle result slots only within the
    budget's free room: it never evicts a graph (the new one included) for
    them; the slots that did not fit grow when a query needs them."""
    s = mk(budget=est() + 50, slots=4, gated=False)  # room for one slot beside the graph's 940 bytes
    rc, a = s.add()
    assert rc == OK and s.resident(a)[0]
    st = s.stats()
    assert (st.evictions, st.slot_bytes) == (0, SLOT)
    t = s.submit(a)[1]
    assert s.wait(t) == OK and s.release(t) == OK


def test_admission_rejected_over_budget(mk, lib):
    s = mk(budget=est(10, 100) - 1)
    rc, _ = s.add(10, 100)
    assert rc == EBUDGET and b"budget" in lib.fake_last_error()
    st = s.stats()
    assert (st.admissions, st.graph_bytes, st.resident_graphs) == (0, 0, 0)
    assert lib.fake_allocated(s.h) == 0 and lib.fake_uploads(s.h) == 0


def test_lru_eviction_and_readmission(mk, lib):
    """A budget of two graphs: the third admission evicts the least recently
    used, a query on the evicted graph admits it again (evicting the next
    least recently used idle graph), and its result comes from its graph."""
    s = mk(budget=2 * est() + 50, slots=1, gated=False)
    a = s.add()[1]
    b = s.add()[1]
    rc, c = s.add()
    assert rc == OK
    assert [s.resident(g)[0] for g in (a, b, c)] == [False, True, True]
    assert s.stats().evictions == 1
    rc, t = s.submit(a, p=3, q=4, mu=2)
    assert rc == OK and s.wait(t, -1) == OK
    assert [s.resident(g)[0] for g in (a, b, c)] == [True, False, True]
    rc, counts, lab, cid, lat = s.result(t)
    assert rc == OK and counts == [10, 3, 4, 2] and lat >= 0
    st = s.stats()
    assert (st.admissions, st.evictions, st.completed) == (4, 2, 1)
    assert st.graph_bytes == sum(s.resident(g)[1] for g in (a, b, c))
    assert st.graph_bytes == 2 * (est() - 100)  # the actual footprints, not the estimates
    assert s.release(t) == OK


def test_eviction_skips_graphs_with_queries_in_flight(mk, lib):
    """With one graph's query running and the other's queued, admitting a
    third graph finds nothing to evict; once both finish it evicts the least
    recently used."""
    s = mk(budget=2 * est() + 4 * SLOT + 50, slots=4, max_batch=1)
    a = s.add()[1]
    b = s.add()[1]
    t1 = s.submit(a)[1]
    s.until(lambda: s.inflight() == 1, "the first batch")
    t2 = s.submit(b)[1]
    assert s.state(t1) == RUNNING and s.state(t2) == PENDING
    rc, _ = s.add()
    assert rc == EBUDGET and b"no idle graph" in lib.fake_last_error()
    assert s.resident(a)[0] and s.resident(b)[0] and s.stats().evictions == 0
    s.complete(k=2)
    assert s.wait(t1) == OK and s.wait(t2) == OK
    rc, c = s.add()
    assert rc == OK and not s.resident(a)[0] and s.resident(b)[0] and s.resident(c)[0]
    for t in (t1, t2):
        assert s.release(t) == OK


def test_admission_waits_for_running_queries(mk):
    """A budget of one graph: a queued query whose graph cannot be admitted
    while another graph's query runs stays pending (no failure) and runs
    once the running query completes and its graph can be evicted."""
    s = mk(budget=est() + 2 * SLOT + 50, slots=2)
    a = s.add()[1]
    b = s.add()[1]  # evicts a
    t1 = s.submit(a, mu=5)[1]  # admits a again, evicting b
    s.until(lambda: s.inflight() == 1, "the first batch")
    t2 = s.submit(b, mu=6)[1]
    time.sleep(0.1)
    assert s.state(t1) == RUNNING and s.state(t2) == PENDING
    s.complete()
    assert s.wait(t1) == OK
    s.until(lambda: s.state(t2) == RUNNING, "the second batch")
    s.complete()
    assert s.wait(t2) == OK
    assert s.result(t2)[1][3] == 6 and s.stats().failed == 0


def test_timed_wait_follows_the_injected_clock(mk, lib):
    s = mk(budget=1 << 30)
    g = s.add()[1]
    t = s.submit(g)[1]
    out = {}
    th = threading.Thread(target=lambda: out.setdefault("rc", s.wait(t, 1000)))
    th.start()
    lib.fake_advance(s.h, 999)
    time.sleep(0.2)
    assert th.is_alive(), "the wait ended before 1000 ms of the device clock passed"
    lib.fake_advance(s.h, 2)
    th.join(5)
    assert not th.is_alive() and out["rc"] == ETIMEDOUT
    # the query itself goes on and completes
    s.complete()
    assert s.wait(t, 1000) == OK
    # a wait that completes before its deadline
    t2 = s.submit(g)[1]
    th = threading.Thread(target=lambda: out.__setitem__("rc2", s.wait(t2, 1000)))
    th.start()
    lib.fake_advance(s.h, 500)
    s.until(lambda: s.inflight() == 1, "the second batch")
    s.complete()
    th.join(5)
    assert out["rc2"] == OK
    # latency comes from the same clock: 500 ms passed between submission and completion
    assert abs(s.result(t2)[4] - 500.0) < 1e-6
    assert s.release(t) == OK and s.release(t2) == OK


def test_cancel_pending_and_refuse_running(mk):
    s = mk(budget=1 << 30, max_batch=1)
    g = s.add()[1]
    t1 = s.submit(g)[1]
    s.until(lambda: s.inflight() == 1, "the first batch")
    t2 = s.submit(g)[1]
    assert s.release(t1) == EINVAL  # running
    assert s.release(t2) == OK  # pending: cancelled
    assert s.state(t2) == ENOTFOUND and s.wait(t2, 0) == ENOTFOUND
    st = s.stats()
    assert (st.cancelled, st.pending) == (1, 0)
    s.complete()
    assert s.wait(t1) == OK and s.release(t1) == OK
    assert s.stats().completed == 1


def test_upload_allocation_failure(mk, lib):
    """The fake allocator refuses an upload: the registration fails with
    ACC_ENOMEM and leaves no bytes reserved; the next one succeeds."""
    s = mk(budget=1 << 30)
    lib.fake_fail(s.h, 1, 0, 0)
    rc, _ = s.add()
    assert rc == ENOMEM
    assert s.stats().graph_bytes == 0 and lib.fake_allocated(s.h) == 0
    rc, g = s.add()
    assert rc == OK and s.resident(g) == (True, est() - 100)


def test_readmission_failure_fails_the_query(mk, lib):
    s = mk(budget=est() + SLOT + 50, slots=1, gated=False)
    a = s.add()[1]
    assert s.add()[0] == OK  # evicts a
    lib.fake_fail(s.h, 1, 0, 0)
    t = s.submit(a)[1]
    assert s.wait(t, -1) == ENOMEM
    assert s.state(t) == FAILED and s.stats().failed == 1
    assert s.release(t) == OK


def test_physical_memory_exhaustion(mk, lib):
    """The budget allows a graph the allocator cannot hold: ACC_ENOMEM from
    the device, not a budget error, and the accounting stays at zero."""
    s = mk(budget=1 << 30, phys=est() - 101)
    rc, _ = s.add()
    assert rc == ENOMEM and s.stats().graph_bytes == 0


def test_device_errors_fail_the_batch(mk, lib):
    s = mk(budget=1 << 30, max_batch=2, gated=False)
    g = s.add()[1]
    lib.fake_fail(s.h, 0, 1, 0)
    t = s.submit(g)[1]
    assert s.wait(t) == EDEVICE and b"launch" in lib.fake_last_error()
    assert s.release(t) == OK
    lib.fake_fail(s.h, 0, 0, 1)
    t = s.submit(g)[1]
    assert s.wait(t) == EDEVICE and b"batch" in lib.fake_last_error()
    assert s.release(t) == OK
    t = s.submit(g, mu=4)[1]
    assert s.wait(t) == OK and s.result(t)[1][3] == 4
    assert s.stats().failed == 2


def test_slots_bound_the_queries_holding_results(mk):
    """Two result slots: a third query waits until a result is released."""
    s = mk(budget=1 << 30, slots=2, max_batch=4, gated=False)
    g = s.add()[1]
    ts = [s.submit(g, mu=k)[1] for k in (1, 2, 3)]
    assert s.wait(ts[0]) == OK and s.wait(ts[1]) == OK
    time.sleep(0.1)
    assert s.state(ts[2]) == PENDING
    assert s.release(ts[0]) == OK
    assert s.wait(ts[2], -1) == OK and s.result(ts[2])[1][3] == 3
    for t in ts[1:]:
        assert s.release(t) == OK


def test_result_slot_reuse_clears_cluster_ids(mk):
    """The second query on a reused slot shows -1 wherever it has no core,
    not the first query's cluster ids."""
    s = mk(budget=1 << 30, slots=1, gated=False)
    g = s.add(40)[1]
    t = s.submit(g, p=0, q=3, mu=1)[1]
    assert s.wait(t) == OK
    _, _, lab1, cid1, _ = s.result(t, 40)
    assert np.all(cid1[lab1 == 1] >= 0) and np.all(cid1[lab1 != 1] == -1)
    assert s.release(t) == OK
    t = s.submit(g, p=1, q=3, mu=1)[1]
    assert s.wait(t) == OK
    _, _, lab2, cid2, _ = s.result(t, 40)
    assert not np.array_equal(lab1, lab2)
    assert np.all(cid2[lab2 != 1] == -1) and np.all(cid2[lab2 == 1] >= 0)


def test_batches_put_index_queries_first(mk, lib):
    """Queries queued behind a running batch go out together, the ones the
    similarity index decides first."""
    s = mk(budget=1 << 30, max_batch=8, slots=8)
    g = s.add()[1]
    t0 = s.submit(g, p=1, q=2)[1]
    s.until(lambda: s.inflight() == 1, "the first batch")
    ts = [s.submit(g, p=p, q=7, mu=2)[1] for p in range(6)]
    s.complete()
    s.until(lambda: s.inflight() == 6, "one batch of six")
    s.complete()
    for t in [t0] + ts:
        assert s.wait(t) == OK
    assert lib.fake_order_violations(s.h) == 0
    st = s.stats()
    assert st.index_queries == sum(1 for p in range(6) if (p + 7) % 2 == 0)
    assert (st.max_batch_seen, st.batches) == (6, 2)


def test_argument_errors(mk):
    s = mk(budget=1 << 30)
    g = s.add()[1]
    assert s.submit(g + 5)[0] == ENOTFOUND
    assert s.submit(g, p=3, q=2)[0] == EINVAL
    assert s.wait(12345, 0) == ENOTFOUND and s.release(12345) == ENOTFOUND
    t = s.submit(g)[1]
    assert s.L.fake_remove_graph(s.h, g) == EINVAL  # a query in flight
    assert s.L.fake_remove_graph(s.h, g + 9) == ENOTFOUND
    s.complete()
    assert s.wait(t) == OK and s.release(t) == OK
    assert s.L.fake_remove_graph(s.h, g) == OK and s.stats().resident_graphs == 0


def test_shutdown_releases_everything(mk, lib):
    """Shutting down with a batch in flight lets it finish, then releases
    every resident graph: only the result slots stay allocated."""
    s = mk(budget=1 << 30, slots=2)
    gs = [s.add()[1] for _ in range(3)]
    s.submit(gs[0])
    s.until(lambda: s.inflight() == 1, "a batch")
    slot_bytes = s.stats().slot_bytes
    lib.fake_shutdown(s.h)
    assert lib.fake_releases(s.h) == 3
    assert lib.fake_allocated(s.h) == slot_bytes
