// Serving loop for SCAN queries over resident graphs (include/acc_cn.h,
// "Serving").
//
// Graphs are registered with a host copy of their CSR and kept resident in
// HBM, with their common-neighbor counts computed once at admission, while
// their device footprint fits the configured budget; admitting a graph that
// does not fit evicts the least recently used resident graphs that have no
// query running (they are admitted again, from the host copy, when a query
// needs them).  Queries are (graph, eps = p/q, mu) triples.  A worker thread
// forms batches of pending queries across graphs and runs each batch as one
// set of SCAN launches (one job table, acc_scan.hip) on one of several
// streams, so a batch is launched while earlier batches are still on the
// device: arrivals are admitted continuously as result slots and streams
// free up.  Each query owns a result slot (device buffers plus pinned host
// memory that the last kernel writes its labels, cluster ids and counters
// into) from launch until the caller releases it.  One thread per stream
// waits for that stream's batch; latency runs from submission to that
// thread seeing the batch complete, with the results already in host memory.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "acc_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct ServedGraph {
  std::vector<uint32_t> off;
  std::vector<int32_t> adj;
  uint32_t n = 0;
  uint64_t m = 0;
  acc_graph *g = nullptr;  // resident copy, or NULL when evicted
  uint64_t bytes = 0;      // device bytes when resident (estimate until uploaded)
  uint64_t last_use = 0;   // LRU clock
  int running = 0;         // queries launched and not complete
  int waiting = 0;         // queries pending
  bool validated = false;  // the host copy passed the upload checks once
};

enum QueryState { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Query {
  int32_t gid = -1;
  uint32_t n = 0;  // the graph's vertex count, kept for results read after the graph is removed
  int64_t p = 0, q = 1;
  int32_t mu = 0;
  Clock::time_point t_submit, t_done;
  int slot = -1;
  int state = kPending;
  int err = ACC_OK;
  std::string err_text;
  int64_t counts[4] = {0, 0, 0, 0};
};

// Result slot: device buffers of one query plus pinned host memory that the
// last SCAN kernel writes the results into (mapped into the device's address
// space: d_* are the device-side addresses of h_*).  The host memory is
// coherent, so the kernel's writes bypass the GPU caches and are all in host
// memory once the batch's completion event has fired.
struct Slot {
  ScanBuffers buf;
  int8_t *h_label = nullptr;
  int32_t *h_cid = nullptr;
  unsigned long long *h_counters = nullptr;
  int8_t *d_label = nullptr;
  int32_t *d_cid = nullptr;
  unsigned long long *d_counters = nullptr;
  uint32_t h_cap = 0;
  bool busy = false;
};

// One stream with at most one batch in flight, and the thread that waits for
// the batch's completion event.
struct Batch {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  JobTable index_jobs{};  // queries the similarity index decides
  JobTable slot_jobs{};   // the others
  std::vector<uint64_t> tickets;
  bool busy = false;
  std::thread waiter;
};

uint64_t graph_estimate(uint32_t n, uint64_t m) {
  // offsets, adjacency, counts, reverse slots, slot sources, similarity index, work plans
  return 4ull * (n + 1) + 4ull * (m + 64) + 4ull * (m + 4) * 2 + 4ull * m + 12ull * m + 16ull * (n + m / 64 + 16) +
         (1ull << 20);
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct acc_server {
  acc_ctx *ctx = nullptr;
  acc_server_config cfg{};
  std::mutex mu;
  std::mutex admit_mu;  // serialises uploads on the context stream
  std::condition_variable cv_work, cv_done;
  std::map<int32_t, ServedGraph> graphs;
  int32_t next_gid = 0;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, Query> queries;
  uint64_t next_ticket = 1;
  std::vector<Slot> slots;
  std::vector<Batch> batches;
  std::condition_variable cv_batch;  // a batch was launched
  std::thread worker;
  bool stop = false;
  uint64_t clock = 0;
  uint64_t graph_bytes = 0;  // resident graphs
  uint64_t slot_bytes = 0;   // result slots
  acc_server_stats st{};
};

namespace {

void free_slot(Slot &s) {
  accscan::free_buffers(&s.buf);
  if (s.h_label) (void)hipHostFree(s.h_label);
  if (s.h_cid) (void)hipHostFree(s.h_cid);
  if (s.h_counters) (void)hipHostFree(s.h_counters);
  s = Slot{};
}

// Device bytes the budget still has room for.
uint64_t room(const acc_server *s) {
  const uint64_t used = s->graph_bytes + s->slot_bytes;
  return used >= s->cfg.hbm_budget ? 0 : s->cfg.hbm_budget - used;
}

// Evicts least recently used resident graphs with no running or pending
// query (other than `keep`) until `need` bytes fit.  Caller holds s->mu.
bool make_room(acc_server *s, uint64_t need, int32_t keep) {
  while (room(s) < need) {
    ServedGraph *victim = nullptr;
    for (auto &kv : s->graphs) {
      ServedGraph &sg = kv.second;
      if (kv.first == keep || !sg.g || sg.running || sg.waiting) continue;
      if (!victim || sg.last_use < victim->last_use) victim = &sg;
    }
    if (!victim) return false;
    acc_graph_release(victim->g);
    victim->g = nullptr;
    s->graph_bytes -= victim->bytes;
    s->st.evictions++;
  }
  return true;
}

// Uploads graph gid and computes its counts.  Called without s->mu held
// (the upload takes milliseconds); s->admit_mu serialises admissions.
int upload_graph(acc_server *s, ServedGraph &sg) {
  acc_graph *g = nullptr;
  int rc = acc_graph_upload(s->ctx, sg.n, sg.m, sg.off.data(), sg.adj.data(),
                            sg.validated ? ACC_UPLOAD_NO_VALIDATE : 0u, &g);
  if (!rc) rc = acc_cn_compute(g, ACC_PATH_AUTO, 0);
  if (!rc) rc = acc_graph_ensure_src(g);
  if (!rc) rc = acc_ctx_synchronize(s->ctx);
  // the index is an optimisation: queries it cannot decide use the per-slot kernels
  if (!rc && !s->cfg.no_index && accscan::build_index(g) == ACC_OK) rc = acc_ctx_synchronize(s->ctx);
  if (rc) {
    if (g) acc_graph_release(g);
    return rc;
  }
  sg.g = g;
  sg.validated = true;
  return ACC_OK;
}

// Makes gid resident: reserves its footprint under the lock, uploads without
// it.  Caller holds `lk` on s->mu; it is released during the upload.
int admit(acc_server *s, std::unique_lock<std::mutex> &lk, int32_t gid) {
  auto it = s->graphs.find(gid);
  if (it == s->graphs.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(gid));
  ServedGraph &sg = it->second;
  if (sg.g) return ACC_OK;
  const uint64_t need = graph_estimate(sg.n, sg.m);
  if (need > s->cfg.hbm_budget) return acc_set_error(ACC_EBUDGET, "graph exceeds the HBM budget");
  if (!make_room(s, need, gid)) return acc_set_error(ACC_EBUDGET, "no idle graph left to evict");
  s->graph_bytes += need;
  sg.bytes = need;
  sg.running++;  // pins the entry against eviction and removal while unlocked
  lk.unlock();
  int rc;
  {
    std::lock_guard<std::mutex> ag(s->admit_mu);
    rc = upload_graph(s, sg);
  }
  lk.lock();
  sg.running--;
  s->graph_bytes -= need;
  if (rc) {
    sg.bytes = 0;
    return rc;
  }
  uint64_t actual = 0;
  (void)acc_graph_info(sg.g, nullptr, nullptr, &actual);
  sg.bytes = actual;
  s->graph_bytes += actual;
  s->st.admissions++;
  return ACC_OK;
}

// Grows slot k's buffers for a graph of n vertices and m slots.  Caller holds s->mu.
int reserve_slot(acc_server *s, Slot &sl, uint32_t n, uint64_t m) {
  const uint64_t before = sl.buf.bytes;
  if (!(sl.buf.counters && n <= sl.buf.n_cap && m <= sl.buf.m_cap)) {
    const uint64_t grow = accscan::buffer_bytes(std::max(n, sl.buf.n_cap), std::max(m, sl.buf.m_cap));
    if (grow > before && !make_room(s, grow - before, -1))
      return acc_set_error(ACC_EBUDGET, "no room for the result buffers of a query");
    ACC_TRY(accscan::reserve_buffers(&sl.buf, n, m));
    s->slot_bytes += sl.buf.bytes - before;
  }
  if (n > sl.h_cap || !sl.h_counters) {
    if (sl.h_label) (void)hipHostFree(sl.h_label);
    if (sl.h_cid) (void)hipHostFree(sl.h_cid);
    sl.h_label = nullptr;
    sl.h_cid = nullptr;
    sl.h_cap = 0;
    const size_t k = std::max<uint32_t>(n, 1);
    if (hipHostMalloc(reinterpret_cast<void **>(&sl.h_label), k, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&sl.h_cid), sizeof(int32_t) * k, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      return acc_set_error(ACC_ENOMEM, "hipHostMalloc of result buffers failed");
    if (!sl.h_counters &&
        hipHostMalloc(reinterpret_cast<void **>(&sl.h_counters), kScanCounters * sizeof(unsigned long long),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      return acc_set_error(ACC_ENOMEM, "hipHostMalloc of result counters failed");
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&sl.d_label), sl.h_label, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&sl.d_cid), sl.h_cid, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&sl.d_counters), sl.h_counters, 0) != hipSuccess)
      return acc_set_error(ACC_EDEVICE, "device address of result buffers unavailable");
    std::fill(sl.h_cid, sl.h_cid + k, -1);  // kernels write only the cores' entries
    sl.h_cap = static_cast<uint32_t>(k);
  }
  return ACC_OK;
}

void fail_query(acc_server *s, Query &q, int rc) {
  q.state = kFailed;
  q.err = rc;
  q.err_text = acc_last_error();
  q.t_done = Clock::now();
  s->st.failed++;
}

// Marks a finished batch's queries done (e: its completion status).  Caller holds s->mu.
void finish_batch(acc_server *s, Batch &b, hipError_t e) {
  const Clock::time_point now = Clock::now();
  for (uint64_t t : b.tickets) {
    auto it = s->queries.find(t);
    if (it == s->queries.end()) continue;
    Query &q = it->second;
    auto gi = s->graphs.find(q.gid);  // pinned by the running count: removal is refused meanwhile
    if (gi != s->graphs.end()) gi->second.running--;
    if (e != hipSuccess) {
      acc_set_error(ACC_EDEVICE, std::string("SCAN batch: ") + hipGetErrorString(e));
      fail_query(s, q, ACC_EDEVICE);
      continue;
    }
    const Slot &sl = s->slots[q.slot];
    for (int i = 0; i < 4; i++) q.counts[i] = static_cast<int64_t>(sl.h_counters[i]);
    q.state = kDone;
    q.t_done = now;
    s->st.completed++;
  }
  b.tickets.clear();
  b.busy = false;
}

// Completion thread of one stream: waits for each launched batch's event.
void waiter_main(acc_server *s, Batch *b) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_batch.wait(lk, [&]() { return s->stop || b->busy; });
    if (!b->busy) return;  // stopping with nothing in flight
    lk.unlock();
    const hipError_t e = hipEventSynchronize(b->done);
    lk.lock();
    finish_batch(s, *b, e);
    s->cv_done.notify_all();
    s->cv_work.notify_one();
  }
}

bool any_running(const acc_server *s) {
  for (const Batch &b : s->batches)
    if (b.busy) return true;
  return false;
}

// Admits the graph of the first pending query whose graph is not resident.
// Returns true when it did something (the caller rescans the queue, which
// may have changed while the lock was released).  Caller holds `lk`.
bool admit_pending(acc_server *s, std::unique_lock<std::mutex> &lk) {
  for (size_t i = 0; i < s->pending.size(); i++) {
    const uint64_t t = s->pending[i];
    Query &q = s->queries[t];
    if (s->graphs[q.gid].g) continue;
    const int rc = admit(s, lk, q.gid);
    if (rc == ACC_EBUDGET && any_running(s)) return false;  // retry once running queries free graphs
    if (rc) {
      auto qi = s->queries.find(t);
      if (qi != s->queries.end() && qi->second.state == kPending) {
        s->graphs[qi->second.gid].waiting--;
        fail_query(s, qi->second, rc);
        auto pi = std::find(s->pending.begin(), s->pending.end(), t);
        if (pi != s->pending.end()) s->pending.erase(pi);
        s->cv_done.notify_all();
      }
    }
    return true;
  }
  return false;
}

// Launches one batch of pending queries whose graphs are resident, if a
// stream and result slots are free.  Caller holds s->mu.
bool launch_batch(acc_server *s) {
  Batch *b = nullptr;
  for (Batch &x : s->batches)
    if (!x.busy) {
      b = &x;
      break;
    }
  if (!b || s->pending.empty()) return false;
  std::vector<uint64_t> take;
  std::vector<int> slot_of;
  const size_t cap = static_cast<size_t>(s->cfg.max_batch);
  size_t next_slot = 0;
  for (size_t i = 0; i < s->pending.size() && take.size() < cap; i++) {
    const uint64_t t = s->pending[i];
    Query &q = s->queries[t];
    ServedGraph &sg = s->graphs[q.gid];
    if (!sg.g) continue;
    while (next_slot < s->slots.size() && s->slots[next_slot].busy) next_slot++;
    if (next_slot == s->slots.size()) break;  // every slot holds a result not yet released
    const int rc = reserve_slot(s, s->slots[next_slot], sg.n, sg.m);
    if (rc == ACC_EBUDGET && (any_running(s) || !take.empty())) break;
    if (rc) {
      sg.waiting--;
      fail_query(s, q, rc);
      s->pending[i] = 0;  // erased below
      s->cv_done.notify_all();
      continue;
    }
    s->slots[next_slot].busy = true;
    take.push_back(t);
    slot_of.push_back(static_cast<int>(next_slot));
    s->pending[i] = 0;
  }
  s->pending.erase(std::remove(s->pending.begin(), s->pending.end(), 0ull), s->pending.end());
  if (take.empty()) return false;
  // queries the similarity index decides go first in the job table, the
  // rest (eps beyond the index's range, graphs without an index) after
  const int nj = static_cast<int>(take.size());
  {
    // each slot was grown for its own query's graph: keep the pairs together
    std::vector<std::pair<uint64_t, int>> ts;
    for (int j = 0; j < nj; j++) ts.emplace_back(take[j], slot_of[j]);
    std::stable_partition(ts.begin(), ts.end(), [&](const std::pair<uint64_t, int> &x) {
      const Query &q = s->queries[x.first];
      return accscan::index_usable(s->graphs[q.gid].g, q.p, q.q);
    });
    for (int j = 0; j < nj; j++) {
      take[j] = ts[j].first;
      slot_of[j] = ts[j].second;
    }
  }
  int ni = 0;
  for (int j = 0; j < nj; j++) {
    Query &q = s->queries[take[j]];
    ServedGraph &sg = s->graphs[q.gid];
    sg.waiting--;
    sg.running++;
    sg.last_use = ++s->clock;
    q.state = kRunning;
    if (accscan::index_usable(sg.g, q.p, q.q)) ni++;
  }
  for (int j = 0; j < nj; j++) {
    Query &q = s->queries[take[j]];
    q.slot = slot_of[j];
    const Slot &sl = s->slots[q.slot];
    ScanJob &job = j < ni ? b->index_jobs.jobs[j] : b->slot_jobs.jobs[j - ni];
    job = accscan::make_job(s->graphs[q.gid].g, sl.buf, q.p, q.q, q.mu);
    job.out_label = sl.d_label;
    job.out_cid = sl.d_cid;
    job.out_counters = sl.d_counters;
  }
  uint32_t vb_i = 0, sb_i = 0, vblocks = 0, sblocks = 0;
  b->index_jobs.nj = ni;
  b->slot_jobs.nj = nj - ni;
  accscan::plan_jobs(b->index_jobs.jobs, ni, s->ctx->num_cus, &vb_i, &sb_i);
  accscan::plan_jobs(b->slot_jobs.jobs, nj - ni, s->ctx->num_cus, &vblocks, &sblocks);
  hipError_t e = hipSuccess;
  int rc = ACC_OK;
  if (ni) rc = accscan::launch_index_jobs(b->index_jobs, vb_i, sb_i, b->stream);
  if (!rc && nj > ni) rc = accscan::launch_jobs(b->slot_jobs, vblocks, sblocks, b->stream);
  if (!rc && (e = hipEventRecord(b->done, b->stream)) != hipSuccess)
    rc = acc_set_error(ACC_EDEVICE, std::string("hipEventRecord: ") + hipGetErrorString(e));
  if (rc) {
    (void)hipStreamSynchronize(b->stream);
    for (uint64_t t : take) {
      Query &q = s->queries[t];
      s->graphs[q.gid].running--;
      fail_query(s, q, rc);
    }
    s->cv_done.notify_all();
    return true;
  }
  b->busy = true;
  b->tickets = take;
  s->cv_batch.notify_all();
  s->st.batches++;
  s->st.index_queries += static_cast<uint64_t>(ni);
  s->st.batched_queries += static_cast<uint64_t>(nj);
  s->st.max_batch_seen = std::max<uint64_t>(s->st.max_batch_seen, static_cast<uint64_t>(nj));
  return true;
}

// Launch thread: admits graphs of pending queries and launches batches
// while streams and result slots are free; completions come from the
// per-stream waiter threads.
void worker_main(acc_server *s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    bool progress = false;
    while (admit_pending(s, lk)) progress = true;
    while (launch_batch(s)) progress = true;
    if (!progress) s->cv_work.wait_for(lk, std::chrono::milliseconds(50));
  }
}

void destroy_server(acc_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_batch.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Batch &b : s->batches)
    if (b.waiter.joinable()) b.waiter.join();
  (void)hipSetDevice(s->ctx->device);
  for (Batch &b : s->batches) {
    if (b.stream) (void)hipStreamSynchronize(b.stream);
    if (b.done) (void)hipEventDestroy(b.done);
    if (b.stream) (void)hipStreamDestroy(b.stream);
  }
  for (Slot &sl : s->slots) free_slot(sl);
  for (auto &kv : s->graphs)
    if (kv.second.g) acc_graph_release(kv.second.g);
  delete s;
}

}  // namespace

extern "C" {

void acc_server_config_default(acc_server_config *cfg) {
  if (!cfg) return;
  cfg->hbm_budget = 64ull << 30;
  cfg->max_batch = kMaxJobs;
  cfg->streams = 4;
  cfg->slots = 32;
  cfg->no_index = 0;
}

int acc_server_create(acc_ctx *ctx, const acc_server_config *cfg, acc_server **out) {
  if (!ctx || !out) return acc_set_error(ACC_EINVAL, "null argument");
  *out = nullptr;
  acc_server_config c;
  acc_server_config_default(&c);
  if (cfg) c = *cfg;
  if (c.max_batch < 1 || c.max_batch > kMaxJobs || c.streams < 1 || c.streams > 16 || c.slots < 1 || c.slots > 4096)
    return acc_set_error(ACC_EINVAL, "server config: max_batch in [1," + std::to_string(kMaxJobs) +
                                         "], streams in [1,16], slots in [1,4096]");
  (void)hipSetDevice(ctx->device);
  auto *s = new (std::nothrow) acc_server;
  if (!s) return acc_set_error(ACC_ENOMEM, "host allocation failed");
  s->ctx = ctx;
  s->cfg = c;
  s->slots.resize(static_cast<size_t>(c.slots));
  s->batches.resize(static_cast<size_t>(c.streams));
  for (Batch &b : s->batches) {
    hipError_t e = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      destroy_server(s);
      return acc_set_error(ACC_EDEVICE, std::string("server setup: ") + hipGetErrorString(e));
    }
  }
  for (Batch &b : s->batches) b.waiter = std::thread(waiter_main, s, &b);
  s->worker = std::thread(worker_main, s);
  *out = s;
  return ACC_OK;
}

void acc_server_destroy(acc_server *s) { destroy_server(s); }

int acc_server_add_graph(acc_server *s, uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst,
                         uint32_t flags, int32_t *graph_id) {
  if (!s || !node_off || (m && !edge_dst) || !graph_id) return acc_set_error(ACC_EINVAL, "null argument");
  if (flags & ~ACC_UPLOAD_NO_VALIDATE)
    return acc_set_error(ACC_EINVAL, "served graphs take only ACC_UPLOAD_NO_VALIDATE (relabel before adding)");
  if (node_off[0] != 0 || node_off[n] != m) return acc_set_error(ACC_EFORMAT, "node_off must run from 0 to m");
  std::unique_lock<std::mutex> lk(s->mu);
  const int32_t gid = s->next_gid++;
  ServedGraph &sg = s->graphs[gid];
  sg.n = n;
  sg.m = m;
  sg.off.assign(node_off, node_off + static_cast<size_t>(n) + 1);
  sg.adj.assign(edge_dst, edge_dst + m);
  sg.validated = (flags & ACC_UPLOAD_NO_VALIDATE) != 0;
  const int rc = admit(s, lk, gid);
  if (rc) {
    s->graphs.erase(gid);
    return rc;
  }
  s->graphs[gid].last_use = ++s->clock;
  // Grow the idle result slots for this graph now, so that queries do not
  // allocate device or pinned memory while they are served (as far as the
  // budget allows; a slot that cannot grow here grows on first use).
  for (Slot &sl : s->slots)
    if (!sl.busy && reserve_slot(s, sl, n, m) != ACC_OK) break;
  *graph_id = gid;
  return ACC_OK;
}

int acc_server_remove_graph(acc_server *s, int32_t graph_id) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->graphs.find(graph_id);
  if (it == s->graphs.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(graph_id));
  if (it->second.running || it->second.waiting) return acc_set_error(ACC_EINVAL, "graph has queries in flight");
  if (it->second.g) {
    acc_graph_release(it->second.g);
    s->graph_bytes -= it->second.bytes;
  }
  s->graphs.erase(it);
  return ACC_OK;
}

int acc_server_graph_info(acc_server *s, int32_t graph_id, uint32_t *n, uint64_t *m, int *resident,
                          uint64_t *device_bytes) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->graphs.find(graph_id);
  if (it == s->graphs.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(graph_id));
  if (n) *n = it->second.n;
  if (m) *m = it->second.m;
  if (resident) *resident = it->second.g != nullptr;
  if (device_bytes) *device_bytes = it->second.g ? it->second.bytes : 0;
  return ACC_OK;
}

int acc_server_submit(acc_server *s, int32_t graph_id, int64_t eps_num, int64_t eps_den, int32_t mu,
                      uint64_t *ticket) {
  if (!s || !ticket) return acc_set_error(ACC_EINVAL, "null argument");
  ACC_TRY(accscan::check_eps(eps_num, eps_den));
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->graphs.find(graph_id);
  if (it == s->graphs.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(graph_id));
  const uint64_t t = s->next_ticket++;
  Query &q = s->queries[t];
  q.gid = graph_id;
  q.n = it->second.n;
  q.p = eps_num;
  q.q = eps_den;
  q.mu = mu;
  q.t_submit = Clock::now();
  it->second.waiting++;
  s->pending.push_back(t);
  s->st.submitted++;
  *ticket = t;
  s->cv_work.notify_one();
  return ACC_OK;
}

int acc_server_wait(acc_server *s, uint64_t ticket, int32_t timeout_ms) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto done = [&]() {
    auto it = s->queries.find(ticket);
    return it == s->queries.end() || it->second.state >= kDone;
  };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
    return acc_set_error(ACC_ETIMEDOUT, "query not complete");
  auto it = s->queries.find(ticket);
  if (it == s->queries.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
  if (it->second.state == kFailed) return acc_set_error(it->second.err, it->second.err_text);
  return ACC_OK;
}

int acc_server_poll(acc_server *s, uint64_t ticket, int *state) {
  if (!s || !state) return acc_set_error(ACC_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->queries.find(ticket);
  if (it == s->queries.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
  *state = it->second.state;
  return ACC_OK;
}

int acc_server_result(acc_server *s, uint64_t ticket, const int8_t **label, const int32_t **cluster_id,
                      uint32_t *n, int64_t *counts, double *latency_ms) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->queries.find(ticket);
  if (it == s->queries.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
  const Query &q = it->second;
  if (q.state == kFailed) return acc_set_error(q.err, q.err_text);
  if (q.state != kDone) return acc_set_error(ACC_EINVAL, "query not complete");
  const Slot &sl = s->slots[q.slot];
  if (label) *label = sl.h_label;
  if (cluster_id) *cluster_id = sl.h_cid;
  if (n) *n = q.n;
  if (counts)
    for (int i = 0; i < 4; i++) counts[i] = q.counts[i];
  if (latency_ms) *latency_ms = ms_between(q.t_submit, q.t_done);
  return ACC_OK;
}

int acc_server_release(acc_server *s, uint64_t ticket) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  int slot = -1;
  uint32_t n = 0;
  bool whole = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->queries.find(ticket);
    if (it == s->queries.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
    Query &q = it->second;
    if (q.state == kRunning) return acc_set_error(ACC_EINVAL, "query is running");
    if (q.state == kPending) {  // a pending query is cancelled
      auto pi = std::find(s->pending.begin(), s->pending.end(), ticket);
      if (pi != s->pending.end()) s->pending.erase(pi);
      auto gi = s->graphs.find(q.gid);  // pinned by the waiting count: removal is refused meanwhile
      if (gi != s->graphs.end()) gi->second.waiting--;
      s->st.cancelled++;
    }
    slot = q.slot;
    if (slot >= 0) {
      // q.n, not the graph's entry: the graph may have been removed since the query finished
      n = std::min(q.n, s->slots[slot].h_cap);
      whole = q.state == kFailed;
    }
    s->queries.erase(it);
  }
  if (slot >= 0) {
    // Back to -1 everywhere before the slot is reused: the next query's
    // kernel writes only its cores' cluster ids.  The slot stays busy (owned
    // by this call) meanwhile, so this runs without the lock.
    Slot &sl = s->slots[slot];
    if (whole) {
      std::fill(sl.h_cid, sl.h_cid + sl.h_cap, -1);
    } else {
      for (uint32_t u = 0; u < n; u++)
        if (sl.h_label[u] == 1) sl.h_cid[u] = -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    sl.busy = false;
  }
  s->cv_work.notify_one();
  return ACC_OK;
}

int acc_server_get_stats(acc_server *s, acc_server_stats *out) {
  if (!s || !out) return acc_set_error(ACC_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->graph_bytes = s->graph_bytes;
  out->slot_bytes = s->slot_bytes;
  out->resident_graphs = 0;
  for (auto &kv : s->graphs) out->resident_graphs += kv.second.g != nullptr;
  out->pending = s->pending.size();
  return ACC_OK;
}

}  // extern "C"
