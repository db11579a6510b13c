// Serving loop for SCAN queries over resident graphs (include/acc_cn.h,
// "Serving"): the HIP side of the policy in serve_core.h.
//
// A resident graph is an acc_graph uploaded with its common-neighbor counts,
// slot sources and similarity index.  Each stream carries at most one batch:
// one set of SCAN launches (one job table per form, acc_scan.hip) and an
// event the stream's waiter thread blocks on.  Each query owns a result slot
// (device buffers plus pinned host memory that the last kernel writes its
// labels, cluster ids and counters into) from launch until the caller
// releases it; latency runs from submission to the waiter seeing the batch
// complete, with the results already in host memory.

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "acc_internal.h"
#include "serve_core.h"

namespace {

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
};

// One stream with at most one batch in flight and its completion event.
struct Stream {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  JobTable index_jobs{};  // queries the similarity index decides
  JobTable slot_jobs{};   // the others
};

uint64_t graph_estimate(uint32_t n, uint64_t m) {
  // offsets, adjacency, counts, reverse slots, slot sources, similarity index, work plans
  return 4ull * (n + 1) + 4ull * (m + 64) + 4ull * (m + 4) * 2 + 4ull * m + 12ull * m + 16ull * (n + m / 64 + 16) +
         (1ull << 20);
}

void free_slot(Slot &s) {
  accscan::free_buffers(&s.buf);
  if (s.h_label) (void)hipHostFree(s.h_label);
  if (s.h_cid) (void)hipHostFree(s.h_cid);
  if (s.h_counters) (void)hipHostFree(s.h_counters);
  s = Slot{};
}

class HipDevice final : public accserve::Device {
 public:
  HipDevice(acc_ctx *ctx, const acc_server_config &cfg) : ctx_(ctx), no_index_(cfg.no_index != 0) {
    slots_.resize(static_cast<size_t>(cfg.slots));
    streams_.resize(static_cast<size_t>(cfg.streams));
  }
  ~HipDevice() override {
    (void)hipSetDevice(ctx_->device);
    for (Stream &b : streams_) {
      if (b.stream) (void)hipStreamSynchronize(b.stream);
      if (b.done) (void)hipEventDestroy(b.done);
      if (b.stream) (void)hipStreamDestroy(b.stream);
    }
    for (Slot &sl : slots_) free_slot(sl);
  }
  int setup() {
    (void)hipSetDevice(ctx_->device);
    for (Stream &b : streams_) {
      hipError_t e = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
      if (e != hipSuccess) return acc_set_error(ACC_EDEVICE, std::string("server setup: ") + hipGetErrorString(e));
    }
    return ACC_OK;
  }

  int check_query(int64_t p, int64_t q, int32_t) override { return accscan::check_eps(p, q); }
  uint64_t estimate(uint32_t n, uint64_t m) override { return graph_estimate(n, m); }

  // Uploads the graph and computes its counts, slot sources and index
  // (serialised by the core; the upload takes milliseconds).
  int upload(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, bool validated, void **handle,
             uint64_t *bytes) override {
    acc_graph *g = nullptr;
    int rc = acc_graph_upload(ctx_, n, m, off, adj, validated ? ACC_UPLOAD_NO_VALIDATE : 0u, &g);
    if (!rc) rc = acc_cn_compute(g, ACC_PATH_AUTO, 0);
    if (!rc) rc = acc_graph_ensure_src(g);
    if (!rc) rc = acc_ctx_synchronize(ctx_);
    // the index is an optimisation: queries it cannot decide use the per-slot kernels
    if (!rc && !no_index_ && accscan::build_index(g) == ACC_OK) rc = acc_ctx_synchronize(ctx_);
    if (rc) {
      if (g) acc_graph_release(g);
      return rc;
    }
    uint64_t actual = 0;
    (void)acc_graph_info(g, nullptr, nullptr, &actual);
    *handle = g;
    *bytes = actual;
    return ACC_OK;
  }
  void release(void *handle) override { acc_graph_release(static_cast<acc_graph *>(handle)); }

  uint64_t slot_growth(int k, uint32_t n, uint64_t m) override {
    const Slot &sl = slots_[k];
    if (sl.buf.counters && n <= sl.buf.n_cap && m <= sl.buf.m_cap) return 0;
    const uint64_t grow = accscan::buffer_bytes(std::max(n, sl.buf.n_cap), std::max(m, sl.buf.m_cap));
    return grow > sl.buf.bytes ? grow - sl.buf.bytes : 0;
  }

  // Grows slot k's buffers for a graph of n vertices and m slots.
  int grow_slot(int k, uint32_t n, uint64_t m, uint64_t *added) override {
    Slot &sl = slots_[k];
    *added = 0;
    if (!(sl.buf.counters && n <= sl.buf.n_cap && m <= sl.buf.m_cap)) {
      const uint64_t before = sl.buf.bytes;
      const int rc = accscan::reserve_buffers(&sl.buf, n, m);
      *added = sl.buf.bytes - before;
      if (rc) return rc;
    }
    if (n > sl.h_cap || !sl.h_counters) {
      if (sl.h_label) (void)hipHostFree(sl.h_label);
      if (sl.h_cid) (void)hipHostFree(sl.h_cid);
      sl.h_label = nullptr;
      sl.h_cid = nullptr;
      sl.h_cap = 0;
      const size_t cap = std::max<uint32_t>(n, 1);
      if (hipHostMalloc(reinterpret_cast<void **>(&sl.h_label), cap, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
          hipHostMalloc(reinterpret_cast<void **>(&sl.h_cid), sizeof(int32_t) * cap, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return acc_set_error(ACC_ENOMEM, "hipHostMalloc of result buffers failed");
      if (!sl.h_counters &&
          hipHostMalloc(reinterpret_cast<void **>(&sl.h_counters), kScanCounters * sizeof(unsigned long long),
                        hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return acc_set_error(ACC_ENOMEM, "hipHostMalloc of result counters failed");
      if (hipHostGetDevicePointer(reinterpret_cast<void **>(&sl.d_label), sl.h_label, 0) != hipSuccess ||
          hipHostGetDevicePointer(reinterpret_cast<void **>(&sl.d_cid), sl.h_cid, 0) != hipSuccess ||
          hipHostGetDevicePointer(reinterpret_cast<void **>(&sl.d_counters), sl.h_counters, 0) != hipSuccess)
        return acc_set_error(ACC_EDEVICE, "device address of result buffers unavailable");
      std::fill(sl.h_cid, sl.h_cid + cap, -1);  // kernels write only the cores' entries
      sl.h_cap = static_cast<uint32_t>(cap);
    }
    return ACC_OK;
  }
  uint32_t slot_cap(int k) override { return slots_[k].h_cap; }

  bool index_usable(void *graph, int64_t p, int64_t q) override {
    return accscan::index_usable(static_cast<const acc_graph *>(graph), p, q);
  }

  int launch(int s, const std::vector<accserve::LaunchItem> &items, int n_index) override {
    Stream &b = streams_[s];
    const int nj = static_cast<int>(items.size());
    for (int j = 0; j < nj; j++) {
      const accserve::LaunchItem &it = items[j];
      const Slot &sl = slots_[it.slot];
      ScanJob &job = j < n_index ? b.index_jobs.jobs[j] : b.slot_jobs.jobs[j - n_index];
      job = accscan::make_job(static_cast<const acc_graph *>(it.graph), sl.buf, it.p, it.q, it.mu);
      job.out_label = sl.d_label;
      job.out_cid = sl.d_cid;
      job.out_counters = sl.d_counters;
    }
    uint32_t vb_i = 0, sb_i = 0, vblocks = 0, sblocks = 0;
    b.index_jobs.nj = n_index;
    b.slot_jobs.nj = nj - n_index;
    accscan::plan_jobs(b.index_jobs.jobs, n_index, ctx_->num_cus, &vb_i, &sb_i);
    accscan::plan_jobs(b.slot_jobs.jobs, nj - n_index, ctx_->num_cus, &vblocks, &sblocks);
    hipError_t e = hipSuccess;
    int rc = ACC_OK;
    if (n_index) rc = accscan::launch_index_jobs(b.index_jobs, vb_i, sb_i, b.stream);
    if (!rc && nj > n_index) rc = accscan::launch_jobs(b.slot_jobs, vblocks, sblocks, b.stream);
    if (!rc && (e = hipEventRecord(b.done, b.stream)) != hipSuccess)
      rc = acc_set_error(ACC_EDEVICE, std::string("hipEventRecord: ") + hipGetErrorString(e));
    if (rc) (void)hipStreamSynchronize(b.stream);
    return rc;
  }

  int wait(int s) override {
    const hipError_t e = hipEventSynchronize(streams_[s].done);
    if (e != hipSuccess) return acc_set_error(ACC_EDEVICE, std::string("SCAN batch: ") + hipGetErrorString(e));
    return ACC_OK;
  }

  void counters(int k, int64_t out[4]) override {
    for (int i = 0; i < 4; i++) out[i] = static_cast<int64_t>(slots_[k].h_counters[i]);
  }
  void results(int k, const int8_t **label, const int32_t **cluster_id) override {
    *label = slots_[k].h_label;
    *cluster_id = slots_[k].h_cid;
  }
  // The cores' entries (label 1) back to -1: the labels are read eight at a
  // time and a word without a core byte is skipped, so the scan costs little
  // more than reading n bytes (a query's release sits on the caller's path).
  void reset(int k, uint32_t n, bool whole) override {
    const Slot &sl = slots_[k];
    int32_t *const cid = sl.h_cid;
    const int8_t *const label = sl.h_label;
    if (whole) {
      std::fill(cid, cid + sl.h_cap, -1);
      return;
    }
    constexpr uint64_t kOnes = 0x0101010101010101ull, kHigh = 0x8080808080808080ull;
    uint32_t u = 0;
    for (; u + 8 <= n; u += 8) {
      uint64_t w;
      std::memcpy(&w, label + u, 8);
      const uint64_t x = w ^ kOnes;  // a zero byte where the label is 1
      if (((x - kOnes) & ~x & kHigh) == 0) continue;
      for (uint32_t v = u; v < u + 8; v++)
        if (label[v] == 1) cid[v] = -1;
    }
    for (; u < n; u++)
      if (label[u] == 1) cid[u] = -1;
  }
  void bind_thread() override { (void)hipSetDevice(ctx_->device); }

 private:
  acc_ctx *ctx_;
  bool no_index_;
  std::vector<Slot> slots_;
  std::vector<Stream> streams_;
};

}  // namespace

struct acc_server {
  HipDevice dev;
  accserve::Core core;
  acc_server(acc_ctx *ctx, const acc_server_config &cfg) : dev(ctx, cfg), core(&dev, cfg) {}
};

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
  auto *s = new (std::nothrow) acc_server(ctx, c);
  if (!s) return acc_set_error(ACC_ENOMEM, "host allocation failed");
  const int rc = s->dev.setup();
  if (rc) {
    delete s;
    return rc;
  }
  s->core.start();
  *out = s;
  return ACC_OK;
}

// The core stops its threads and releases the graphs first, then the
// device side synchronises its streams and frees the result slots.
void acc_server_destroy(acc_server *s) { delete s; }

int acc_server_add_graph(acc_server *s, uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst,
                         uint32_t flags, int32_t *graph_id) {
  if (!s || !node_off || (m && !edge_dst) || !graph_id) return acc_set_error(ACC_EINVAL, "null argument");
  if (flags & ~ACC_UPLOAD_NO_VALIDATE)
    return acc_set_error(ACC_EINVAL, "served graphs take only ACC_UPLOAD_NO_VALIDATE (relabel before adding)");
  if (node_off[0] != 0 || node_off[n] != m) return acc_set_error(ACC_EFORMAT, "node_off must run from 0 to m");
  return s->core.add_graph(n, m, node_off, edge_dst, (flags & ACC_UPLOAD_NO_VALIDATE) != 0, graph_id);
}

int acc_server_remove_graph(acc_server *s, int32_t graph_id) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  return s->core.remove_graph(graph_id);
}

int acc_server_graph_info(acc_server *s, int32_t graph_id, uint32_t *n, uint64_t *m, int *resident,
                          uint64_t *device_bytes) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  return s->core.graph_info(graph_id, n, m, resident, device_bytes);
}

int acc_server_submit(acc_server *s, int32_t graph_id, int64_t eps_num, int64_t eps_den, int32_t mu,
                      uint64_t *ticket) {
  if (!s || !ticket) return acc_set_error(ACC_EINVAL, "null argument");
  return s->core.submit(graph_id, eps_num, eps_den, mu, ticket);
}

int acc_server_wait(acc_server *s, uint64_t ticket, int32_t timeout_ms) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  return s->core.wait(ticket, timeout_ms);
}

int acc_server_poll(acc_server *s, uint64_t ticket, int *state) {
  if (!s || !state) return acc_set_error(ACC_EINVAL, "null argument");
  return s->core.poll(ticket, state);
}

int acc_server_result(acc_server *s, uint64_t ticket, const int8_t **label, const int32_t **cluster_id,
                      uint32_t *n, int64_t *counts, double *latency_ms) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  return s->core.result(ticket, label, cluster_id, n, counts, latency_ms);
}

int acc_server_release(acc_server *s, uint64_t ticket) {
  if (!s) return acc_set_error(ACC_EINVAL, "null server");
  return s->core.release(ticket);
}

int acc_server_get_stats(acc_server *s, acc_server_stats *out) {
  if (!s || !out) return acc_set_error(ACC_EINVAL, "null argument");
  s->core.stats(out);
  return ACC_OK;
}

}  // extern "C"
