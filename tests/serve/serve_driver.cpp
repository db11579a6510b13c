// CPU driver of the serving core (acctricnt_amd/csrc/serve_core.h, the code
// behind acc_server_* in libacc_cn.so) over a fake device, for the tests in
// tests/test_serve_host.py.  No GPU is used.
//
// The fake device has
//  - an allocator with a hard limit of its own (beside the server's budget),
//    which uploads and result slots draw from, and failures injected on
//    demand (uploads, launches, batch completions);
//  - batches that complete only when the test lets them (or at once), each
//    writing results derived from its query (counters n, p, q, mu; labels
//    and cluster ids from the vertex id), like the SCAN kernels writing only
//    the cores' cluster ids;
//  - a clock that moves only when the test moves it.
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "serve_core.h"

namespace {
thread_local std::string g_err;
}

int acc_set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
extern "C" const char *acc_last_error(void) { return g_err.c_str(); }

namespace {

struct FakeGraph {
  uint32_t n;
  uint64_t m;
  uint64_t bytes;
};

class FakeDevice final : public accserve::Device {
 public:
  FakeDevice(uint64_t phys_limit, int streams, int slots, bool gated)
      : phys_limit_(phys_limit), gated_(gated), gate_(static_cast<size_t>(streams), 0),
        inflight_(static_cast<size_t>(streams)), slots_(static_cast<size_t>(slots)) {}

  int check_query(int64_t p, int64_t q, int32_t) override {
    if (p < 0 || q <= 0 || p > q) return acc_set_error(ACC_EINVAL, "eps must be p/q with 0 <= p <= q");
    return ACC_OK;
  }
  uint64_t estimate(uint32_t n, uint64_t m) override { return 1000 + 8ull * m + 4ull * n; }
  int upload(uint32_t n, uint64_t m, const uint32_t *, const int32_t *, bool, void **handle,
             uint64_t *bytes) override {
    std::lock_guard<std::mutex> lk(mu_);
    uploads_++;
    if (fail_uploads_ > 0) {
      fail_uploads_--;
      return acc_set_error(ACC_ENOMEM, "fake allocator: upload refused");
    }
    const uint64_t b = estimate(n, m) - 100;  // the actual footprint, below the estimate
    if (allocated_ + b > phys_limit_) return acc_set_error(ACC_ENOMEM, "fake allocator: out of memory");
    allocated_ += b;
    *handle = new FakeGraph{n, m, b};
    *bytes = b;
    return ACC_OK;
  }
  void release(void *handle) override {
    std::lock_guard<std::mutex> lk(mu_);
    auto *g = static_cast<FakeGraph *>(handle);
    allocated_ -= g->bytes;
    releases_++;
    delete g;
  }
  static uint64_t slot_bytes_for(uint32_t n, uint64_t m) { return 100 + n + 4ull * m; }
  uint64_t slot_growth(int k, uint32_t n, uint64_t m) override {
    const FSlot &s = slots_[k];
    if (s.bytes && n <= s.n_cap && m <= s.m_cap) return 0;
    const uint64_t want = slot_bytes_for(std::max(n, s.n_cap), std::max(m, s.m_cap));
    return want > s.bytes ? want - s.bytes : 0;
  }
  int grow_slot(int k, uint32_t n, uint64_t m, uint64_t *added) override {
    std::lock_guard<std::mutex> lk(mu_);
    FSlot &s = slots_[k];
    *added = 0;
    if (s.bytes && n <= s.n_cap && m <= s.m_cap) return ACC_OK;
    const uint32_t nn = std::max(n, s.n_cap);
    const uint64_t mm = std::max(m, s.m_cap);
    const uint64_t want = slot_bytes_for(nn, mm);
    if (allocated_ - s.bytes + want > phys_limit_) return acc_set_error(ACC_ENOMEM, "fake allocator: no room for a slot");
    allocated_ = allocated_ - s.bytes + want;
    *added = want - s.bytes;
    s.bytes = want;
    s.n_cap = nn;
    s.m_cap = mm;
    s.label.resize(std::max<uint32_t>(nn, 1), 0);
    s.cid.resize(std::max<uint32_t>(nn, 1), -1);
    return ACC_OK;
  }
  uint32_t slot_cap(int k) override { return static_cast<uint32_t>(slots_[k].cid.size()); }
  bool index_usable(void *, int64_t p, int64_t q) override { return (p + q) % 2 == 0; }
  int launch(int s, const std::vector<accserve::LaunchItem> &items, int n_index) override {
    std::lock_guard<std::mutex> lk(mu_);
    if (fail_launches_ > 0) {
      fail_launches_--;
      return acc_set_error(ACC_EDEVICE, "fake launch failure");
    }
    for (size_t j = 0; j < items.size(); j++)
      if (index_usable(items[j].graph, items[j].p, items[j].q) != (static_cast<int>(j) < n_index)) order_violations_++;
    inflight_[s] = items;
    launches_++;
    return ACC_OK;
  }
  int wait(int s) override {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&]() { return !gated_ || gate_[s] > 0; });
    if (gated_) gate_[s]--;
    const bool fail = fail_waits_ > 0;
    if (fail) fail_waits_--;
    // the "kernels" write every query's labels and its cores' cluster ids
    for (const accserve::LaunchItem &it : inflight_[s]) {
      const auto *g = static_cast<const FakeGraph *>(it.graph);
      FSlot &sl = slots_[it.slot];
      for (uint32_t u = 0; u < g->n; u++) {
        const int8_t lab = static_cast<int8_t>((u + it.p + it.mu) % 3 == 0 ? 1 : 0);
        sl.label[u] = lab;
        if (lab == 1) sl.cid[u] = static_cast<int32_t>((u + it.q) % 7);
      }
      sl.counters[0] = g->n;
      sl.counters[1] = it.p;
      sl.counters[2] = it.q;
      sl.counters[3] = it.mu;
    }
    inflight_[s].clear();
    if (fail) return acc_set_error(ACC_EDEVICE, "fake batch failure");
    return ACC_OK;
  }
  void counters(int k, int64_t out[4]) override {
    for (int i = 0; i < 4; i++) out[i] = slots_[k].counters[i];
  }
  void results(int k, const int8_t **label, const int32_t **cluster_id) override {
    *label = slots_[k].label.data();
    *cluster_id = slots_[k].cid.data();
  }
  void reset(int k, uint32_t n, bool whole) override {
    FSlot &sl = slots_[k];
    if (whole) {
      std::fill(sl.cid.begin(), sl.cid.end(), -1);
    } else {
      for (uint32_t u = 0; u < n; u++)
        if (sl.label[u] == 1) sl.cid[u] = -1;
    }
  }
  accserve::Clock::time_point now() override {
    return accserve::Clock::time_point(std::chrono::microseconds(now_us_.load()));
  }

  // test controls
  void complete(int s, int k) {
    std::lock_guard<std::mutex> lk(mu_);
    if (s < 0)
      for (int &g : gate_) g += k;
    else
      gate_[s] += k;
    cv_.notify_all();
  }
  void ungate() {
    std::lock_guard<std::mutex> lk(mu_);
    gated_ = false;
    cv_.notify_all();
  }
  int inflight(int s) {
    std::lock_guard<std::mutex> lk(mu_);
    return static_cast<int>(inflight_[s].size());
  }
  void fail(int uploads, int launches, int waits) {
    std::lock_guard<std::mutex> lk(mu_);
    fail_uploads_ = uploads;
    fail_launches_ = launches;
    fail_waits_ = waits;
  }
  void advance(int64_t ms) { now_us_ += ms * 1000; }
  uint64_t allocated() {
    std::lock_guard<std::mutex> lk(mu_);
    return allocated_;
  }
  int order_violations() {
    std::lock_guard<std::mutex> lk(mu_);
    return order_violations_;
  }
  int uploads() {
    std::lock_guard<std::mutex> lk(mu_);
    return uploads_;
  }
  int releases() {
    std::lock_guard<std::mutex> lk(mu_);
    return releases_;
  }

 private:
  struct FSlot {
    uint32_t n_cap = 0;
    uint64_t m_cap = 0;
    uint64_t bytes = 0;
    std::vector<int8_t> label;
    std::vector<int32_t> cid;
    int64_t counters[4] = {0, 0, 0, 0};
  };
  std::mutex mu_;
  std::condition_variable cv_;
  uint64_t phys_limit_;
  uint64_t allocated_ = 0;
  bool gated_;
  std::vector<int> gate_;
  std::vector<std::vector<accserve::LaunchItem>> inflight_;
  std::vector<FSlot> slots_;
  int fail_uploads_ = 0, fail_launches_ = 0, fail_waits_ = 0;
  int uploads_ = 0, releases_ = 0, launches_ = 0, order_violations_ = 0;
  std::atomic<int64_t> now_us_{1000000};
};

struct Fake {
  FakeDevice dev;
  accserve::Core core;
  Fake(uint64_t phys, const acc_server_config &cfg, bool gated)
      : dev(phys, cfg.streams, cfg.slots, gated), core(&dev, cfg) {}
};

}  // namespace

extern "C" {

void *fake_create(uint64_t budget, uint64_t phys_limit, int streams, int slots, int max_batch, int gated) {
  acc_server_config cfg{};
  cfg.hbm_budget = budget;
  cfg.streams = streams;
  cfg.slots = slots;
  cfg.max_batch = max_batch;
  auto *f = new Fake(phys_limit, cfg, gated != 0);
  f->core.start();
  return f;
}
// Lets every batch complete, then stops the core (threads, resident graphs).
void fake_shutdown(void *h) {
  auto *f = static_cast<Fake *>(h);
  f->dev.ungate();
  f->core.shutdown();
}
void fake_destroy(void *h) {
  fake_shutdown(h);
  delete static_cast<Fake *>(h);
}
int fake_add_graph(void *h, uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, int32_t *gid) {
  return static_cast<Fake *>(h)->core.add_graph(n, m, off, adj, false, gid);
}
int fake_remove_graph(void *h, int32_t gid) { return static_cast<Fake *>(h)->core.remove_graph(gid); }
int fake_graph_info(void *h, int32_t gid, int *resident, uint64_t *bytes) {
  return static_cast<Fake *>(h)->core.graph_info(gid, nullptr, nullptr, resident, bytes);
}
int fake_submit(void *h, int32_t gid, int64_t p, int64_t q, int32_t mu, uint64_t *ticket) {
  return static_cast<Fake *>(h)->core.submit(gid, p, q, mu, ticket);
}
int fake_wait(void *h, uint64_t ticket, int32_t timeout_ms) {
  return static_cast<Fake *>(h)->core.wait(ticket, timeout_ms);
}
int fake_poll(void *h, uint64_t ticket, int *state) { return static_cast<Fake *>(h)->core.poll(ticket, state); }
// Copies a completed query's counters, labels and cluster ids (cap entries at most).
int fake_result(void *h, uint64_t ticket, int64_t *counts, int8_t *label, int32_t *cid, uint32_t cap,
                double *latency_ms) {
  const int8_t *l = nullptr;
  const int32_t *c = nullptr;
  uint32_t n = 0;
  const int rc = static_cast<Fake *>(h)->core.result(ticket, &l, &c, &n, counts, latency_ms);
  if (rc) return rc;
  const uint32_t k = std::min(n, cap);
  std::memcpy(label, l, k);
  std::memcpy(cid, c, 4ull * k);
  return ACC_OK;
}
int fake_release(void *h, uint64_t ticket) { return static_cast<Fake *>(h)->core.release(ticket); }
void fake_stats(void *h, acc_server_stats *out) { static_cast<Fake *>(h)->core.stats(out); }
void fake_complete(void *h, int stream, int k) { static_cast<Fake *>(h)->dev.complete(stream, k); }
int fake_inflight(void *h, int stream) { return static_cast<Fake *>(h)->dev.inflight(stream); }
void fake_fail(void *h, int uploads, int launches, int waits) { static_cast<Fake *>(h)->dev.fail(uploads, launches, waits); }
void fake_advance(void *h, int64_t ms) {
  auto *f = static_cast<Fake *>(h);
  f->dev.advance(ms);
  f->core.clock_moved();
}
uint64_t fake_allocated(void *h) { return static_cast<Fake *>(h)->dev.allocated(); }
int fake_order_violations(void *h) { return static_cast<Fake *>(h)->dev.order_violations(); }
int fake_uploads(void *h) { return static_cast<Fake *>(h)->dev.uploads(); }
int fake_releases(void *h) { return static_cast<Fake *>(h)->dev.releases(); }
const char *fake_last_error(void) { return acc_last_error(); }

}  // extern "C"
