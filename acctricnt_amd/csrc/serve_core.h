// Serving loop for SCAN queries over resident graphs: the policy, host only.
//
// Graphs are registered with a host copy of their CSR and kept resident on
// the device, with their common-neighbor counts computed once at admission,
// while their device footprint fits the configured budget; admitting a graph
// that does not fit evicts the least recently used resident graphs that have
// no query running or pending (they are admitted again, from the host copy,
// when a query needs them).  Queries are (graph, eps = p/q, mu) triples.  A
// worker thread forms batches of pending queries across graphs and launches
// each batch on one of several streams as soon as a stream and result slots
// are free, so batches overlap on the device; one waiter thread per stream
// blocks until that stream's batch completes and marks its queries done.
// Each query owns a result slot from launch until the caller releases it;
// releasing a pending query cancels it.
//
// Everything that touches the device -- uploads, result slots, launches,
// completion -- goes through the Device interface, and so does the clock
// that stamps latencies and bounds timed waits.  acc_server.hip implements
// Device over HIP; a CPU test implements it with a fake allocator, gated
// completions and a manual clock, and drives this same code.

#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "acc_cn.h"

// Sets the calling thread's last error (acc_last_error) and returns code.
int acc_set_error(int code, const std::string &msg);

namespace accserve {

using Clock = std::chrono::steady_clock;

// One query of a launched batch: its resident graph, result slot and parameters.
struct LaunchItem {
  void *graph;
  int slot;
  int64_t p, q;
  int32_t mu;
};

// The device side of the serving loop.  Calls come from the API's callers
// and from the core's threads, always with the core's lock held except
// upload (serialised by the core), wait and reset.
class Device {
 public:
  virtual ~Device() = default;
  // Rejects an invalid query (ACC_EINVAL with acc_set_error) or returns ACC_OK.
  virtual int check_query(int64_t p, int64_t q, int32_t mu) = 0;
  // Device bytes reserved for a graph of n vertices and m slots before its upload.
  virtual uint64_t estimate(uint32_t n, uint64_t m) = 0;
  // Makes a graph resident and computes its counts (and similarity index):
  // *handle and its actual footprint *bytes on success.
  virtual int upload(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, bool validated, void **handle,
                     uint64_t *bytes) = 0;
  virtual void release(void *handle) = 0;
  // Device bytes result slot k must add to hold a query on (n, m); 0 when it fits.
  virtual uint64_t slot_growth(int k, uint32_t n, uint64_t m) = 0;
  // Grows slot k for (n, m); *added = the device bytes it added (also on failure).
  virtual int grow_slot(int k, uint32_t n, uint64_t m, uint64_t *added) = 0;
  // Vertices slot k's host results hold.
  virtual uint32_t slot_cap(int k) = 0;
  // Whether the graph's similarity index decides this eps.
  virtual bool index_usable(void *graph, int64_t p, int64_t q) = 0;
  // Launches a batch on stream s (the first n_index items decided by the
  // index) and records its completion; on failure nothing stays in flight.
  virtual int launch(int s, const std::vector<LaunchItem> &items, int n_index) = 0;
  // Blocks until stream s's batch completes: ACC_OK, or an error code with acc_set_error.
  virtual int wait(int s) = 0;
  // A completed query's counters (cores, clusters, hubs, outliers) in slot k.
  virtual void counters(int k, int64_t out[4]) = 0;
  virtual void results(int k, const int8_t **label, const int32_t **cluster_id) = 0;
  // Returns slot k's cluster ids to -1 after a query on n vertices (every entry when whole).
  virtual void reset(int k, uint32_t n, bool whole) = 0;
  // Called first in every thread the core starts.
  virtual void bind_thread() {}
  virtual Clock::time_point now() { return Clock::now(); }
};

// Waits on cv for at most d (spurious wake-ups allowed: callers recheck).
// Under ThreadSanitizer the wait is taken on the system clock: GCC 11's
// sanitizer runtime does not intercept the steady-clock wait
// (pthread_cond_clockwait) and would take the mutex as still held by the
// waiting thread.
template <class Rep, class Period>
void wait_at_most(std::condition_variable &cv, std::unique_lock<std::mutex> &lk,
                  std::chrono::duration<Rep, Period> d) {
#if defined(__SANITIZE_THREAD__)
  cv.wait_until(lk, std::chrono::system_clock::now() + d);
#else
  cv.wait_for(lk, d);
#endif
}

enum QueryState { kPending = ACC_QUERY_PENDING, kRunning = ACC_QUERY_RUNNING, kDone = ACC_QUERY_DONE,
                  kFailed = ACC_QUERY_FAILED };

class Core {
 public:
  Core(Device *dev, const acc_server_config &cfg) : dev_(dev), cfg_(cfg) {
    slots_.resize(static_cast<size_t>(cfg.slots));
    batches_.resize(static_cast<size_t>(cfg.streams));
  }
  Core(const Core &) = delete;
  Core &operator=(const Core &) = delete;
  ~Core() { shutdown(); }

  // Starts the worker and the per-stream waiter threads.
  void start() {
    for (size_t b = 0; b < batches_.size(); b++) batches_[b].waiter = std::thread(&Core::waiter_main, this, b);
    worker_ = std::thread(&Core::worker_main, this);
  }

  // Stops the threads (batches in flight complete first) and releases every resident graph.
  void shutdown() {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (down_) return;
      stop_ = true;
    }
    cv_work_.notify_all();
    cv_batch_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (Batch &b : batches_)
      if (b.waiter.joinable()) b.waiter.join();
    std::lock_guard<std::mutex> g(mu_);
    for (auto &kv : graphs_)
      if (kv.second.g) dev_->release(kv.second.g);
    graphs_.clear();
    down_ = true;
  }

  int add_graph(uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst, bool validated,
                int32_t *graph_id) {
    std::unique_lock<std::mutex> lk(mu_);
    const int32_t gid = next_gid_++;
    ServedGraph &sg = graphs_[gid];
    sg.n = n;
    sg.m = m;
    sg.off.assign(node_off, node_off + static_cast<size_t>(n) + 1);
    sg.adj.assign(edge_dst, edge_dst + m);
    sg.validated = validated;
    const int rc = admit(lk, gid);
    if (rc) {
      graphs_.erase(gid);
      return rc;
    }
    graphs_[gid].last_use = ++clock_;
    // Grow the idle result slots for this graph now, so that queries do not
    // allocate while they are served, as far as the budget's free room
    // allows (evicting nothing, the new graph included); a slot that cannot
    // grow here grows on first use.
    for (size_t k = 0; k < slots_.size(); k++)
      if (!slots_[k].busy && reserve_slot(static_cast<int>(k), n, m, false) != ACC_OK) break;
    *graph_id = gid;
    return ACC_OK;
  }

  int remove_graph(int32_t graph_id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = graphs_.find(graph_id);
    if (it == graphs_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(graph_id));
    if (it->second.running || it->second.waiting) return acc_set_error(ACC_EINVAL, "graph has queries in flight");
    if (it->second.g) {
      dev_->release(it->second.g);
      graph_bytes_ -= it->second.bytes;
    }
    graphs_.erase(it);
    return ACC_OK;
  }

  int graph_info(int32_t graph_id, uint32_t *n, uint64_t *m, int *resident, uint64_t *device_bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = graphs_.find(graph_id);
    if (it == graphs_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(graph_id));
    if (n) *n = it->second.n;
    if (m) *m = it->second.m;
    if (resident) *resident = it->second.g != nullptr;
    if (device_bytes) *device_bytes = it->second.g ? it->second.bytes : 0;
    return ACC_OK;
  }

  int submit(int32_t graph_id, int64_t eps_num, int64_t eps_den, int32_t mu, uint64_t *ticket) {
    const int rc = dev_->check_query(eps_num, eps_den, mu);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = graphs_.find(graph_id);
    if (it == graphs_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(graph_id));
    const uint64_t t = next_ticket_++;
    Query &q = queries_[t];
    q.gid = graph_id;
    q.n = it->second.n;
    q.p = eps_num;
    q.q = eps_den;
    q.mu = mu;
    q.t_submit = dev_->now();
    it->second.waiting++;
    pending_.push_back(t);
    st_.submitted++;
    *ticket = t;
    cv_work_.notify_one();
    return ACC_OK;
  }

  // ACC_OK when the query is done, its error when it failed, ACC_ETIMEDOUT
  // when timeout_ms (< 0: no limit) of the device's clock pass first.
  int wait(uint64_t ticket, int32_t timeout_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto done = [&]() {
      auto it = queries_.find(ticket);
      return it == queries_.end() || it->second.state >= kDone;
    };
    if (timeout_ms < 0) {
      cv_done_.wait(lk, done);
    } else {
      const Clock::time_point deadline = dev_->now() + std::chrono::milliseconds(timeout_ms);
      while (!done()) {
        const Clock::time_point now = dev_->now();
        if (now >= deadline) return acc_set_error(ACC_ETIMEDOUT, "query not complete");
        wait_at_most(cv_done_, lk, deadline - now);
      }
    }
    auto it = queries_.find(ticket);
    if (it == queries_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
    if (it->second.state == kFailed) return acc_set_error(it->second.err, it->second.err_text);
    return ACC_OK;
  }

  // Wakes timed waits to look at the clock again (a Device whose clock is
  // moved by hand calls it after each move).
  void clock_moved() {
    std::lock_guard<std::mutex> lk(mu_);
    cv_done_.notify_all();
  }

  int poll(uint64_t ticket, int *state) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = queries_.find(ticket);
    if (it == queries_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
    *state = it->second.state;
    return ACC_OK;
  }

  int result(uint64_t ticket, const int8_t **label, const int32_t **cluster_id, uint32_t *n, int64_t *counts,
             double *latency_ms) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = queries_.find(ticket);
    if (it == queries_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
    const Query &q = it->second;
    if (q.state == kFailed) return acc_set_error(q.err, q.err_text);
    if (q.state != kDone) return acc_set_error(ACC_EINVAL, "query not complete");
    const int8_t *l = nullptr;
    const int32_t *c = nullptr;
    dev_->results(q.slot, &l, &c);
    if (label) *label = l;
    if (cluster_id) *cluster_id = c;
    if (n) *n = q.n;
    if (counts)
      for (int i = 0; i < 4; i++) counts[i] = q.counts[i];
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(q.t_done - q.t_submit).count();
    return ACC_OK;
  }

  int release(uint64_t ticket) {
    int slot = -1;
    uint32_t n = 0;
    bool whole = false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = queries_.find(ticket);
      if (it == queries_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown ticket");
      Query &q = it->second;
      if (q.state == kRunning) return acc_set_error(ACC_EINVAL, "query is running");
      if (q.state == kPending) {  // a pending query is cancelled
        auto pi = std::find(pending_.begin(), pending_.end(), ticket);
        if (pi != pending_.end()) pending_.erase(pi);
        auto gi = graphs_.find(q.gid);  // pinned by the waiting count: removal is refused meanwhile
        if (gi != graphs_.end()) gi->second.waiting--;
        st_.cancelled++;
      }
      slot = q.slot;
      if (slot >= 0) {
        // q.n, not the graph's entry: the graph may have been removed since the query finished
        n = std::min(q.n, dev_->slot_cap(slot));
        whole = q.state == kFailed;
      }
      queries_.erase(it);
    }
    if (slot >= 0) {
      // Back to -1 before the slot is reused: the next query writes only its
      // cores' cluster ids.  The slot stays busy (owned by this call)
      // meanwhile, so this runs without the lock.
      dev_->reset(slot, n, whole);
      std::lock_guard<std::mutex> lk(mu_);
      slots_[slot].busy = false;
    }
    cv_work_.notify_one();
    return ACC_OK;
  }

  void stats(acc_server_stats *out) {
    std::lock_guard<std::mutex> lk(mu_);
    *out = st_;
    out->graph_bytes = graph_bytes_;
    out->slot_bytes = slot_bytes_;
    out->resident_graphs = 0;
    for (auto &kv : graphs_) out->resident_graphs += kv.second.g != nullptr;
    out->pending = pending_.size();
  }

 private:
  struct ServedGraph {
    std::vector<uint32_t> off;
    std::vector<int32_t> adj;
    uint32_t n = 0;
    uint64_t m = 0;
    void *g = nullptr;       // resident copy, or NULL when evicted
    uint64_t bytes = 0;      // device bytes when resident (estimate until uploaded)
    uint64_t last_use = 0;   // LRU clock
    int running = 0;         // queries launched and not complete, plus an admission in progress
    int waiting = 0;         // queries pending
    bool validated = false;  // the host copy passed the upload checks once
  };
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
  struct SlotState {
    bool busy = false;  // holds a query from launch to release
  };
  struct Batch {
    std::vector<uint64_t> tickets;
    bool busy = false;
    std::thread waiter;
  };

  // Device bytes the budget still has room for.
  uint64_t room() const {
    const uint64_t used = graph_bytes_ + slot_bytes_;
    return used >= cfg_.hbm_budget ? 0 : cfg_.hbm_budget - used;
  }

  // Evicts least recently used resident graphs with no running or pending
  // query (other than `keep`) until `need` bytes fit.  Caller holds mu_.
  bool make_room(uint64_t need, int32_t keep) {
    while (room() < need) {
      ServedGraph *victim = nullptr;
      for (auto &kv : graphs_) {
        ServedGraph &sg = kv.second;
        if (kv.first == keep || !sg.g || sg.running || sg.waiting) continue;
        if (!victim || sg.last_use < victim->last_use) victim = &sg;
      }
      if (!victim) return false;
      dev_->release(victim->g);
      victim->g = nullptr;
      graph_bytes_ -= victim->bytes;
      st_.evictions++;
    }
    return true;
  }

  // Makes gid resident: reserves its footprint under the lock, uploads
  // without it (admit_mu_ serialises uploads).  Caller holds lk on mu_.
  int admit(std::unique_lock<std::mutex> &lk, int32_t gid) {
    auto it = graphs_.find(gid);
    if (it == graphs_.end()) return acc_set_error(ACC_ENOTFOUND, "unknown graph id " + std::to_string(gid));
    ServedGraph &sg = it->second;
    if (sg.g) return ACC_OK;
    const uint64_t need = dev_->estimate(sg.n, sg.m);
    if (need > cfg_.hbm_budget) return acc_set_error(ACC_EBUDGET, "graph exceeds the HBM budget");
    if (!make_room(need, gid)) return acc_set_error(ACC_EBUDGET, "no idle graph left to evict");
    graph_bytes_ += need;
    sg.bytes = need;
    sg.running++;  // pins the entry against eviction and removal while unlocked
    lk.unlock();
    int rc;
    void *handle = nullptr;
    uint64_t actual = 0;
    {
      std::lock_guard<std::mutex> ag(admit_mu_);
      rc = dev_->upload(sg.n, sg.m, sg.off.data(), sg.adj.data(), sg.validated, &handle, &actual);
    }
    lk.lock();
    sg.running--;
    graph_bytes_ -= need;
    if (rc) {
      sg.bytes = 0;
      return rc;
    }
    sg.g = handle;
    sg.validated = true;
    sg.bytes = actual;
    graph_bytes_ += actual;
    st_.admissions++;
    return ACC_OK;
  }

  // Grows slot k for a graph of n vertices and m slots, evicting idle
  // graphs for room when may_evict.  Caller holds mu_.
  int reserve_slot(int k, uint32_t n, uint64_t m, bool may_evict) {
    const uint64_t need = dev_->slot_growth(k, n, m);
    if (need && (may_evict ? !make_room(need, -1) : room() < need))
      return acc_set_error(ACC_EBUDGET, "no room for the result buffers of a query");
    uint64_t added = 0;
    const int rc = dev_->grow_slot(k, n, m, &added);
    slot_bytes_ += added;
    return rc;
  }

  void fail_query(Query &q, int rc) {
    q.state = kFailed;
    q.err = rc;
    q.err_text = acc_last_error();
    q.t_done = dev_->now();
    st_.failed++;
  }

  // Marks a finished batch's queries done (rc: its completion status).  Caller holds mu_.
  void finish_batch(Batch &b, int rc) {
    const Clock::time_point now = dev_->now();
    for (uint64_t t : b.tickets) {
      auto it = queries_.find(t);
      if (it == queries_.end()) continue;
      Query &q = it->second;
      auto gi = graphs_.find(q.gid);  // pinned by the running count: removal is refused meanwhile
      if (gi != graphs_.end()) gi->second.running--;
      if (rc) {
        fail_query(q, rc);
        continue;
      }
      dev_->counters(q.slot, q.counts);
      q.state = kDone;
      q.t_done = now;
      st_.completed++;
    }
    b.tickets.clear();
    b.busy = false;
  }

  // Completion thread of stream index bi: waits for each launched batch.
  void waiter_main(size_t bi) {
    dev_->bind_thread();
    Batch &b = batches_[bi];
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_batch_.wait(lk, [&]() { return stop_ || b.busy; });
      if (!b.busy) return;  // stopping with nothing in flight
      lk.unlock();
      const int rc = dev_->wait(static_cast<int>(bi));
      lk.lock();
      finish_batch(b, rc);
      cv_done_.notify_all();
      cv_work_.notify_one();
    }
  }

  bool any_running() const {
    for (const Batch &b : batches_)
      if (b.busy) return true;
    return false;
  }

  // Admits the graph of the first pending query whose graph is not resident.
  // Returns true when it did something (the caller rescans the queue, which
  // may have changed while the lock was released).  Caller holds lk.
  bool admit_pending(std::unique_lock<std::mutex> &lk) {
    for (size_t i = 0; i < pending_.size(); i++) {
      const uint64_t t = pending_[i];
      Query &q = queries_[t];
      if (graphs_[q.gid].g) continue;
      const int rc = admit(lk, q.gid);
      if (rc == ACC_EBUDGET && any_running()) return false;  // retry once running queries free graphs
      if (rc) {
        auto qi = queries_.find(t);
        if (qi != queries_.end() && qi->second.state == kPending) {
          graphs_[qi->second.gid].waiting--;
          fail_query(qi->second, rc);
          auto pi = std::find(pending_.begin(), pending_.end(), t);
          if (pi != pending_.end()) pending_.erase(pi);
          cv_done_.notify_all();
        }
      }
      return true;
    }
    return false;
  }

  // Launches one batch of pending queries whose graphs are resident, if a
  // stream and result slots are free.  Caller holds mu_.
  bool launch_batch() {
    int bi = -1;
    for (size_t x = 0; x < batches_.size(); x++)
      if (!batches_[x].busy) {
        bi = static_cast<int>(x);
        break;
      }
    if (bi < 0 || pending_.empty()) return false;
    std::vector<uint64_t> take;
    std::vector<int> slot_of;
    const size_t cap = static_cast<size_t>(cfg_.max_batch);
    size_t next_slot = 0;
    for (size_t i = 0; i < pending_.size() && take.size() < cap; i++) {
      const uint64_t t = pending_[i];
      Query &q = queries_[t];
      ServedGraph &sg = graphs_[q.gid];
      if (!sg.g) continue;
      while (next_slot < slots_.size() && slots_[next_slot].busy) next_slot++;
      if (next_slot == slots_.size()) break;  // every slot holds a result not yet released
      const int rc = reserve_slot(static_cast<int>(next_slot), sg.n, sg.m, true);
      if (rc == ACC_EBUDGET && (any_running() || !take.empty())) break;
      if (rc) {
        sg.waiting--;
        fail_query(q, rc);
        pending_[i] = 0;  // erased below
        cv_done_.notify_all();
        continue;
      }
      slots_[next_slot].busy = true;
      take.push_back(t);
      slot_of.push_back(static_cast<int>(next_slot));
      pending_[i] = 0;
    }
    pending_.erase(std::remove(pending_.begin(), pending_.end(), 0ull), pending_.end());
    if (take.empty()) return false;
    // queries the similarity index decides go first in the batch, the rest
    // (eps beyond the index's range, graphs without an index) after; each
    // slot was grown for its own query's graph, so the pairs stay together
    const int nj = static_cast<int>(take.size());
    std::vector<std::pair<uint64_t, int>> ts;
    for (int j = 0; j < nj; j++) ts.emplace_back(take[j], slot_of[j]);
    std::stable_partition(ts.begin(), ts.end(), [&](const std::pair<uint64_t, int> &x) {
      const Query &q = queries_[x.first];
      return dev_->index_usable(graphs_[q.gid].g, q.p, q.q);
    });
    int ni = 0;
    std::vector<LaunchItem> items;
    for (int j = 0; j < nj; j++) {
      take[j] = ts[j].first;
      Query &q = queries_[take[j]];
      ServedGraph &sg = graphs_[q.gid];
      sg.waiting--;
      sg.running++;
      sg.last_use = ++clock_;
      q.state = kRunning;
      q.slot = ts[j].second;
      if (dev_->index_usable(sg.g, q.p, q.q)) ni++;
      items.push_back(LaunchItem{sg.g, q.slot, q.p, q.q, q.mu});
    }
    const int rc = dev_->launch(bi, items, ni);
    if (rc) {
      for (uint64_t t : take) {
        Query &q = queries_[t];
        graphs_[q.gid].running--;
        fail_query(q, rc);
      }
      cv_done_.notify_all();
      return true;
    }
    Batch &b = batches_[bi];
    b.busy = true;
    b.tickets = take;
    cv_batch_.notify_all();
    st_.batches++;
    st_.index_queries += static_cast<uint64_t>(ni);
    st_.batched_queries += static_cast<uint64_t>(nj);
    st_.max_batch_seen = std::max<uint64_t>(st_.max_batch_seen, static_cast<uint64_t>(nj));
    return true;
  }

  // Launch thread: admits graphs of pending queries and launches batches
  // while streams and result slots are free; completions come from the
  // waiter threads.
  void worker_main() {
    dev_->bind_thread();
    std::unique_lock<std::mutex> lk(mu_);
    while (!stop_) {
      bool progress = false;
      // stop_ is checked before each step, under the lock: once it is set no
      // batch is launched, so every batch in flight has a waiter to finish it
      while (!stop_ && admit_pending(lk)) progress = true;
      while (!stop_ && launch_batch()) progress = true;
      if (!progress && !stop_) wait_at_most(cv_work_, lk, std::chrono::milliseconds(50));
    }
  }

  Device *dev_;
  acc_server_config cfg_;
  std::mutex mu_;
  std::mutex admit_mu_;  // serialises uploads
  std::condition_variable cv_work_, cv_done_, cv_batch_;
  std::map<int32_t, ServedGraph> graphs_;
  int32_t next_gid_ = 0;
  std::deque<uint64_t> pending_;
  std::unordered_map<uint64_t, Query> queries_;
  uint64_t next_ticket_ = 1;
  std::vector<SlotState> slots_;
  std::vector<Batch> batches_;
  std::thread worker_;
  bool stop_ = false;
  bool down_ = false;
  uint64_t clock_ = 0;
  uint64_t graph_bytes_ = 0;  // resident graphs
  uint64_t slot_bytes_ = 0;   // result slots
  acc_server_stats st_{};
};

}  // namespace accserve
