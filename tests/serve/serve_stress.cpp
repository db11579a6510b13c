// Concurrent clients of the serving core over the fake device of
// serve_driver.cpp, for a ThreadSanitizer run (tests/test_sanitize.py).
#include <atomic>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "acc_cn.h"

extern "C" {
void *fake_create(uint64_t budget, uint64_t phys_limit, int streams, int slots, int max_batch, int gated);
void fake_destroy(void *h);
int fake_add_graph(void *h, uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, int32_t *gid);
int fake_remove_graph(void *h, int32_t gid);
int fake_submit(void *h, int32_t gid, int64_t p, int64_t q, int32_t mu, uint64_t *ticket);
int fake_wait(void *h, uint64_t ticket, int32_t timeout_ms);
int fake_result(void *h, uint64_t ticket, int64_t *counts, int8_t *label, int32_t *cid, uint32_t cap,
                double *latency_ms);
int fake_release(void *h, uint64_t ticket);
void fake_stats(void *h, acc_server_stats *out);
void fake_advance(void *h, int64_t ms);
void fake_fail(void *h, int uploads, int launches, int waits);
}

int main() {
  const uint32_t n = 32;
  std::vector<uint32_t> off(n + 1, 0);
  std::vector<int32_t> adj(1, 0);
  const uint64_t graph = 1000 + 4ull * n;  // the fake device's estimate for (32, 0)
  void *h = fake_create(3 * graph + 8 * 132 + 64, 1ull << 40, 4, 8, 6, 0);
  std::vector<int32_t> gids(6);
  for (int32_t &g : gids) fake_add_graph(h, n, 0, off.data(), adj.data(), &g);
  std::atomic<int> bad{0}, done{0}, stop{0};
  std::vector<std::thread> clients;
  for (int c = 0; c < 6; c++) {
    clients.emplace_back([&, c]() {
      std::mt19937 rng(c + 1);
      for (int i = 0; i < 300; i++) {
        const int32_t g = gids[rng() % gids.size()];
        const int64_t p = rng() % 10, q = 10;
        const int32_t mu = 1 + rng() % 4;
        uint64_t t = 0;
        if (fake_submit(h, g, p, q, mu, &t) != ACC_OK) continue;  // the graph may be gone
        const int how = rng() % 4;
        if (how == 0) {
          fake_release(h, t);  // cancels it if still pending; refused while running
          if (fake_wait(h, t, -1) == ACC_ENOTFOUND) continue;
        } else if (how == 1) {
          fake_wait(h, t, 5);
        }
        const int rc = fake_wait(h, t, -1);
        if (rc == ACC_OK) {
          int64_t counts[4];
          int8_t lab[32];
          int32_t cid[32];
          double lat = 0;
          if (fake_result(h, t, counts, lab, cid, 32, &lat) != ACC_OK || counts[1] != p || counts[3] != mu) bad++;
          done++;
        }
        fake_release(h, t);
      }
    });
  }
  std::thread admin([&]() {
    std::mt19937 rng(99);
    while (!stop.load()) {
      int32_t g = -1;
      if (fake_add_graph(h, n, 0, off.data(), adj.data(), &g) == ACC_OK) fake_remove_graph(h, g);
      if (rng() % 8 == 0) fake_fail(h, 1, 0, rng() % 2);
      fake_advance(h, 1);
      std::this_thread::yield();
    }
  });
  for (auto &t : clients) t.join();
  stop = 1;
  admin.join();
  acc_server_stats st{};
  fake_stats(h, &st);
  fake_destroy(h);
  std::printf("completed %d, evictions %llu, admissions %llu, failed %llu, cancelled %llu\n", done.load(),
              static_cast<unsigned long long>(st.evictions), static_cast<unsigned long long>(st.admissions),
              static_cast<unsigned long long>(st.failed), static_cast<unsigned long long>(st.cancelled));
  if (bad.load() || done.load() == 0) {
    std::printf("stress failed: %d wrong results\n", bad.load());
    return 1;
  }
  std::printf("stress ok\n");
  return 0;
}
