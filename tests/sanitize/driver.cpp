// Host-only driver for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of the graph library (acctricnt_amd/csrc/graph_host.cpp) and the
// oracle (oracle/oracle.c): every entry point on small generated, written,
// read, converted, relabelled and malformed graphs.  Exits 0 when every
// result is as expected; a sanitizer report aborts the run.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "acc_graph.h"
#include "oracle.h"

static int failures = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                   \
    }                                                               \
  } while (0)

static void exercise(acc_hgraph *g, const std::string &dir) {
  const uint32_t n = acc_hgraph_n(g);
  const uint64_t m = acc_hgraph_m(g);
  const uint32_t *off = acc_hgraph_offsets(g);
  const int32_t *adj = acc_hgraph_adj(g);
  EXPECT(acc_csr_validate(n, m, off, adj) == ACC_OK);
  std::vector<int32_t> a(m ? m : 1), b(m ? m : 1), c(m ? m : 1, -1);
  oracle_cn_merge(n, m, off, adj, a.data(), 2);
  oracle_cn_bitmap(n, m, off, adj, b.data(), 3);
  for (uint64_t i = 0; i < m; i++) EXPECT(a[i] == b[i]);
  uint64_t t = oracle_triangle_count(n, off, adj, 2), tb = oracle_triangle_count_bitmap(n, off, adj, 3);
  EXPECT(t == tb);
  uint64_t s = 0;
  for (uint64_t i = 0; i < m; i++) s += a[i] - 2;
  EXPECT(s == 6 * t);
  std::vector<uint32_t> slots;
  std::vector<int32_t> srcs;
  for (uint32_t u = 0; u < n; u++)
    for (uint32_t j = off[u]; j < off[u + 1]; j++)
      if ((int32_t)u < adj[j] && j % 3 == 0) {
        slots.push_back(j);
        srcs.push_back((int32_t)u);
      }
  oracle_cn_merge_slots(n, off, adj, c.data(), slots.data(), srcs.data(), slots.size(), 2);
  for (uint32_t j : slots) EXPECT(c[j] == a[j]);
  std::vector<int32_t> r(m ? m : 1, 2);
  oracle_cn_merge_range(n, off, adj, r.data(), 0, n / 2, 2);
  oracle_cn_merge_range(n, off, adj, r.data(), n / 2, n, 2);
  for (uint64_t i = 0; i < m; i++) EXPECT(r[i] == a[i]);
  uint64_t bad = 1, first = 0;
  EXPECT(acc_csr_verify_counts(n, m, off, adj, a.data(), 2, 2, &bad, &first) == ACC_OK);
  EXPECT(bad == 0 && first == m);
  std::vector<uint32_t> rev(m ? m : 1);
  EXPECT(acc_hgraph_reverse_slots(g, rev.data()) == ACC_OK);
  for (uint64_t i = 0; i < m; i++) EXPECT(a[rev[i]] == a[i]);
  // SCAN in both decision forms, clusters, result files
  std::vector<uint8_t> sim(m ? m : 1), sim2(m ? m : 1);
  std::vector<int32_t> cc(n ? n : 1), lab(n ? n : 1), lab2(n ? n : 1), root(n ? n : 1), cid(n ? n : 1);
  oracle_scan_core_int(n, off, adj, a.data(), 3, 10, 2, sim.data(), cc.data(), lab.data());
  oracle_scan_core_fp(n, off, adj, a.data(), 0.3, 2, sim2.data(), cc.data(), lab2.data());
  for (uint64_t i = 0; i < m; i++) EXPECT(sim[i] == sim2[i]);
  int64_t k[4];
  EXPECT(oracle_scan_cluster(n, off, adj, sim.data(), lab.data(), nullptr, root.data(), cid.data(), &k[0], &k[1],
                             &k[2], &k[3]) == 0);
  EXPECT(k[0] + k[2] + k[3] == (int64_t)n);
  const std::string res = dir + "/oracle-result.txt";
  EXPECT(oracle_scan_write_result(res.c_str(), n, off, adj, sim.data(), lab.data(), cid.data(), nullptr) == 0);
  std::vector<int32_t> pairs;
  for (uint32_t u = 0; u < n; u++)
    for (uint32_t j = off[u]; j < off[u + 1]; j++)
      if (lab[u] == 1 && lab[adj[j]] != 1 && sim[j]) {
        pairs.push_back(cid[u]);
        pairs.push_back(adj[j]);
      }
  const std::string res2 = dir + "/lib-result.txt";
  EXPECT(acc_write_scan_result(res2.c_str(), n, lab.data(), cid.data(), nullptr, pairs.data(), pairs.size() / 2) ==
         ACC_OK);
  char path[512];
  EXPECT(acc_scan_result_path(dir.c_str(), 0.3, 2, path, sizeof(path)) == ACC_OK);
  EXPECT(acc_scan_result_path(dir.c_str(), 0.3, 2, path, 4) == ACC_EINVAL);
  char eps[64];
  oracle_eps_string(0.3, eps);
  EXPECT(std::strcmp(eps, "0.3") == 0);
  // file round trip and relabel
  EXPECT(acc_hgraph_write_bins(g, dir.c_str()) == ACC_OK);
  acc_hgraph *h = nullptr;
  EXPECT(acc_hgraph_read_bins(dir.c_str(), &h) == ACC_OK);
  if (h) {
    EXPECT(acc_hgraph_n(h) == n && acc_hgraph_m(h) == m);
    std::vector<int32_t> old(n ? n : 1);
    EXPECT(acc_hgraph_relabel_degree(h, old.data()) == ACC_OK);
    EXPECT(acc_csr_validate(acc_hgraph_n(h), acc_hgraph_m(h), acc_hgraph_offsets(h), acc_hgraph_adj(h)) == ACC_OK);
    EXPECT(oracle_triangle_count(acc_hgraph_n(h), acc_hgraph_offsets(h), acc_hgraph_adj(h), 2) == t);
    acc_hgraph_free(h);
  }
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  acc_hgraph *g = nullptr;
  EXPECT(acc_gen_rmat(1 << 11, 11, 16u << 11, 0.57, 0.19, 0.19, 1, &g) == ACC_OK);
  if (g) exercise(g, dir), acc_hgraph_free(g);
  EXPECT(acc_gen_er(3000, 20000, 2, &g) == ACC_OK);
  if (g) exercise(g, dir), acc_hgraph_free(g);
  EXPECT(acc_gen_chung_lu(4000, 2.1, 12, 500, 3, &g) == ACC_OK);
  if (g) exercise(g, dir), acc_hgraph_free(g);
  EXPECT(acc_gen_rmat(1500, 11, 4000, 0.57, 0.19, 0.19, 4, &g) == ACC_OK);  // n not a power of two
  if (g) exercise(g, dir), acc_hgraph_free(g);
  // edge lists with self-loops, duplicates and both directions; an empty graph
  const int32_t s[] = {0, 1, 2, 2, 3, 1, 4, 0, 2};
  const int32_t d[] = {1, 2, 0, 2, 1, 2, 4, 3, 3};
  EXPECT(acc_hgraph_from_edges(6, 9, s, d, &g) == ACC_OK);
  if (g) exercise(g, dir), acc_hgraph_free(g);
  g = nullptr;
  EXPECT(acc_hgraph_from_edges(0, 0, nullptr, nullptr, &g) == ACC_EINVAL && g == nullptr);  // no vertices
  EXPECT(acc_hgraph_from_edges(1, 0, nullptr, nullptr, &g) == ACC_OK);  // one isolated vertex
  if (g) exercise(g, dir), acc_hgraph_free(g);
  g = nullptr;
  EXPECT(acc_hgraph_from_edges(3, 9, s, d, &g) == ACC_EINVAL && g == nullptr);  // endpoints out of range
  // malformed CSRs are rejected
  const uint32_t off[] = {0, 2, 4, 6};
  const int32_t loop[] = {0, 1, 0, 2, 1, 2}, unsorted[] = {2, 1, 0, 2, 0, 1}, asym[] = {1, 2, 0, 2, 0, 0},
                range[] = {1, 7, 0, 2, 0, 1};
  for (const int32_t *a : {loop, unsorted, asym, range}) EXPECT(acc_csr_validate(3, 6, off, a) == ACC_EFORMAT);
  g = nullptr;
  EXPECT(acc_hgraph_from_csr(3, 6, off, range, &g) == ACC_EFORMAT && g == nullptr);
  // converter on a text edge list
  const std::string txt = dir + "/edges.txt";
  FILE *f = std::fopen(txt.c_str(), "w");
  if (f) {
    std::fputs("# comment\n0 1\n1 2\n2 0\n3 3\n1 2\n0 3\n", f);
    std::fclose(f);
  }
  EXPECT(acc_convert_edge_list(txt.c_str(), dir.c_str()) == ACC_OK);
  g = nullptr;
  EXPECT(acc_hgraph_read_bins(dir.c_str(), &g) == ACC_OK);
  if (g) {
    EXPECT(acc_hgraph_n(g) == 4 && acc_hgraph_m(g) == 6);
    acc_hgraph_free(g);
  }
  g = nullptr;
  EXPECT(acc_hgraph_read_bins((dir + "/missing").c_str(), &g) == ACC_EIO && g == nullptr);
  std::printf("sanitized driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
