/*
 * oracle.c -- CPU restatement of the reference's all-edge common-neighbor
 * counting, triangle counting and SCAN post-processing.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP path and the timed
 * CPU baseline; nothing in the product library links or calls it.
 *
 * Pinning: every function here is checked in tests/test_oracle.py against
 * closed forms (cliques, wheels, windmills, bipartite graphs, cycles, stars),
 * against scipy.sparse ((A·A)∘A and trace(A^3)/6), against brute-force set
 * intersection and a brute-force SCAN written from the definitions, and
 * against each other (merge == bitmap, sum of counts == 6T, integer
 * similarity == long-double similarity).  oracle_cn_merge_slots is compared
 * with scipy on every u<v slot and on random subsets (only the chosen slots
 * and their reverse slots written); both triangle counts with closed forms,
 * scipy's trace and the golden totals.  The reference's own build is not
 * available here: its set-intersection and SCAN translation units include
 * the un-vendored sparsepp header (scan-query/set-inter/csr_set_intersection_serial.h,
 * scan-query/scan_xp.cpp:7), so no _ref build exists to compare with.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_CORE 1
#define ORACLE_HUB 2
#define CHUNK_SLOTS 6000 /* schedule(dynamic, 6000), scan-query/scan_xp.cpp:684 */

/* ComputeCNNaiveStdMerge(Graph*, off_u, end_u, off_v, end_v),
 * scan-query/set-inter/csr_set_intersection_serial.cpp:137-165 */
static int std_merge(const int32_t *adj, uint64_t a, uint64_t ae, uint64_t b, uint64_t be) {
  int c = 0;
  if (a >= ae || b >= be) return 0;
  for (;;) {
    if (adj[a] < adj[b]) {
      if (++a >= ae) return c;
    } else if (adj[a] > adj[b]) {
      if (++b >= be) return c;
    } else {
      c++;
      ++a;
      ++b;
      if (a >= ae || b >= be) return c;
    }
  }
}

/* ComputeCNNaive(Graph*, off_u, end_u, off_v, end_v),
 * scan-query/set-inter/csr_set_intersection_serial.cpp:77-106 */
static int naive_merge(const int32_t *adj, uint64_t a, uint64_t ae, uint64_t b, uint64_t be) {
  int c = 0;
  if (a >= ae || b >= be) return 0;
  for (;;) {
    while (adj[a] < adj[b])
      if (++a >= ae) return c;
    while (adj[a] > adj[b])
      if (++b >= be) return c;
    if (adj[a] == adj[b]) {
      c++;
      ++a;
      ++b;
      if (a >= ae || b >= be) return c;
    }
  }
}

/* First position in adj[lo, hi) whose value is >= val. */
static uint64_t lower_bound_i32(const int32_t *adj, uint64_t lo, uint64_t hi, int32_t val) {
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (adj[mid] < val)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* Source vertex of slot i: the u with off[u] <= i < off[u+1] (the result of
 * SCAN_XP::FindSrc, scan-query/scan_xp.cpp:94-110, found here by binary
 * search so it does not depend on per-thread history). */
static uint32_t src_of_slot(const uint32_t *off, uint32_t n, uint64_t i) {
  uint32_t lo = 0, hi = n; /* invariant: off[lo] <= i < off[hi] */
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  while (i >= off[lo + 1]) lo++;
  return lo;
}

static int threads_or_default(int nthreads) { return nthreads > 0 ? nthreads : omp_get_max_threads(); }

void oracle_cn_merge_range(uint32_t n, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                           uint32_t u_lo, uint32_t u_hi, int nthreads) {
  if (u_hi > n) u_hi = n;
  if (u_lo >= u_hi) return;
  const uint64_t s_lo = off[u_lo], s_hi = off[u_hi];
  const int64_t nchunks = (int64_t)((s_hi - s_lo + CHUNK_SLOTS - 1) / CHUNK_SLOTS);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads_or_default(nthreads))
  for (int64_t ch = 0; ch < nchunks; ch++) {
    uint64_t i0 = s_lo + (uint64_t)ch * CHUNK_SLOTS;
    uint64_t i1 = i0 + CHUNK_SLOTS < s_hi ? i0 + CHUNK_SLOTS : s_hi;
    uint32_t u = src_of_slot(off, n, i0);
    for (uint64_t i = i0; i < i1; i++) {
      while (i >= off[u + 1]) u++;
      int32_t v = adj[i];
      if ((int32_t)u < v) {
        int32_t c = 2 + std_merge(adj, off[u], off[u + 1], off[v], off[v + 1]);
        cnt[i] = c;
        cnt[lower_bound_i32(adj, off[v], off[v + 1], (int32_t)u)] = c;
      }
    }
  }
}

void oracle_cn_merge_slots(uint32_t n, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                           const uint32_t *slots, const int32_t *srcs, uint64_t k, int nthreads) {
  (void)n;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads_or_default(nthreads))
  for (int64_t t = 0; t < (int64_t)k; t++) {
    const uint64_t i = slots[t];
    const int32_t u = srcs[t], v = adj[i];
    if (u < v) {
      int32_t c = 2 + std_merge(adj, off[u], off[u + 1], off[v], off[v + 1]);
      cnt[i] = c;
      cnt[lower_bound_i32(adj, off[v], off[v + 1], u)] = c;
    }
  }
}

void oracle_cn_merge(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                     int nthreads) {
  /* constructor initialisation to 2, scan-query/scan_xp.cpp:72-81 */
#pragma omp parallel for schedule(static) num_threads(threads_or_default(nthreads))
  for (int64_t i = 0; i < (int64_t)m; i++) cnt[i] = 2;
  oracle_cn_merge_range(n, off, adj, cnt, 0, n, nthreads);
}

void oracle_cn_bitmap(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                      int nthreads) {
  const int nt = threads_or_default(nthreads);
  const uint64_t words = ((uint64_t)n + 63) / 64;
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)m; i++) cnt[i] = 2;
  const int64_t nchunks = (int64_t)((m + CHUNK_SLOTS - 1) / CHUNK_SLOTS);
#pragma omp parallel num_threads(nt)
  {
    uint64_t *bits = (uint64_t *)calloc(words ? words : 1, sizeof(uint64_t));
    int64_t last_u = -1;
#pragma omp for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nchunks; ch++) {
      uint64_t i0 = (uint64_t)ch * CHUNK_SLOTS;
      uint64_t i1 = i0 + CHUNK_SLOTS < m ? i0 + CHUNK_SLOTS : m;
      uint32_t u = src_of_slot(off, n, i0);
      for (uint64_t i = i0; i < i1; i++) {
        while (i >= off[u + 1]) u++;
        if ((int64_t)u != last_u) {
          if (last_u >= 0)
            for (uint64_t j = off[last_u]; j < off[last_u + 1]; j++) bits[adj[j] >> 6] = 0;
          for (uint64_t j = off[u]; j < off[u + 1]; j++) bits[adj[j] >> 6] |= 1ULL << (adj[j] & 63);
          last_u = u;
        }
        int32_t v = adj[i];
        uint32_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
        if (du > dv || (du == dv && (int32_t)u < v)) {
          int32_t c = 2;
          for (uint64_t j = off[v]; j < off[v + 1]; j++) c += (int32_t)((bits[adj[j] >> 6] >> (adj[j] & 63)) & 1);
          cnt[i] = c;
          cnt[lower_bound_i32(adj, off[v], off[v + 1], (int32_t)u)] = c;
        }
      }
    }
    free(bits);
  }
}

uint64_t oracle_triangle_count(uint32_t n, const uint32_t *off, const int32_t *adj, int nthreads) {
  const int nt = threads_or_default(nthreads);
  uint32_t *end = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t u = 0; u < (int64_t)n; u++) end[u] = (uint32_t)lower_bound_i32(adj, off[u], off[u + 1], (int32_t)u);
  uint64_t t = 0;
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : t) num_threads(nt)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    for (uint64_t j = off[u]; j < end[u]; j++) {
      int32_t v = adj[j];
      t += (uint64_t)naive_merge(adj, off[u], end[u], off[v], end[v]);
    }
  }
  free(end);
  return t;
}

uint64_t oracle_triangle_count_bitmap(uint32_t n, const uint32_t *off, const int32_t *adj, int nthreads) {
  const int nt = threads_or_default(nthreads);
  const uint64_t words = ((uint64_t)n + 63) / 64;
  /* node_off_end: end of N<(u) (scan_xp_triangle_cnt.cpp:669-673) */
  uint32_t *end = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t u = 0; u < (int64_t)n; u++) end[u] = (uint32_t)lower_bound_i32(adj, off[u], off[u + 1], (int32_t)u);
  uint64_t t = 0;
#pragma omp parallel num_threads(nt) reduction(+ : t)
  {
    uint64_t *bits = (uint64_t *)calloc(words ? words : 1, sizeof(uint64_t));
#pragma omp for schedule(dynamic, 100)
    for (int64_t u = 0; u < (int64_t)n; u++) {
      /* set N<(u) (:683-689), probe N<(v) for every v in N<(u) (:691-698), clear (:701-705) */
      for (uint64_t j = off[u]; j < end[u]; j++) bits[adj[j] >> 6] |= 1ULL << (adj[j] & 63);
      for (uint64_t j = off[u]; j < end[u]; j++) {
        const int32_t v = adj[j];
        for (uint64_t k = off[v]; k < end[v]; k++) t += (bits[adj[k] >> 6] >> (adj[k] & 63)) & 1;
      }
      for (uint64_t j = off[u]; j < end[u]; j++) bits[adj[j] >> 6] = 0;
    }
    free(bits);
  }
  free(end);
  return t;
}

static void core_finish(uint32_t n, const int32_t *core_count, int mu, int32_t *label) {
  for (uint32_t u = 0; u < n; u++) label[u] = core_count[u] >= mu ? ORACLE_CORE : 0;
}

void oracle_scan_core_fp(uint32_t n, const uint32_t *off, const int32_t *adj, const int32_t *cnt,
                         double eps, int mu, uint8_t *sim, int32_t *core_count, int32_t *label) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    int32_t cc = 0;
    long double du = (long double)(off[u + 1] - off[u]) + 1;
    for (uint64_t i = off[u]; i < off[u + 1]; i++) {
      int32_t v = adj[i];
      long double dv = (long double)(off[v + 1] - off[v]) + 1;
      double s = (double)((long double)cnt[i] / sqrtl(du * dv));
      sim[i] = s >= eps;
      cc += sim[i];
    }
    core_count[u] = cc;
  }
  core_finish(n, core_count, mu, label);
}

void oracle_scan_core_int(uint32_t n, const uint32_t *off, const int32_t *adj, const int32_t *cnt,
                          int64_t eps_num, int64_t eps_den, int mu, uint8_t *sim, int32_t *core_count,
                          int32_t *label) {
  const unsigned __int128 p2 = (unsigned __int128)((uint64_t)eps_num * (uint64_t)eps_num);
  const unsigned __int128 q2 = (unsigned __int128)((uint64_t)eps_den * (uint64_t)eps_den);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    int32_t cc = 0;
    uint64_t du = (uint64_t)(off[u + 1] - off[u]) + 1;
    for (uint64_t i = off[u]; i < off[u + 1]; i++) {
      int32_t v = adj[i];
      uint64_t dv = (uint64_t)(off[v + 1] - off[v]) + 1;
      uint64_t c = (uint64_t)cnt[i];
      unsigned __int128 lhs = q2 * (unsigned __int128)(c * c);
      unsigned __int128 rhs = p2 * (unsigned __int128)(du * dv);
      sim[i] = lhs >= rhs;
      cc += sim[i];
    }
    core_count[u] = cc;
  }
  core_finish(n, core_count, mu, label);
}

/* UnionFind::FindRoot with path compression, scan-query/util/union_find.h:17-19 */
static int32_t uf_find(int32_t *parent, int32_t x) {
  int32_t r = x;
  while (parent[r] != r) r = parent[r];
  while (parent[x] != r) {
    int32_t nx = parent[x];
    parent[x] = r;
    x = nx;
  }
  return r;
}

/* UnionFind::UnionThreadSafe, union_find.h:21-34: the larger root is linked
 * under the smaller one, so every root is the minimum id of its set. */
static void uf_union(int32_t *parent, int32_t a, int32_t b) {
  int32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
  if (ra == rb) return;
  if (ra < rb)
    parent[rb] = ra;
  else
    parent[ra] = rb;
}

int oracle_scan_cluster(uint32_t n, const uint32_t *off, const int32_t *adj, const uint8_t *sim,
                        int32_t *label, const int32_t *old_id, int32_t *root, int32_t *cluster_id,
                        int64_t *core_num, int64_t *cluster_num, int64_t *hub_num, int64_t *outlier_num) {
  int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) parent[i] = (int32_t)i;
  /* ClusterCore, scan-query/scan_xp.cpp:1915-1928 */
  for (uint32_t u = 0; u < n; u++) {
    if (label[u] != ORACLE_CORE) continue;
    for (uint64_t j = off[u]; j < off[u + 1]; j++)
      if (label[adj[j]] == ORACLE_CORE && sim[j]) uf_union(parent, (int32_t)u, adj[j]);
  }
  for (uint32_t u = 0; u < n; u++) root[u] = label[u] == ORACLE_CORE ? uf_find(parent, (int32_t)u) : -1;
  /* LabelNonCore + CheckHub, scan-query/scan_xp.cpp:1930-1955: a non-core
   * vertex whose CORE neighbours lie in two or more distinct sets is a hub. */
  int64_t cores = 0, hubs = 0, outliers = 0, clusters = 0;
  int32_t *new_label = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
  for (uint32_t a = 0; a < n; a++) {
    new_label[a] = label[a];
    if (label[a] == ORACLE_CORE) {
      cores++;
      continue;
    }
    int32_t first = -1;
    int hub = 0;
    for (uint64_t j = off[a]; j < off[a + 1] && !hub; j++) {
      int32_t v = adj[j];
      if (label[v] != ORACLE_CORE) continue;
      if (first < 0)
        first = root[v];
      else if (root[v] != first)
        hub = 1;
    }
    if (hub) new_label[a] = ORACLE_HUB;
  }
  memcpy(label, new_label, sizeof(int32_t) * n);
  free(new_label);
  /* PostProcess counters, scan-query/scan_xp.cpp:1957-1981 */
  for (uint32_t u = 0; u < n; u++) {
    if (label[u] == ORACLE_CORE && root[u] == (int32_t)u) clusters++;
    if (label[u] == ORACLE_HUB) hubs++;
    if (label[u] == 0) outliers++;
  }
  /* MarkClusterMinEleAsId (REORDERING branch), scan-query/scan_xp.cpp:1983-2020:
   * cluster id = minimum old id among the cores of the set. */
  int32_t *cl = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) cl[i] = (int32_t)n;
  for (uint32_t u = 0; u < n; u++) {
    if (label[u] != ORACLE_CORE) continue;
    int32_t o = old_id ? old_id[u] : (int32_t)u;
    if (o < cl[root[u]]) cl[root[u]] = o;
  }
  for (uint32_t u = 0; u < n; u++) cluster_id[u] = label[u] == ORACLE_CORE ? cl[root[u]] : -1;
  free(cl);
  free(parent);
  if (core_num) *core_num = cores;
  if (cluster_num) *cluster_num = clusters;
  if (hub_num) *hub_num = hubs;
  if (outlier_num) *outlier_num = outliers;
  return 0;
}

typedef struct {
  int32_t a, b;
} pair_i32;

static int cmp_pair(const void *x, const void *y) {
  const pair_i32 *p = (const pair_i32 *)x, *q = (const pair_i32 *)y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  return 0;
}

int oracle_scan_write_result(const char *path, uint32_t n, const uint32_t *off, const int32_t *adj,
                             const uint8_t *sim, const int32_t *label, const int32_t *cluster_id,
                             const int32_t *old_id) {
  FILE *f = fopen(path, "w");
  if (!f) return -4;
  fputs("c/n vertex_id cluster_id\n", f);
  /* cores: (old id, cluster id) sorted by old id, scan-query/util/graph.cpp:185-199 */
  size_t nc = 0;
  for (uint32_t u = 0; u < n; u++) nc += label[u] == ORACLE_CORE;
  pair_i32 *cores = (pair_i32 *)malloc(sizeof(pair_i32) * (nc ? nc : 1));
  size_t k = 0;
  for (uint32_t u = 0; u < n; u++)
    if (label[u] == ORACLE_CORE) {
      cores[k].a = old_id ? old_id[u] : (int32_t)u;
      cores[k].b = cluster_id[u];
      k++;
    }
  qsort(cores, nc, sizeof(pair_i32), cmp_pair);
  for (size_t i = 0; i < nc; i++) fprintf(f, "c %d %d\n", cores[i].a, cores[i].b);
  free(cores);
  /* non-cores similar to a core: sorted unique (cluster id, old id),
   * scan-query/scan_xp.cpp:2029-2049, util/graph.cpp:201-213 */
  size_t np = 0;
  for (uint32_t u = 0; u < n; u++) {
    if (label[u] != ORACLE_CORE) continue;
    for (uint64_t j = off[u]; j < off[u + 1]; j++)
      if (label[adj[j]] != ORACLE_CORE && sim[j]) np++;
  }
  pair_i32 *pairs = (pair_i32 *)malloc(sizeof(pair_i32) * (np ? np : 1));
  k = 0;
  for (uint32_t u = 0; u < n; u++) {
    if (label[u] != ORACLE_CORE) continue;
    for (uint64_t j = off[u]; j < off[u + 1]; j++) {
      int32_t v = adj[j];
      if (label[v] != ORACLE_CORE && sim[j]) {
        pairs[k].a = cluster_id[u];
        pairs[k].b = old_id ? old_id[v] : v;
        k++;
      }
    }
  }
  qsort(pairs, np, sizeof(pair_i32), cmp_pair);
  for (size_t i = 0; i < np; i++) {
    if (i > 0 && pairs[i].a == pairs[i - 1].a && pairs[i].b == pairs[i - 1].b) continue;
    fprintf(f, "n %d %d\n", pairs[i].b, pairs[i].a);
  }
  free(pairs);
  return fclose(f) == 0 ? 0 : -4;
}

int oracle_threads(void) { return omp_get_max_threads(); }

void oracle_eps_string(double eps, char *buf) {
  snprintf(buf, 64, "%f", eps);
  size_t len = strlen(buf);
  while (len > 0 && (buf[len - 1] == '0' || buf[len - 1] == 'u')) buf[--len] = '\0';
}
