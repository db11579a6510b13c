// SCAN structural clustering on the resident counts (include/acc_cn.h,
// "SCAN queries").  Restates on the GPU the reference's post-processing:
//   similarity + core check   SCAN_XP::CheckCoreCompSimCore (scan-query/scan_xp.cpp:112-141)
//   core clustering           SCAN_XP::ClusterCore (:1915-1928) + UnionFind (util/union_find.h:4-36)
//   hubs                      SCAN_XP::CheckHub / LabelNonCore (:1930-1955)
//   counters                  SCAN_XP::PostProcess (:1957-1981)
//   cluster ids               SCAN_XP::MarkClusterMinEleAsId, REORDERING branch (:1983-2020)
//   non-core memberships      SCAN_XP::PrepareResultOutput (:2022-2049)
// The similarity decision is the exact integer form of the reference's
// long-double test: cn/sqrt((d_u+1)(d_v+1)) >= p/q  <=>  q^2 cn^2 >= p^2 (d_u+1)(d_v+1).
// Union-find links the larger root under the smaller one, so every root is
// the minimum vertex id of its cluster whatever the order of the unions, as
// in the reference.

#include <algorithm>
#include <new>

#include "acc_internal.h"

namespace {

constexpr int kScanWG = 256;

__device__ __forceinline__ int32_t uf_load(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; parent[y] <= y holds for every y, so the walk terminates.
__device__ __forceinline__ int32_t uf_find(int32_t *parent, int32_t x) {
  for (;;) {
    const int32_t p = uf_load(&parent[x]);
    if (p == x) return x;
    const int32_t gp = uf_load(&parent[p]);
    if (gp != p) __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
  }
}

// Links the sets of a and b, the larger root under the smaller.
__device__ __forceinline__ void uf_union(int32_t *parent, int32_t a, int32_t b) {
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a < b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    // a > b: parent[a] == a expected
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

__global__ __launch_bounds__(256) void k_fill_src(const uint32_t *__restrict__ off, uint32_t n,
                                                  int32_t *__restrict__ src) {
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * 256u + threadIdx.x) >> 6);
  const uint32_t nwaves = gridDim.x * 4u;
  const int lane = threadIdx.x & 63;
  for (uint32_t u = wave; u < n; u += nwaves) {
    const uint32_t b = off[u], e = off[u + 1];
    for (uint32_t i = b + lane; i < e; i += 64) src[i] = static_cast<int32_t>(u);
  }
}

__global__ __launch_bounds__(256) void k_scan_init(int32_t *__restrict__ core_cnt, uint32_t n,
                                                   unsigned long long *__restrict__ counters) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  for (uint32_t u = t; u < n; u += gridDim.x * 256u) core_cnt[u] = 0;
  if (t < 8) counters[t] = 0;
}

// Similarity of every slot and per-vertex counts of similar neighbors.  Each
// wave covers 64 consecutive slots; runs of equal source are reduced with a
// ballot and added with one atomic per run.
__global__ __launch_bounds__(kScanWG) void k_scan_sim(const uint32_t *__restrict__ off,
                                                      const int32_t *__restrict__ adj,
                                                      const int32_t *__restrict__ src,
                                                      const int32_t *__restrict__ cnt, uint64_t m, int32_t add,
                                                      unsigned long long p2, unsigned long long q2,
                                                      uint8_t *__restrict__ sim, int32_t *__restrict__ core_cnt) {
  const int lane = threadIdx.x & 63;
  const uint64_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * static_cast<uint64_t>(kScanWG) + threadIdx.x) >> 6);
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kScanWG / 64);
  for (uint64_t base = wave * 64; base < m; base += nwaves * 64) {
    const uint64_t i = base + lane;
    const bool valid = i < m;
    int32_t u = -1;
    bool s = false;
    if (valid) {
      u = src[i];
      const int32_t v = adj[i];
      const uint64_t c = static_cast<uint64_t>(cnt[i] + add);
      const uint64_t du1 = static_cast<uint64_t>(off[u + 1] - off[u]) + 1;
      const uint64_t dv1 = static_cast<uint64_t>(off[v + 1] - off[v]) + 1;
      const unsigned __int128 lhs = static_cast<unsigned __int128>(q2) * (c * c);
      const unsigned __int128 rhs = static_cast<unsigned __int128>(p2) * (du1 * dv1);
      s = lhs >= rhs;
      sim[i] = s ? 1 : 0;
    }
    const int32_t prev = __shfl_up(u, 1, 64);
    const bool head = valid && (lane == 0 || prev != u);
    const unsigned long long heads = __ballot(head);
    const unsigned long long sims = __ballot(s);
    if (head) {
      const unsigned long long above = heads & ~((2ull << lane) - 1);  // heads strictly above lane
      const int end = above ? __builtin_ctzll(above) : 64;
      const unsigned long long run = (end == 64 ? ~0ull : ((1ull << end) - 1)) & ~((1ull << lane) - 1);
      const int k = __popcll(sims & run);
      if (k) atomicAdd(&core_cnt[u], k);
    }
  }
}

__global__ __launch_bounds__(256) void k_scan_label(const int32_t *__restrict__ core_cnt, uint32_t n, int32_t mu,
                                                    int32_t *__restrict__ label, int32_t *__restrict__ parent,
                                                    int32_t *__restrict__ rmin, int32_t *__restrict__ rmax,
                                                    int32_t *__restrict__ cl) {
  for (uint32_t u = blockIdx.x * 256u + threadIdx.x; u < n; u += gridDim.x * 256u) {
    label[u] = core_cnt[u] >= mu ? 1 : 0;
    parent[u] = static_cast<int32_t>(u);
    rmin[u] = INT32_MAX;
    rmax[u] = -1;
    cl[u] = INT32_MAX;
  }
}

// ClusterCore: union over similar core-core edges (each undirected edge once).
__global__ __launch_bounds__(256) void k_scan_union(const int32_t *__restrict__ adj, const int32_t *__restrict__ src,
                                                    const uint8_t *__restrict__ sim, const int32_t *__restrict__ label,
                                                    uint64_t m, int32_t *parent) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < m; i += gridDim.x * 256ull) {
    const int32_t u = src[i];
    const int32_t v = adj[i];
    if (u < v && sim[i] && label[u] == 1 && label[v] == 1) uf_union(parent, u, v);
  }
}

__global__ __launch_bounds__(256) void k_scan_compress(const int32_t *__restrict__ label, uint32_t n,
                                                       int32_t *parent) {
  for (uint32_t u = blockIdx.x * 256u + threadIdx.x; u < n; u += gridDim.x * 256u)
    if (label[u] == 1)
      __hip_atomic_store(&parent[u], uf_find(parent, static_cast<int32_t>(u)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// CheckHub: per non-core vertex, the smallest and largest root among its core neighbors.
__global__ __launch_bounds__(256) void k_scan_hub(const int32_t *__restrict__ adj, const int32_t *__restrict__ src,
                                                  const int32_t *__restrict__ label,
                                                  const int32_t *__restrict__ parent, uint64_t m,
                                                  int32_t *__restrict__ rmin, int32_t *__restrict__ rmax) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < m; i += gridDim.x * 256ull) {
    const int32_t a = src[i];
    const int32_t v = adj[i];
    if (label[a] != 1 && label[v] == 1) {
      const int32_t r = parent[v];
      if (r < rmin[a]) atomicMin(&rmin[a], r);
      if (r > rmax[a]) atomicMax(&rmax[a], r);
    }
  }
}

// Labels hubs, takes the minimum old id of every cluster and counts
// cores, clusters, hubs and outliers (counters[0..3]).
__global__ __launch_bounds__(256) void k_scan_finalize(uint32_t n, int32_t *__restrict__ label,
                                                       const int32_t *__restrict__ parent,
                                                       const int32_t *__restrict__ rmin,
                                                       const int32_t *__restrict__ rmax,
                                                       const int32_t *__restrict__ old_id, int32_t *__restrict__ cl,
                                                       unsigned long long *__restrict__ counters) {
  unsigned long long cores = 0, clusters = 0, hubs = 0, outliers = 0;
  for (uint32_t u = blockIdx.x * 256u + threadIdx.x; u < n; u += gridDim.x * 256u) {
    if (label[u] == 1) {
      cores++;
      const int32_t r = parent[u];
      if (r == static_cast<int32_t>(u)) clusters++;
      const int32_t o = old_id ? old_id[u] : static_cast<int32_t>(u);
      atomicMin(&cl[r], o);
    } else if (rmax[u] >= 0 && rmin[u] != rmax[u]) {
      label[u] = 2;
      hubs++;
    } else {
      outliers++;
    }
  }
  __shared__ unsigned long long part[4][4];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cores += __shfl_xor(cores, o, 64);
    clusters += __shfl_xor(clusters, o, 64);
    hubs += __shfl_xor(hubs, o, 64);
    outliers += __shfl_xor(outliers, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    part[w][0] = cores;
    part[w][1] = clusters;
    part[w][2] = hubs;
    part[w][3] = outliers;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const unsigned long long t = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                 part[3][threadIdx.x];
    if (t) atomicAdd(&counters[threadIdx.x], t);
  }
}

__global__ __launch_bounds__(256) void k_scan_cid(uint32_t n, const int32_t *__restrict__ label,
                                                  const int32_t *__restrict__ parent,
                                                  const int32_t *__restrict__ cl, int32_t *__restrict__ cid) {
  for (uint32_t u = blockIdx.x * 256u + threadIdx.x; u < n; u += gridDim.x * 256u)
    cid[u] = label[u] == 1 ? cl[parent[u]] : -1;
}

// Non-core memberships: (cluster id of core u, old id of v) for every similar
// slot (u, v) with u core and v not core, emitted in any order; counters[4]
// counts them, capacity cap.  Every lane takes its position from the same
// counter with an increment of 0 or 1, so positions are unique without a
// lane-0 broadcast.
__global__ __launch_bounds__(256) void k_scan_pairs(const int32_t *__restrict__ adj, const int32_t *__restrict__ src,
                                                    const uint8_t *__restrict__ sim,
                                                    const int32_t *__restrict__ label,
                                                    const int32_t *__restrict__ cid,
                                                    const int32_t *__restrict__ old_id, uint64_t m,
                                                    int2 *__restrict__ pairs, uint64_t cap,
                                                    unsigned long long *__restrict__ counters) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < m; i += gridDim.x * 256ull) {
    if (!sim[i]) continue;
    const int32_t u = src[i];
    const int32_t v = adj[i];
    if (label[u] != 1 || label[v] == 1) continue;
    const unsigned long long pos = atomicAdd(&counters[4], 1ull);
    if (pos < cap) pairs[pos] = make_int2(cid[u], old_id ? old_id[v] : v);
  }
}

inline unsigned grid_for(uint64_t items, int cus) {
  const uint64_t want = (items + 255) / 256;
  const uint64_t c = static_cast<uint64_t>(cus > 0 ? cus : 256) * 8;
  return static_cast<unsigned>(want < c ? (want ? want : 1) : c);
}

}  // namespace

struct acc_scan_ws {
  acc_graph *g = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;  // counts of the graph are complete
  uint8_t *d_sim = nullptr;
  int32_t *d_core_cnt = nullptr;
  int32_t *d_label = nullptr;
  int32_t *d_parent = nullptr;
  int32_t *d_rmin = nullptr;
  int32_t *d_rmax = nullptr;
  int32_t *d_cl = nullptr;
  int32_t *d_cid = nullptr;
  unsigned long long *d_counters = nullptr;  // core, cluster, hub, outlier, pairs
  int2 *d_pairs = nullptr;
  uint64_t pairs_cap = 0;
  uint64_t device_bytes = 0;
  bool have_result = false;
};

namespace {

void free_ws(acc_scan_ws *w) {
  if (!w) return;
  (void)hipSetDevice(w->g->ctx->device);
  if (w->stream) (void)hipStreamSynchronize(w->stream);
  for (void *p : {static_cast<void *>(w->d_sim), static_cast<void *>(w->d_core_cnt),
                  static_cast<void *>(w->d_label), static_cast<void *>(w->d_parent), static_cast<void *>(w->d_rmin),
                  static_cast<void *>(w->d_rmax), static_cast<void *>(w->d_cl), static_cast<void *>(w->d_cid),
                  static_cast<void *>(w->d_counters), static_cast<void *>(w->d_pairs)})
    if (p) (void)hipFree(p);
  if (w->ready) (void)hipEventDestroy(w->ready);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
}

template <typename T>
int ws_alloc(acc_scan_ws *w, T **p, uint64_t count) {
  const size_t bytes = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return acc_set_error(ACC_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  w->device_bytes += bytes;
  return ACC_OK;
}

}  // namespace

// Builds the slot -> source array of a resident graph once.
int acc_graph_ensure_src(acc_graph *g) {
  if (g->d_src) return ACC_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&g->d_src), sizeof(int32_t) * (g->m ? g->m : 1));
  if (e != hipSuccess) {
    g->d_src = nullptr;
    return acc_set_error(ACC_ENOMEM, "hipMalloc of the slot source array failed");
  }
  g->device_bytes += sizeof(int32_t) * (g->m ? g->m : 1);
  hipLaunchKernelGGL(k_fill_src, dim3(grid_for(g->n, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream, g->d_off,
                     g->n, g->d_src);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

extern "C" {

int acc_graph_set_vertex_ids(acc_graph *g, const int32_t *old_id) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  (void)hipSetDevice(g->ctx->device);
  if (!old_id) {
    if (g->d_old) (void)hipFree(g->d_old);
    g->d_old = nullptr;
    return ACC_OK;
  }
  for (uint32_t u = 0; u < g->n; u++)
    if (old_id[u] < 0) return acc_set_error(ACC_EINVAL, "vertex ids must be non-negative");
  if (!g->d_old) {
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&g->d_old), sizeof(int32_t) * (g->n ? g->n : 1));
    if (e != hipSuccess) {
      g->d_old = nullptr;
      return acc_set_error(ACC_ENOMEM, "hipMalloc of the vertex id array failed");
    }
    g->device_bytes += sizeof(int32_t) * (g->n ? g->n : 1);
  }
  ACC_HIP_TRY(hipMemcpyAsync(g->d_old, old_id, sizeof(int32_t) * g->n, hipMemcpyHostToDevice, g->ctx->stream));
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  return ACC_OK;
}

int acc_scan_ws_create(acc_graph *g, acc_scan_ws **out) {
  if (!g || !out) return acc_set_error(ACC_EINVAL, "null argument");
  *out = nullptr;
  (void)hipSetDevice(g->ctx->device);
  auto *w = new (std::nothrow) acc_scan_ws;
  if (!w) return acc_set_error(ACC_ENOMEM, "host allocation failed");
  w->g = g;
  int rc;
  auto bail = [&](int code) {
    free_ws(w);
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(acc_set_error(ACC_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  e = hipEventCreateWithFlags(&w->ready, hipEventDisableTiming);
  if (e != hipSuccess) return bail(acc_set_error(ACC_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e)));
  if ((rc = ws_alloc(w, &w->d_sim, g->m))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_core_cnt, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_label, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_parent, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_rmin, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_rmax, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_cl, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_cid, g->n))) return bail(rc);
  if ((rc = ws_alloc(w, &w->d_counters, 8))) return bail(rc);
  if ((rc = acc_graph_ensure_src(g))) return bail(rc);
  *out = w;
  return ACC_OK;
}

void acc_scan_ws_destroy(acc_scan_ws *w) { free_ws(w); }

int acc_scan_ws_info(const acc_scan_ws *w, uint64_t *device_bytes, void **stream) {
  if (!w) return acc_set_error(ACC_EINVAL, "null workspace");
  if (device_bytes) *device_bytes = w->device_bytes;
  if (stream) *stream = w->stream;
  return ACC_OK;
}

int acc_scan_run(acc_scan_ws *w, int64_t eps_num, int64_t eps_den, int32_t mu, uint32_t flags) {
  if (!w) return acc_set_error(ACC_EINVAL, "null workspace");
  if (eps_num < 0 || eps_den <= 0 || eps_num > (1ll << 31) || eps_den > (1ll << 31))
    return acc_set_error(ACC_EINVAL, "eps must be eps_num/eps_den with 0 <= eps_num and 0 < eps_den <= 2^31");
  acc_graph *g = w->g;
  (void)hipSetDevice(g->ctx->device);
  if (!g->have_result) ACC_TRY(acc_cn_compute(g, ACC_PATH_AUTO, 0));
  // order this query after the graph's counts
  ACC_HIP_TRY(hipEventRecord(w->ready, g->ctx->stream));
  ACC_HIP_TRY(hipStreamWaitEvent(w->stream, w->ready, 0));
  const int cus = g->ctx->num_cus;
  hipStream_t s = w->stream;
  const unsigned long long p2 = static_cast<unsigned long long>(eps_num) * static_cast<unsigned long long>(eps_num);
  const unsigned long long q2 = static_cast<unsigned long long>(eps_den) * static_cast<unsigned long long>(eps_den);
  const int32_t add = g->cnt_plus2 ? 0 : 2;
  hipLaunchKernelGGL(k_scan_init, dim3(grid_for(g->n, cus)), dim3(256), 0, s, w->d_core_cnt, g->n, w->d_counters);
  hipLaunchKernelGGL(k_scan_sim, dim3(grid_for(g->m, cus)), dim3(kScanWG), 0, s, g->d_off, g->d_adj, g->d_src,
                     g->d_cnt, g->m, add, p2, q2, w->d_sim, w->d_core_cnt);
  hipLaunchKernelGGL(k_scan_label, dim3(grid_for(g->n, cus)), dim3(256), 0, s, w->d_core_cnt, g->n, mu, w->d_label,
                     w->d_parent, w->d_rmin, w->d_rmax, w->d_cl);
  hipLaunchKernelGGL(k_scan_union, dim3(grid_for(g->m, cus)), dim3(256), 0, s, g->d_adj, g->d_src, w->d_sim,
                     w->d_label, g->m, w->d_parent);
  hipLaunchKernelGGL(k_scan_compress, dim3(grid_for(g->n, cus)), dim3(256), 0, s, w->d_label, g->n, w->d_parent);
  hipLaunchKernelGGL(k_scan_hub, dim3(grid_for(g->m, cus)), dim3(256), 0, s, g->d_adj, g->d_src, w->d_label,
                     w->d_parent, g->m, w->d_rmin, w->d_rmax);
  hipLaunchKernelGGL(k_scan_finalize, dim3(grid_for(g->n, cus)), dim3(256), 0, s, g->n, w->d_label, w->d_parent,
                     w->d_rmin, w->d_rmax, g->d_old, w->d_cl, w->d_counters);
  hipLaunchKernelGGL(k_scan_cid, dim3(grid_for(g->n, cus)), dim3(256), 0, s, g->n, w->d_label, w->d_parent, w->d_cl,
                     w->d_cid);
  ACC_HIP_TRY(hipGetLastError());
  (void)flags;
  w->have_result = true;
  return ACC_OK;
}

int acc_scan_synchronize(acc_scan_ws *w) {
  if (!w) return acc_set_error(ACC_EINVAL, "null workspace");
  (void)hipSetDevice(w->g->ctx->device);
  ACC_HIP_TRY(hipStreamSynchronize(w->stream));
  return ACC_OK;
}

int acc_scan_results(acc_scan_ws *w, int32_t *label, int32_t *cluster_id, uint8_t *sim, int32_t *core_count,
                     int64_t *counts) {
  if (!w) return acc_set_error(ACC_EINVAL, "null workspace");
  if (!w->have_result) return acc_set_error(ACC_EINVAL, "no SCAN query has run on this workspace");
  acc_graph *g = w->g;
  (void)hipSetDevice(g->ctx->device);
  hipStream_t s = w->stream;
  if (label) ACC_HIP_TRY(hipMemcpyAsync(label, w->d_label, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, s));
  if (cluster_id) ACC_HIP_TRY(hipMemcpyAsync(cluster_id, w->d_cid, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, s));
  if (sim && g->m) ACC_HIP_TRY(hipMemcpyAsync(sim, w->d_sim, g->m, hipMemcpyDeviceToHost, s));
  if (core_count)
    ACC_HIP_TRY(hipMemcpyAsync(core_count, w->d_core_cnt, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, s));
  unsigned long long c[8] = {0};
  ACC_HIP_TRY(hipMemcpyAsync(c, w->d_counters, sizeof(c), hipMemcpyDeviceToHost, s));
  ACC_HIP_TRY(hipStreamSynchronize(s));
  if (counts)
    for (int i = 0; i < 4; i++) counts[i] = static_cast<int64_t>(c[i]);
  return ACC_OK;
}

int acc_scan_pairs(acc_scan_ws *w, int32_t *pairs, uint64_t cap, uint64_t *count) {
  if (!w || !count) return acc_set_error(ACC_EINVAL, "null argument");
  if (!w->have_result) return acc_set_error(ACC_EINVAL, "no SCAN query has run on this workspace");
  acc_graph *g = w->g;
  (void)hipSetDevice(g->ctx->device);
  hipStream_t s = w->stream;
  // every slot is at most one pair: size the buffer for m once
  if (!w->d_pairs) {
    int rc = ws_alloc(w, &w->d_pairs, g->m);
    if (rc) return rc;
    w->pairs_cap = g->m;
  }
  ACC_HIP_TRY(hipMemsetAsync(w->d_counters + 4, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_scan_pairs, dim3(grid_for(g->m, g->ctx->num_cus)), dim3(256), 0, s, g->d_adj, g->d_src,
                     w->d_sim, w->d_label, w->d_cid, g->d_old, g->m, w->d_pairs, w->pairs_cap, w->d_counters);
  ACC_HIP_TRY(hipGetLastError());
  unsigned long long k = 0;
  ACC_HIP_TRY(hipMemcpyAsync(&k, w->d_counters + 4, sizeof(k), hipMemcpyDeviceToHost, s));
  ACC_HIP_TRY(hipStreamSynchronize(s));
  *count = k;
  if (pairs && k) {
    const uint64_t take = std::min<uint64_t>(k, cap);
    ACC_HIP_TRY(hipMemcpy(pairs, w->d_pairs, sizeof(int2) * take, hipMemcpyDeviceToHost));
  }
  return ACC_OK;
}

}  // extern "C"
