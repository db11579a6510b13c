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
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "acc_internal.h"

namespace {

constexpr int kScanWG = 256;

__device__ __forceinline__ uint32_t wave_excl_scan32(uint32_t v, int lane) {
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

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

// Root of x without writes.  Used once the unions are done: a path-halving
// store racing with another thread's store of a root could leave a pointer
// to a non-root ancestor behind.
__device__ __forceinline__ int32_t uf_root(const int32_t *parent, int32_t x) {
  for (;;) {
    const int32_t p = uf_load(&parent[x]);
    if (p == x) return x;
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

// Unites a with every lane's v where `valid`, a being the same on every lane
// (a wave-uniform call): the roots involved are linked under their minimum
// m in one step each -- distinct roots, so the CASes rarely collide, where 64
// separate unions through a's root would queue on it.  A root that stopped
// being one meanwhile takes the general union.  Larger roots still go under
// smaller ones.
__device__ __forceinline__ void uf_union_wave(int32_t *parent, int32_t a, int32_t v, bool valid) {
  const int32_t ra = uf_find(parent, a);
  const int32_t rv = valid ? uf_find(parent, v) : INT32_MAX;
  int32_t mn = rv < ra ? rv : ra;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t t = __shfl_xor(mn, o, 64);
    mn = t < mn ? t : mn;
  }
  if (valid && rv != mn && atomicCAS(&parent[rv], rv, mn) != rv) uf_union(parent, rv, mn);
  if ((threadIdx.x & 63) == 0 && ra != mn && atomicCAS(&parent[ra], ra, mn) != ra) uf_union(parent, ra, mn);
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

// Every SCAN kernel below runs a table of jobs (one job = one (graph, eps,
// mu) query) in one launch: job j owns the workgroups [sb0, sb0 + sbn) of a
// slot-sized launch and [vb0, vb0 + vbn) of a vertex-sized one, and those
// workgroups stride over its m slots or n vertices.  The job of a workgroup
// is found by a binary search of the table with workgroup-uniform values.
__device__ __forceinline__ ScanJob job_of(const JobTable &jt, bool slots, uint32_t *local, uint32_t *count) {
  const uint32_t blk = blockIdx.x;
  int lo = 0, hi = jt.nj;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((slots ? jt.jobs[mid].sb0 : jt.jobs[mid].vb0) <= blk) lo = mid; else hi = mid;
  }
  const ScanJob J = jt.jobs[lo];
  *local = blk - (slots ? J.sb0 : J.vb0);
  *count = slots ? J.sbn : J.vbn;
  return J;
}

__global__ __launch_bounds__(256) void k_sj_init(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  const uint32_t t = lb * 256u + threadIdx.x;
  for (uint32_t u = t; u < J.n; u += nb * 256u) J.core_cnt[u] = 0;
  if (t < kScanCounters) J.counters[t] = 0;
}

// Similarity of every slot and per-vertex counts of similar neighbors.  Each
// wave covers 64 consecutive slots; runs of equal source are reduced with a
// ballot and added with one atomic per run.
__global__ __launch_bounds__(kScanWG) void k_sj_sim(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, true, &lb, &nb);
  const uint32_t *__restrict__ off = J.off;
  const int32_t *__restrict__ adj = J.adj;
  const int32_t *__restrict__ src = J.src;
  const int32_t *__restrict__ cnt = J.cnt;
  const uint64_t m = J.m;
  const int32_t add = J.add;
  const unsigned long long p2 = J.p2, q2 = J.q2;
  const int lane = threadIdx.x & 63;
  const uint64_t wave = __builtin_amdgcn_readfirstlane((static_cast<uint64_t>(lb) * kScanWG + threadIdx.x) >> 6);
  const uint64_t nwaves = static_cast<uint64_t>(nb) * (kScanWG / 64);
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
      J.sim[i] = s ? 1 : 0;
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
      if (k) atomicAdd(&J.core_cnt[u], k);
    }
  }
}

__global__ __launch_bounds__(256) void k_sj_label(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  for (uint32_t u = lb * 256u + threadIdx.x; u < J.n; u += nb * 256u) {
    J.label[u] = J.core_cnt[u] >= J.mu ? 1 : 0;
    J.parent[u] = static_cast<int32_t>(u);
    J.rmin[u] = INT32_MAX;
    J.rmax[u] = -1;
    J.cl[u] = INT32_MAX;
  }
}

// ClusterCore: union over similar core-core edges (each undirected edge once).
__global__ __launch_bounds__(256) void k_sj_union(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, true, &lb, &nb);
  for (uint64_t i = lb * 256ull + threadIdx.x; i < J.m; i += nb * 256ull) {
    const int32_t u = J.src[i];
    const int32_t v = J.adj[i];
    if (u < v && J.sim[i] && J.label[u] == 1 && J.label[v] == 1) uf_union(J.parent, u, v);
  }
}

// Points every core at its root.  Only roots are stored, and every pointer
// leads to the root, so concurrent walks stay correct.
__global__ __launch_bounds__(256) void k_sj_compress(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  for (uint32_t u = lb * 256u + threadIdx.x; u < J.n; u += nb * 256u)
    if (J.label[u] == 1)
      __hip_atomic_store(&J.parent[u], uf_root(J.parent, static_cast<int32_t>(u)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// CheckHub: per non-core vertex, the smallest and largest root among its core neighbors.
__global__ __launch_bounds__(256) void k_sj_hub(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, true, &lb, &nb);
  for (uint64_t i = lb * 256ull + threadIdx.x; i < J.m; i += nb * 256ull) {
    const int32_t a = J.src[i];
    const int32_t v = J.adj[i];
    if (J.label[a] != 1 && J.label[v] == 1) {
      const int32_t r = J.parent[v];
      if (r < J.rmin[a]) atomicMin(&J.rmin[a], r);
      if (r > J.rmax[a]) atomicMax(&J.rmax[a], r);
    }
  }
}

// Labels hubs, takes the minimum old id of every cluster and counts
// cores, clusters, hubs and outliers (counters[0..3]).
__global__ __launch_bounds__(256) void k_sj_finalize(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  unsigned long long cores = 0, clusters = 0, hubs = 0, outliers = 0;
  for (uint32_t u = lb * 256u + threadIdx.x; u < J.n; u += nb * 256u) {
    if (J.label[u] == 1) {
      cores++;
      const int32_t r = J.parent[u];
      if (r == static_cast<int32_t>(u)) clusters++;
      const int32_t o = J.old ? J.old[u] : static_cast<int32_t>(u);
      atomicMin(&J.cl[r], o);
    } else if (J.rmax[u] >= 0 && J.rmin[u] != J.rmax[u]) {
      J.label[u] = 2;
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
    if (t) atomicAdd(&J.counters[threadIdx.x], t);
  }
}

// Cluster ids; with host-mapped outputs set, the labels (as int8), the
// cores' cluster ids and the four counters go straight to host memory
// instead (the host keeps every other entry of out_cid at -1).
__global__ __launch_bounds__(256) void k_sj_cid(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  if (J.out_cid) {
    for (uint32_t u = lb * 256u + threadIdx.x; u < J.n; u += nb * 256u) {
      const int32_t l = J.label[u];
      if (l == 1) J.out_cid[u] = J.cl[J.parent[u]];
      J.out_label[u] = static_cast<int8_t>(l);
    }
    if (lb == 0 && threadIdx.x < 4) J.out_counters[threadIdx.x] = J.counters[threadIdx.x];
    // the host reads these once the batch's completion event fires
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    return;
  }
  for (uint32_t u = lb * 256u + threadIdx.x; u < J.n; u += nb * 256u)
    J.cid[u] = J.label[u] == 1 ? J.cl[J.parent[u]] : -1;
}

__global__ __launch_bounds__(64) void k_sj_zero_pairs(const JobTable jt) {
  if (static_cast<int>(threadIdx.x) < jt.nj) jt.jobs[threadIdx.x].counters[4] = 0;
}

// Non-core memberships: (cluster id of core u, old id of v) for every similar
// slot (u, v) with u core and v not core, emitted in any order; counters[4]
// counts them, capacity pairs_cap.  Every lane takes its position from the
// same counter with an increment of 0 or 1, so positions are unique without a
// lane-0 broadcast.
__global__ __launch_bounds__(256) void k_sj_pairs(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, true, &lb, &nb);
  for (uint64_t i = lb * 256ull + threadIdx.x; i < J.m; i += nb * 256ull) {
    if (!J.sim[i]) continue;
    const int32_t u = J.src[i];
    const int32_t v = J.adj[i];
    if (J.label[u] != 1 || J.label[v] == 1) continue;
    const unsigned long long pos = atomicAdd(&J.counters[4], 1ull);
    if (pos < J.pairs_cap) J.pairs[pos] = make_int2(J.cid[u], J.old ? J.old[v] : v);
  }
}

// ---------------------------------------------------------------------------
// Similarity index.  key(x) = float(double(x)) of x = cn^2 / ((d_u+1)(d_v+1)),
// computed as a correctly rounded double division of two exact doubles
// (cn < 2^26 and degrees below 2^26 keep both exact) and a round-to-nearest
// conversion: a composition of monotone roundings, so key is monotone
// nondecreasing in x.  With e = p^2 / q^2 (p, q < 2^26) and ke = key(e):
// key(x) > ke implies x >= e (similar), key(x) < ke implies x < e (not
// similar), and key(x) == ke is decided by the exact integer test.  Every
// vertex's slots are sorted by key, descending, once per graph; a query then
// finds its similar neighbors as a prefix of that order.
__device__ __forceinline__ float sim_key(unsigned long long num, unsigned long long den) {
  return static_cast<float>(static_cast<double>(num) / static_cast<double>(den));
}

__global__ __launch_bounds__(256) void k_idx_key(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                 const int32_t *__restrict__ src, const int32_t *__restrict__ cnt,
                                                 uint64_t m, int32_t add, float *__restrict__ key,
                                                 uint32_t *__restrict__ slot) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < m; i += gridDim.x * 256ull) {
    const int32_t u = src[i], v = adj[i];
    const unsigned long long c = static_cast<unsigned long long>(cnt[i] + add);
    const unsigned long long d = static_cast<unsigned long long>(off[u + 1] - off[u] + 1) *
                                 static_cast<unsigned long long>(off[v + 1] - off[v] + 1);
    key[i] = sim_key(c * c, d);
    slot[i] = static_cast<uint32_t>(i);
  }
}

__global__ __launch_bounds__(256) void k_idx_gather(const int32_t *__restrict__ adj,
                                                    const uint32_t *__restrict__ slot, uint64_t m,
                                                    int32_t *__restrict__ nbr) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < m; i += gridDim.x * 256ull) nbr[i] = adj[slot[i]];
}

// Exact similarity of slot s = (u, v): q^2 cn^2 >= p^2 (d_u+1)(d_v+1).
__device__ __forceinline__ bool exact_sim(const ScanJob &J, uint32_t s, uint64_t du1, int32_t v) {
  const uint64_t c = static_cast<uint64_t>(J.cnt[s] + J.add);
  const uint64_t dv1 = static_cast<uint64_t>(J.off[v + 1] - J.off[v]) + 1;
  const unsigned __int128 lhs = static_cast<unsigned __int128>(J.q2) * (c * c);
  const unsigned __int128 rhs = static_cast<unsigned __int128>(J.p2) * (du1 * dv1);
  return lhs >= rhs;
}

__global__ __launch_bounds__(64) void k_sx_zero(const JobTable jt) {
  if (threadIdx.x < kScanCounters) jt.jobs[blockIdx.x].counters[threadIdx.x] = 0;
}

// First position in [lo, hi) of the descending keys K whose key is < ke
// (strict) or <= ke (!strict).
__device__ __forceinline__ uint32_t key_bound(const float *K, uint32_t lo, uint32_t hi, float ke, bool strict) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const float k = K[mid];
    if (strict ? k >= ke : k > ke) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Appends `items` work items (u, 0), (u, kItem), ... per lane to J.work,
// counted in *count: one atomic per wave, each lane writing at its prefix.
__device__ __forceinline__ void append_items(const ScanJob &J, unsigned long long *count, uint32_t u,
                                             uint32_t items) {
  const int lane = threadIdx.x & 63;
  const uint32_t pre = wave_excl_scan32(items, lane);
  const uint32_t tot = __builtin_amdgcn_readfirstlane(__shfl(pre + items, 63));
  if (!tot) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(count, static_cast<unsigned long long>(tot));
  base = __shfl(base, 0);
  for (uint32_t k = 0; k < items; k++)
    J.work[base + pre + k] = make_int2(static_cast<int32_t>(u), static_cast<int32_t>(k * kItem));
}

// Items of every core's possibly-similar prefix (counters[8]).
__device__ __forceinline__ void union_items(const ScanJob &J, uint32_t u, bool core, uint32_t end) {
  append_items(J, &J.counters[8], u, core ? (end + kItem - 1) / kItem : 0);
}

// Core check from the index, with ke = key(eps^2) and K = u's sorted keys:
// the mu-th largest key decides unless it ties with ke (K[mu-1] > ke: at
// least mu similar neighbors; < ke: fewer than mu), in which case the tie
// block takes the exact test.  send[u] = the end of a core's possibly-similar
// prefix (keys >= ke), 0 for non-cores; the prefixes are cut into the union's
// work items.  Degree sums over cores and non-cores go to counters[5] and
// counters[6], the total prefix length to counters[7].
__global__ __launch_bounds__(256) void k_sx_core(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  const float ke = sim_key(J.p2, J.q2);
  unsigned long long core_deg = 0, other_deg = 0, prefix = 0;
  const int lane = threadIdx.x & 63;
  // wave-uniform loop: the item append at its end is a wave operation that
  // every lane reaches
  for (uint32_t u0 = lb * 256u + (threadIdx.x & ~63u); u0 < J.n; u0 += nb * 256u) {
    const uint32_t u = u0 + lane;
    bool core = false;
    uint32_t end = 0;
    if (u < J.n) {
      const uint32_t base = J.off[u];
      const uint32_t d = J.off[u + 1] - base;
      const float *K = J.skey + base;
      int32_t count = 0;
      if (J.mu <= 0) {
        core = true;
        end = key_bound(K, 0, d, ke, true);
      } else if (d >= static_cast<uint32_t>(J.mu)) {
        const uint32_t t = static_cast<uint32_t>(J.mu) - 1;
        const float k = K[t];
        if (k > ke) {
          core = true;
          end = key_bound(K, t + 1, d, ke, true);
        } else if (k == ke) {
          const uint32_t lo = key_bound(K, 0, t, ke, false);
          end = key_bound(K, t + 1, d, ke, true);
          count = static_cast<int32_t>(lo);
          for (uint32_t i = lo; i < end; i++)
            if (exact_sim(J, J.sslot[base + i], static_cast<uint64_t>(d) + 1, J.snbr[base + i])) count++;
          core = count >= J.mu;
        }
      }
      // each array only where a later kernel reads it: parent, cl and send
      // of cores, rmin / rmax of non-cores; core counts only for workspaces
      if (!J.out_cid) J.core_cnt[u] = count;
      J.label[u] = core ? 1 : 0;
      if (core) {
        J.parent[u] = static_cast<int32_t>(u);
        J.cl[u] = INT32_MAX;
        J.send[u] = end;
      } else {
        end = 0;
        J.rmin[u] = INT32_MAX;
        J.rmax[u] = -1;
      }
      if (core) {
        core_deg += d;
        prefix += end;
      } else {
        other_deg += d;
      }
    }
    union_items(J, u, core, end);
  }
  core_deg = wave_sum_u64(core_deg);
  other_deg = wave_sum_u64(other_deg);
  prefix = wave_sum_u64(prefix);
  if (lane == 0) {
    if (core_deg) atomicAdd(&J.counters[5], core_deg);
    if (other_deg) atomicAdd(&J.counters[6], other_deg);
    if (prefix) atomicAdd(&J.counters[7], prefix);
  }
}

// Whether the index path's clustering kernels take work items (the lists
// they touch hold under a quarter of the slots) or sweep every slot.
__device__ __forceinline__ bool use_items(unsigned long long touched, uint64_t m) { return touched * 4 < m; }


// ClusterCore from the index, in three steps after the idea of Afforest
// (Sutton, Ben-Nun and Barak, IPDPS 2018): every core first links to the
// first two similar cores in its prefix (k_sx_link); the root holding most
// of 1024 sampled cores is taken as the giant cluster's (k_sx_giant); then
// only cores outside that cluster unite with every similar core of their
// prefix, in both directions (k_sx_union).  Similarity is symmetric, so an
// edge whose one end sits in the giant cluster is processed from its other
// end, and the clusters are those of uniting every similar core pair.
__global__ __launch_bounds__(256) void k_sx_link(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  if (J.counters[7] == 0) return;
  const float ke = sim_key(J.p2, J.q2);
  for (uint32_t u = lb * 256u + threadIdx.x; u < J.n; u += nb * 256u) {
    if (J.label[u] != 1) continue;
    const uint32_t base = J.off[u];
    const uint32_t end = base + min(J.send[u], 64u);
    int linked = 0;
    for (uint32_t i = base; i < end && linked < 2; i++) {
      const int32_t v = J.snbr[i];
      if (J.label[v] != 1) continue;
      if (J.skey[i] == ke && !exact_sim(J, J.sslot[i], static_cast<uint64_t>(J.off[u + 1] - base) + 1, v)) continue;
      uf_union(J.parent, static_cast<int32_t>(u), v);
      linked++;
    }
  }
}

// One workgroup per job: the most frequent root among up to 1024 sampled
// cores (an LDS hash table of counts) into counters[10]; ~0 with no core.
__global__ __launch_bounds__(1024) void k_sx_giant(const JobTable jt) {
  __shared__ int32_t keys[2048];
  __shared__ int32_t cnts[2048];
  __shared__ unsigned long long best[16];
  const ScanJob &J = jt.jobs[blockIdx.x];
  const int t = threadIdx.x;
  for (int i = t; i < 2048; i += 1024) {
    keys[i] = -1;
    cnts[i] = 0;
  }
  __syncthreads();
  if (J.counters[7] != 0 && J.n) {
    const uint32_t u = static_cast<uint32_t>((static_cast<uint64_t>(t) * J.n) / 1024u);
    if (J.label[u] == 1) {
      const int32_t r = uf_root(J.parent, static_cast<int32_t>(u));
      uint32_t h = (static_cast<uint32_t>(r) * 0x9E3779B1u) >> 21;
      for (;;) {
        const int32_t k = atomicCAS(&keys[h], -1, r);
        if (k == -1 || k == r) {
          atomicAdd(&cnts[h], 1);
          break;
        }
        h = (h + 1) & 2047;
      }
    }
  }
  __syncthreads();
  // (count, root) packed so that the maximum is the most frequent root
  unsigned long long m = 0;
  for (int i = t; i < 2048; i += 1024)
    if (cnts[i]) {
      const unsigned long long c = (static_cast<unsigned long long>(cnts[i]) << 32) | static_cast<uint32_t>(keys[i]);
      m = c > m ? c : m;
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(m, o, 64);
    m = x > m ? x : m;
  }
  if ((t & 63) == 0) best[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    unsigned long long b = 0;
    for (int i = 0; i < 16; i++) b = best[i] > b ? best[i] : b;
    J.counters[10] = b ? (b & 0xffffffffull) : ~0ull;
  }
}

// One wave per work item written by k_sx_core (up to kItem consecutive
// positions of one core's prefix); items of cores already in the giant
// cluster are skipped.  Each round of 64 positions unites the core with its
// similar cores through one uf_union_wave.  Loops are wave-uniform.
__global__ __launch_bounds__(256) void k_sx_union(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, true, &lb, &nb);
  if (J.counters[7] == 0) return;
  const float ke = sim_key(J.p2, J.q2);
  const int32_t giant = static_cast<int32_t>(J.counters[10]);  // -1: none
  const int lane = threadIdx.x & 63;
  const uint32_t gw = __builtin_amdgcn_readfirstlane(lb * 4u + (threadIdx.x >> 6));
  const uint32_t nw = nb * 4u;
  const uint32_t items = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(J.counters[8]));
  for (uint32_t it = gw; it < items; it += nw) {
    const int2 w = J.work[it];
    const int32_t a = __builtin_amdgcn_readfirstlane(w.x);
    if (__builtin_amdgcn_readfirstlane(uf_root(J.parent, a)) == giant) continue;
    const uint32_t start = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(w.y));
    const uint32_t base = __builtin_amdgcn_readfirstlane(J.off[a]);
    const uint32_t end = base + min(start + kItem, static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(J.send[a])));
    const uint64_t da1 = static_cast<uint64_t>(J.off[a + 1] - base) + 1;
    for (uint32_t i0 = base + start; i0 < end; i0 += 64) {
      const uint32_t i = i0 + lane;
      int32_t v = 0;
      bool valid = false;
      if (i < end) {
        v = J.snbr[i];
        valid = v != a && J.label[v] == 1 && (J.skey[i] != ke || exact_sim(J, J.sslot[i], da1, v));
      }
      uf_union_wave(J.parent, a, v, valid);
    }
  }
}

// CheckHub side: the side with fewer adjacency slots (counters[5] over
// cores, counters[6] over non-cores); cores push their root to non-core
// neighbors, or non-cores pull their core neighbors' roots.
__device__ __forceinline__ bool hub_push(const ScanJob &J) { return J.counters[5] <= J.counters[6]; }

// Points every core at its root (as k_sj_compress) and, when CheckHub takes
// work items, writes the items of every adjacency on the chosen side
// (counters[9]).  Wave-uniform loop.
__global__ __launch_bounds__(256) void k_sx_compress(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, false, &lb, &nb);
  const bool push = hub_push(J);
  const unsigned long long side = push ? J.counters[5] : J.counters[6];
  const bool items = side != 0 && use_items(side, J.m);
  const int lane = threadIdx.x & 63;
  for (uint32_t u0 = lb * 256u + (threadIdx.x & ~63u); u0 < J.n; u0 += nb * 256u) {
    const uint32_t u = u0 + lane;
    const bool in = u < J.n;
    const bool core = in && J.label[u] == 1;
    if (core)
      __hip_atomic_store(&J.parent[u], uf_root(J.parent, static_cast<int32_t>(u)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (items)
      append_items(J, &J.counters[9], u, in && core == push ? (J.off[u + 1] - J.off[u] + kItem - 1) / kItem : 0);
  }
}

// CheckHub.  Item form: one wave per work item of the chosen side; slot
// form: every slot, four independent slots per thread and round.  Both
// leave the smallest and largest root of each non-core's core neighbors in
// rmin / rmax.
__device__ __forceinline__ void hub_at(const ScanJob &J, bool push, int32_t a, int32_t ra, int32_t x) {
  if ((J.label[x] == 1) == push) return;
  const int32_t r = push ? ra : J.parent[x];
  const int32_t y = push ? x : a;
  if (r < J.rmin[y]) atomicMin(&J.rmin[y], r);
  if (r > J.rmax[y]) atomicMax(&J.rmax[y], r);
}

__global__ __launch_bounds__(256) void k_sx_hub(const JobTable jt) {
  uint32_t lb, nb;
  const ScanJob J = job_of(jt, true, &lb, &nb);
  const bool push = hub_push(J);
  const unsigned long long side = push ? J.counters[5] : J.counters[6];
  if (side == 0) return;
  if (use_items(side, J.m)) {
    const int lane = threadIdx.x & 63;
    const uint32_t items = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(J.counters[9]));
    const uint32_t nw = nb * 4u;
    for (uint32_t it = __builtin_amdgcn_readfirstlane(lb * 4u + (threadIdx.x >> 6)); it < items; it += nw) {
      const int2 w = J.work[it];
      const int32_t a = __builtin_amdgcn_readfirstlane(w.x);
      const uint32_t base = J.off[a];
      const uint32_t end = min(base + static_cast<uint32_t>(w.y) + kItem, J.off[a + 1]);
      const int32_t ra = push ? J.parent[a] : 0;
      for (uint32_t i = base + static_cast<uint32_t>(w.y) + lane; i < end; i += 64) hub_at(J, push, a, ra, J.adj[i]);
    }
    return;
  }
  const uint64_t stride = nb * 256ull;
  for (uint64_t i0 = lb * 256ull + threadIdx.x; i0 < J.m; i0 += 4 * stride) {
    int32_t a[4], w[4];
    bool on[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint64_t i = i0 + q * stride;
      a[q] = i < J.m ? J.src[i] : -1;
      w[q] = i < J.m ? J.adj[i] : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) on[q] = a[q] >= 0 && (J.label[a[q]] == 1) == push;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (on[q]) hub_at(J, push, a[q], push ? J.parent[a[q]] : 0, w[q]);
  }
}

inline unsigned grid_for(uint64_t items, int cus) {
  const uint64_t want = (items + 255) / 256;
  const uint64_t c = static_cast<uint64_t>(cus > 0 ? cus : 256) * 8;
  return static_cast<unsigned>(want < c ? (want ? want : 1) : c);
}

}  // namespace

namespace accscan {

// Workgroup ranges of every job: a job gets one workgroup per 256 items, at
// most 8 per CU, at least one.
void plan_jobs(ScanJob *jobs, int nj, int cus, uint32_t *vblocks, uint32_t *sblocks) {
  uint32_t v = 0, s = 0;
  for (int j = 0; j < nj; j++) {
    jobs[j].vb0 = v;
    jobs[j].vbn = grid_for(jobs[j].n, cus);
    jobs[j].sb0 = s;
    jobs[j].sbn = grid_for(jobs[j].m, cus);
    v += jobs[j].vbn;
    s += jobs[j].sbn;
  }
  *vblocks = v;
  *sblocks = s;
}

int launch_jobs(const JobTable &jt, uint32_t vblocks, uint32_t sblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_sj_init, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_sim, dim3(sblocks), dim3(kScanWG), 0, s, jt);
  hipLaunchKernelGGL(k_sj_label, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_union, dim3(sblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_compress, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_hub, dim3(sblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_finalize, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_cid, dim3(vblocks), dim3(256), 0, s, jt);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_index_jobs(const JobTable &jt, uint32_t vblocks, uint32_t sblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_sx_zero, dim3(jt.nj), dim3(64), 0, s, jt);
  hipLaunchKernelGGL(k_sx_core, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sx_link, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sx_giant, dim3(jt.nj), dim3(1024), 0, s, jt);
  hipLaunchKernelGGL(k_sx_union, dim3(sblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sx_compress, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sx_hub, dim3(sblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_finalize, dim3(vblocks), dim3(256), 0, s, jt);
  hipLaunchKernelGGL(k_sj_cid, dim3(vblocks), dim3(256), 0, s, jt);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

bool index_usable(const acc_graph *g, int64_t p, int64_t q) {
  const int64_t lim = int64_t(1) << 26;
  return g->d_skey && p < lim && q < lim;
}

int build_index(acc_graph *g) {
  if (g->d_skey) return ACC_OK;
  if (!g->have_result) return acc_set_error(ACC_EINVAL, "index needs the graph's counts");
  if (g->max_deg + 2 >= (1u << 26)) return acc_set_error(ACC_EINVAL, "degrees too large for the similarity index");
  ACC_TRY(acc_graph_ensure_src(g));
  const uint64_t m = g->m;
  if (m >= (1ull << 31)) return acc_set_error(ACC_EINVAL, "too many slots for the similarity index");
  hipStream_t st = g->ctx->stream;
  const int cus = g->ctx->num_cus;
  float *key_in = nullptr;
  void *temp = nullptr;
  size_t temp_bytes = 0;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(st);
    if (key_in) (void)hipFree(key_in);
    if (temp) (void)hipFree(temp);
  };
  auto fail = [&](const char *what, hipError_t e) {
    cleanup();
    for (void **p : {reinterpret_cast<void **>(&g->d_skey), reinterpret_cast<void **>(&g->d_snbr),
                     reinterpret_cast<void **>(&g->d_sslot)})
      if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
      }
    return acc_set_error(e == hipErrorOutOfMemory ? ACC_ENOMEM : ACC_EDEVICE,
                         std::string(what) + ": " + hipGetErrorString(e));
  };
  const size_t mm = m ? m : 1;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&key_in), sizeof(float) * mm);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&g->d_skey), sizeof(float) * mm);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&g->d_sslot), sizeof(uint32_t) * mm);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&g->d_snbr), sizeof(int32_t) * mm);
  if (e != hipSuccess) return fail("index allocation", e);
  if (m) {
    // the sort is double-buffered (keys between key_in and d_skey, slots
    // between d_snbr, free until the gather, and d_sslot), so its temporary
    // storage holds no copy of the pairs
    uint32_t *const slot_in = reinterpret_cast<uint32_t *>(g->d_snbr);
    hipLaunchKernelGGL(k_idx_key, dim3(grid_for(m, cus)), dim3(256), 0, st, g->d_off, g->d_adj, g->d_src, g->d_cnt,
                       m, g->cnt_plus2 ? 0 : 2, key_in, slot_in);
    e = hipGetLastError();
    if (e != hipSuccess) return fail("index keys", e);
    hipcub::DoubleBuffer<float> kb(key_in, g->d_skey);
    hipcub::DoubleBuffer<uint32_t> sb(slot_in, g->d_sslot);
    e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(nullptr, temp_bytes, kb, sb, static_cast<int>(m),
                                                              static_cast<int>(g->n), g->d_off, g->d_off + 1, 0,
                                                              32, st);
    if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes ? temp_bytes : 1);
    kb = hipcub::DoubleBuffer<float>(key_in, g->d_skey);  // as before the size query
    sb = hipcub::DoubleBuffer<uint32_t>(slot_in, g->d_sslot);
    if (e == hipSuccess)
      e = hipcub::DeviceSegmentedRadixSort::SortPairsDescending(temp, temp_bytes, kb, sb, static_cast<int>(m),
                                                                static_cast<int>(g->n), g->d_off, g->d_off + 1, 0,
                                                                32, st);
    if (e == hipSuccess && kb.Current() != g->d_skey)
      e = hipMemcpyAsync(g->d_skey, kb.Current(), sizeof(float) * m, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && sb.Current() != g->d_sslot)
      e = hipMemcpyAsync(g->d_sslot, sb.Current(), sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail("index sort", e);
    hipLaunchKernelGGL(k_idx_gather, dim3(grid_for(m, cus)), dim3(256), 0, st, g->d_adj, g->d_sslot, m, g->d_snbr);
    e = hipGetLastError();
    if (e != hipSuccess) return fail("index gather", e);
  }
  cleanup();
  g->device_bytes += 12 * mm;
  return ACC_OK;
}

int launch_pairs(const JobTable &jt, uint32_t sblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_sj_zero_pairs, dim3(1), dim3(64), 0, s, jt);
  hipLaunchKernelGGL(k_sj_pairs, dim3(sblocks), dim3(256), 0, s, jt);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

}  // namespace accscan

struct acc_scan_ws {
  acc_graph *g = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;  // counts of the graph are complete
  ScanBuffers buf;
  JobTable jt{};  // the last query as a one-job table
  uint32_t vblocks = 0, sblocks = 0;
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
  accscan::free_buffers(&w->buf);
  if (w->d_pairs) (void)hipFree(w->d_pairs);
  if (w->ready) (void)hipEventDestroy(w->ready);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
}

}  // namespace

namespace accscan {

void free_buffers(ScanBuffers *b) {
  for (void *p : {static_cast<void *>(b->sim), static_cast<void *>(b->send), static_cast<void *>(b->work),
                  static_cast<void *>(b->core_cnt),
                  static_cast<void *>(b->label),
                  static_cast<void *>(b->parent), static_cast<void *>(b->rmin), static_cast<void *>(b->rmax),
                  static_cast<void *>(b->cl), static_cast<void *>(b->cid), static_cast<void *>(b->counters)})
    if (p) (void)hipFree(p);
  *b = ScanBuffers{};
}

// Grows the per-query buffers to hold a graph of n vertices and m slots.
int reserve_buffers(ScanBuffers *b, uint32_t n, uint64_t m) {
  if (b->counters && n <= b->n_cap && m <= b->m_cap) return ACC_OK;
  const uint32_t nc = std::max(n, b->n_cap);
  const uint64_t mc = std::max(m, b->m_cap);
  free_buffers(b);
  auto alloc = [&](auto **p, uint64_t count) -> int {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(*p)>>;
    const size_t bytes = sizeof(T) * (count ? count : 1);
    if (hipMalloc(reinterpret_cast<void **>(p), bytes) != hipSuccess) {
      *p = nullptr;
      return acc_set_error(ACC_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes of SCAN buffers failed");
    }
    b->bytes += bytes;
    return ACC_OK;
  };
  int rc = ACC_OK;
  if (!rc) rc = alloc(&b->sim, mc);
  if (!rc) rc = alloc(&b->send, nc);
  if (!rc) rc = alloc(&b->core_cnt, nc);
  if (!rc) rc = alloc(&b->label, nc);
  if (!rc) rc = alloc(&b->parent, nc);
  if (!rc) rc = alloc(&b->rmin, nc);
  if (!rc) rc = alloc(&b->rmax, nc);
  if (!rc) rc = alloc(&b->cl, nc);
  if (!rc) rc = alloc(&b->cid, nc);
  if (!rc) rc = alloc(&b->counters, kScanCounters);
  if (!rc) rc = alloc(&b->work, static_cast<uint64_t>(nc) + mc / kItem + 64);
  if (rc) {
    free_buffers(b);
    return rc;
  }
  b->n_cap = nc;
  b->m_cap = mc;
  return ACC_OK;
}

uint64_t buffer_bytes(uint32_t n, uint64_t m) {
  const uint64_t nn = n ? n : 1, mm = m ? m : 1;
  return mm + 9 * sizeof(int32_t) * nn + sizeof(int2) * (nn + mm / kItem + 64) +
         kScanCounters * sizeof(unsigned long long);
}

int check_eps(int64_t eps_num, int64_t eps_den) {
  if (eps_num < 0 || eps_den <= 0 || eps_num > (1ll << 31) || eps_den > (1ll << 31))
    return acc_set_error(ACC_EINVAL, "eps must be eps_num/eps_den with 0 <= eps_num and 0 < eps_den <= 2^31");
  return ACC_OK;
}

// The job of one query of graph g with the given buffers.
ScanJob make_job(const acc_graph *g, const ScanBuffers &b, int64_t eps_num, int64_t eps_den, int32_t mu) {
  ScanJob j{};
  j.off = g->d_off;
  j.adj = g->d_adj;
  j.src = g->d_src;
  j.cnt = g->d_cnt;
  j.old = g->d_old;
  j.sim = b.sim;
  j.skey = g->d_skey;
  j.snbr = g->d_snbr;
  j.sslot = g->d_sslot;
  j.send = b.send;
  j.work = b.work;
  j.core_cnt = b.core_cnt;
  j.label = b.label;
  j.parent = b.parent;
  j.rmin = b.rmin;
  j.rmax = b.rmax;
  j.cl = b.cl;
  j.cid = b.cid;
  j.counters = b.counters;
  j.p2 = static_cast<unsigned long long>(eps_num) * static_cast<unsigned long long>(eps_num);
  j.q2 = static_cast<unsigned long long>(eps_den) * static_cast<unsigned long long>(eps_den);
  j.m = g->m;
  j.n = g->n;
  j.add = g->cnt_plus2 ? 0 : 2;
  j.mu = mu;
  return j;
}

}  // namespace accscan

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
    if (g->d_old) {
      (void)hipFree(g->d_old);
      g->device_bytes -= sizeof(int32_t) * (g->n ? g->n : 1);
    }
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
  if ((rc = accscan::reserve_buffers(&w->buf, g->n, g->m))) return bail(rc);
  w->device_bytes = w->buf.bytes;
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
  ACC_TRY(accscan::check_eps(eps_num, eps_den));
  acc_graph *g = w->g;
  (void)hipSetDevice(g->ctx->device);
  if (!g->have_result || g->counts_share) {
    // SCAN consumes every slot's count: one share's counts (acc_graph_set_shard) would
    // cluster on partial counts, so a share that the counting pass would use is refused
    const bool share = g->shard_lo != 0 || g->shard_hi != g->n;
    if (share && g->auto_path == ACC_PATH_FORWARD)
      return acc_set_error(ACC_EINVAL, "SCAN needs the whole graph's counts, but the forward path is restricted "
                                       "to one share (acc_graph_set_shard(g, 1, 0, NULL) restores the whole graph)");
    ACC_TRY(acc_cn_compute(g, ACC_PATH_AUTO, 0));
  }
  // order this query after the graph's counts
  ACC_HIP_TRY(hipEventRecord(w->ready, g->ctx->stream));
  ACC_HIP_TRY(hipStreamWaitEvent(w->stream, w->ready, 0));
  w->jt.jobs[0] = accscan::make_job(g, w->buf, eps_num, eps_den, mu);
  w->jt.nj = 1;
  accscan::plan_jobs(w->jt.jobs, 1, g->ctx->num_cus, &w->vblocks, &w->sblocks);
  ACC_TRY(accscan::launch_jobs(w->jt, w->vblocks, w->sblocks, w->stream));
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
  const ScanBuffers &b = w->buf;
  if (label) ACC_HIP_TRY(hipMemcpyAsync(label, b.label, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, s));
  if (cluster_id) ACC_HIP_TRY(hipMemcpyAsync(cluster_id, b.cid, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, s));
  if (sim && g->m) ACC_HIP_TRY(hipMemcpyAsync(sim, b.sim, g->m, hipMemcpyDeviceToHost, s));
  if (core_count)
    ACC_HIP_TRY(hipMemcpyAsync(core_count, b.core_cnt, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, s));
  unsigned long long c[8] = {0};
  ACC_HIP_TRY(hipMemcpyAsync(c, b.counters, sizeof(c), hipMemcpyDeviceToHost, s));
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
    const size_t bytes = sizeof(int2) * (g->m ? g->m : 1);
    if (hipMalloc(reinterpret_cast<void **>(&w->d_pairs), bytes) != hipSuccess) {
      w->d_pairs = nullptr;
      return acc_set_error(ACC_ENOMEM, "hipMalloc of the pair buffer failed");
    }
    w->device_bytes += bytes;
    w->pairs_cap = g->m;
  }
  w->jt.jobs[0].pairs = w->d_pairs;
  w->jt.jobs[0].pairs_cap = w->pairs_cap;
  ACC_TRY(accscan::launch_pairs(w->jt, w->sblocks, s));
  unsigned long long k = 0;
  ACC_HIP_TRY(hipMemcpyAsync(&k, w->buf.counters + 4, sizeof(k), hipMemcpyDeviceToHost, s));
  ACC_HIP_TRY(hipStreamSynchronize(s));
  *count = k;
  if (pairs && k) {
    const uint64_t take = std::min<uint64_t>(k, cap);
    ACC_HIP_TRY(hipMemcpy(pairs, w->d_pairs, sizeof(int2) * take, hipMemcpyDeviceToHost));
  }
  return ACC_OK;
}

}  // extern "C"
