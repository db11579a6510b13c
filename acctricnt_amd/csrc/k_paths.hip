// The two per-edge counting paths of the reference, MI355X-native:
//
// ACC_PATH_BITMAP -- the reference's BMP algorithm
//   (set_intersection_GPU_bitmap_warp_per_vertex,
//    scan-query-cuda/scan_xp_multi_gpu_multi_pass_common.cu:308-407):
//   for each source u a membership set of N(u) is built once, then every
//   neighbor list N(v), v > u, is probed against it.  Sets of at most
//   kLdsCap ids live in an LDS hash table; hub sources use an n-bit bitmap
//   in HBM owned by one workgroup of a persistent launch (the reference's
//   per-SM bitmap pool, without its %smid slot search).
// ACC_PATH_MERGE -- the reference's MPS kernels
//   (set_intersection_GPU_shared_merge_filtered / _galloping_filtered, same
//    file :68-176): one wave per edge; balanced pairs merge the shorter list
//   against 64-element LDS-staged windows of the longer one with a binary
//   search per lane and a ballot/popcount; skewed pairs (ratio above
//   kSkew) binary-search the longer list in global memory per element.
//
// Both write 2 + |N(u) ∩ N(v)| (or the raw count) into the upper slot (u, v)
// and its reverse slot (the position of u in N(v)), with no atomics on the
// counts, and add the raw counts to the triangle sum.

#include "acc_internal.h"

namespace {

constexpr int kWG = accfwd::kWG;
constexpr int kWaves = accfwd::kWaves;
constexpr int kTBits = accfwd::kTableBits;
constexpr int kT = 1 << kTBits;
constexpr uint32_t kChunk = accfwd::kChunk;
constexpr uint32_t kSkew = 32;

__device__ __forceinline__ uint32_t hslot(int32_t a, int bits) {
  return (static_cast<uint32_t>(a) * 0x9E3779B1u) >> (32 - bits);
}

__device__ __forceinline__ int wsum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ bool hash_has(const int32_t *keys, int bits, uint32_t mask, int32_t a) {
  uint32_t h = hslot(a, bits);
  for (;;) {
    const int32_t k = keys[h];
    if (k == a) return true;
    if (k == -1) return false;
    h = (h + 1) & mask;
  }
}

// Position of x in the sorted list adj[b, e) (x must be present); all lanes
// run the same search, so the loads are broadcasts.
__device__ __forceinline__ uint32_t find_pos(const int32_t *__restrict__ adj, uint32_t b, uint32_t e, int32_t x) {
  while (b < e) {
    const uint32_t mid = b + ((e - b) >> 1);
    if (adj[mid] < x)
      b = mid + 1;
    else
      e = mid;
  }
  return b;
}

__device__ __forceinline__ void block_add_sum(unsigned long long *sum, unsigned long long s) {
  __shared__ unsigned long long part[kWaves];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWaves; w++) t += part[w];
    if (t) atomicAdd(sum, t);
  }
}

// Stores the count of edge (u, v) into its upper slot j and reverse slot p.
__device__ __forceinline__ void store_pair(int32_t *cnt, uint32_t j, uint32_t p, int c, int plus) {
  cnt[j] = c + plus;
  cnt[p] = c + plus;
}

// ---------------------------------------------------------------- BITMAP path
// LDS-set units: one workgroup per chunk of at most kChunk upper slots of a
// source u with d_u <= kCap.
__global__ __launch_bounds__(kWG) void k_bmp_lds(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                 const FwdUnit *__restrict__ units, int32_t *cnt, int plus,
                                                 unsigned long long *sum) {
  __shared__ int32_t keys[kT];
  __shared__ uint32_t s_cb[kChunk], s_ce[kChunk];
  const FwdUnit un = units[blockIdx.x];
  const int32_t u = un.b;
  const uint32_t su = off[u], du = off[u + 1] - su;
  const uint32_t nup = un.j1 - un.j0;
  int bits = 6;
  while ((1u << bits) < 2 * du) bits++;
  const uint32_t mask = (1u << bits) - 1;
  for (uint32_t i = threadIdx.x; i <= mask; i += kWG) keys[i] = -1;
  for (uint32_t i = threadIdx.x; i < nup; i += kWG) {
    const int32_t v = adj[un.j0 + i];
    s_cb[i] = off[v];
    s_ce[i] = off[v + 1];
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < du; i += kWG) {
    const int32_t a = adj[su + i];
    uint32_t h = hslot(a, bits);
    while (atomicCAS(&keys[h], -1, a) != -1) h = (h + 1) & mask;
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  unsigned long long s = 0;
  for (uint32_t k = wave; k < nup; k += kWaves) {
    const uint32_t vb = __builtin_amdgcn_readfirstlane(s_cb[k]);
    const uint32_t ve = __builtin_amdgcn_readfirstlane(s_ce[k]);
    int hits = 0;
    uint32_t p = ~0u;
    for (uint32_t base = vb; base < ve; base += 64) {
      const uint32_t i = base + lane;
      const int32_t a = i < ve ? adj[i] : -1;
      if (a >= 0 && hash_has(keys, bits, mask, a)) hits++;
      const unsigned long long mu = __ballot(a == u);
      if (mu) p = base + static_cast<uint32_t>(__builtin_ctzll(mu));
    }
    const int c = wsum(hits);
    if (lane == 0) {
      s += static_cast<unsigned long long>(c);
      if (p < ve) store_pair(cnt, un.j0 + k, p, c, plus);
    }
  }
  block_add_sum(sum, s);
}

// Hub units: persistent workgroups; workgroup w owns bitmap w of the pool
// (words = ceil(n/32) each, zero between units) and takes units w, w+G, ...
__global__ __launch_bounds__(kWG) void k_bmp_hub(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                 const FwdUnit *__restrict__ units, uint32_t n_units,
                                                 uint32_t *__restrict__ pool, uint64_t words, int32_t *cnt, int plus,
                                                 unsigned long long *sum) {
  __shared__ uint32_t s_cb[kChunk], s_ce[kChunk];
  uint32_t *bm = pool + static_cast<uint64_t>(blockIdx.x) * words;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  unsigned long long s = 0;
  for (uint32_t ui = blockIdx.x; ui < n_units; ui += gridDim.x) {
    const FwdUnit un = units[ui];
    const int32_t u = un.b;
    const uint32_t su = off[u], eu = off[u + 1];
    const uint32_t nup = un.j1 - un.j0;
    for (uint32_t i = threadIdx.x; i < nup; i += kWG) {
      const int32_t v = adj[un.j0 + i];
      s_cb[i] = off[v];
      s_ce[i] = off[v + 1];
    }
    for (uint32_t i = su + threadIdx.x; i < eu; i += kWG) {
      const int32_t a = adj[i];
      atomicOr(&bm[a >> 5], 1u << (a & 31));
    }
    __threadfence();
    __syncthreads();
    for (uint32_t k = wave; k < nup; k += kWaves) {
      const uint32_t vb = __builtin_amdgcn_readfirstlane(s_cb[k]);
      const uint32_t ve = __builtin_amdgcn_readfirstlane(s_ce[k]);
      int hits = 0;
      uint32_t p = ~0u;
      for (uint32_t base = vb; base < ve; base += 64) {
        const uint32_t i = base + lane;
        const int32_t a = i < ve ? adj[i] : -1;
        if (a >= 0) {
          // atomic load: the bitmap was written by other waves through L2
          const uint32_t w = __hip_atomic_load(&bm[a >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          hits += (w >> (a & 31)) & 1u;
        }
        const unsigned long long mu = __ballot(a == u);
        if (mu) p = base + static_cast<uint32_t>(__builtin_ctzll(mu));
      }
      const int c = wsum(hits);
      if (lane == 0) {
        s += static_cast<unsigned long long>(c);
        if (p < ve) store_pair(cnt, un.j0 + k, p, c, plus);
      }
    }
    __syncthreads();
    for (uint32_t i = su + threadIdx.x; i < eu; i += kWG)
      __hip_atomic_store(&bm[adj[i] >> 5], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    __syncthreads();
  }
  block_add_sum(sum, s);
}

// ---------------------------------------------------------------- MERGE path
// One wave per upper slot of the unit list; 8 waves per workgroup round-robin.
__global__ __launch_bounds__(kWG) void k_merge(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                               const FwdUnit *__restrict__ units, int32_t *cnt, int plus,
                                               unsigned long long *sum) {
  __shared__ int32_t win[kWaves][64];
  const FwdUnit un = units[blockIdx.x];
  const int32_t u = un.b;
  const uint32_t ub = off[u], ue = off[u + 1];
  const uint32_t nup = un.j1 - un.j0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  int32_t *w = win[wave];
  unsigned long long s = 0;
  for (uint32_t k = wave; k < nup; k += kWaves) {
    const uint32_t j = un.j0 + k;
    const int32_t v = adj[j];
    const uint32_t vb = off[v], ve = off[v + 1];
    // A: shorter list, B: longer list
    uint32_t ab = ub, ae = ue, bb = vb, be = ve;
    if (ae - ab > be - bb) {
      ab = vb;
      ae = ve;
      bb = ub;
      be = ue;
    }
    const uint32_t la = ae - ab, lb = be - bb;
    int hits = 0;
    if (la == 0) {
      // no common neighbor
    } else if (lb > kSkew * la) {
      // skewed: binary search of each A element in B
      for (uint32_t i0 = ab; i0 < ae; i0 += 64) {
        const uint32_t i = i0 + lane;
        if (i < ae) {
          const int32_t x = adj[i];
          const uint32_t q = find_pos(adj, bb, be, x);
          hits += (q < be && adj[q] == x) ? 1 : 0;
        }
      }
    } else {
      // windowed merge: A chunk in registers, B window staged in LDS
      uint32_t ia = ab, ib = bb;
      while (ia < ae && ib < be) {
        const int32_t x = ia + lane < ae ? adj[ia + lane] : INT32_MAX;
        const int32_t y = ib + lane < be ? adj[ib + lane] : INT32_MAX;
        w[lane] = y;
        const uint32_t na = ae - ia < 64 ? ae - ia : 64;
        const uint32_t nb = be - ib < 64 ? be - ib : 64;
        const int32_t amax = __shfl(x, static_cast<int>(na - 1), 64);
        const int32_t bmax = __shfl(y, static_cast<int>(nb - 1), 64);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (x != INT32_MAX) {
          int lo = 0, hi = static_cast<int>(nb);
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (w[mid] < x)
              lo = mid + 1;
            else
              hi = mid;
          }
          hits += (lo < static_cast<int>(nb) && w[lo] == x) ? 1 : 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int32_t am = __builtin_amdgcn_readfirstlane(amax);
        const int32_t bm = __builtin_amdgcn_readfirstlane(bmax);
        if (am <= bm) ia += 64;
        if (bm <= am) ib += 64;
      }
    }
    const int c = wsum(hits);
    const uint32_t p = find_pos(adj, vb, ve, u);
    if (lane == 0) {
      s += static_cast<unsigned long long>(c);
      if (p < ve) store_pair(cnt, j, p, c, plus);
    }
  }
  block_add_sum(sum, s);
}

__global__ void k_zero_sum(unsigned long long *sum) {
  if (threadIdx.x == 0) *sum = 0;
}

}  // namespace

namespace accpaths {

int launch_bitmap(acc_graph *g, int plus) {
  hipStream_t st = g->ctx->stream;
  hipLaunchKernelGGL(k_zero_sum, dim3(1), dim3(64), 0, st, g->d_sum);
  if (g->n_bmp_lds) {
    hipLaunchKernelGGL(k_bmp_lds, dim3(g->n_bmp_lds), dim3(kWG), 0, st, g->d_off, g->d_adj, g->d_bmp_lds, g->d_cnt,
                       plus, g->d_sum);
    ACC_HIP_TRY(hipGetLastError());
  }
  if (g->n_bmp_hub) {
    hipLaunchKernelGGL(k_bmp_hub, dim3(g->bmp_pool_slots), dim3(kWG), 0, st, g->d_off, g->d_adj, g->d_bmp_hub,
                       g->n_bmp_hub, g->d_bmp_pool, g->bmp_pool_words, g->d_cnt, plus, g->d_sum);
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

int launch_merge(acc_graph *g, int plus) {
  hipStream_t st = g->ctx->stream;
  hipLaunchKernelGGL(k_zero_sum, dim3(1), dim3(64), 0, st, g->d_sum);
  if (g->n_edge_units) {
    hipLaunchKernelGGL(k_merge, dim3(g->n_edge_units), dim3(kWG), 0, st, g->d_off, g->d_adj, g->d_edge_units,
                       g->d_cnt, plus, g->d_sum);
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

}  // namespace accpaths
