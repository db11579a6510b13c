// Forward-path kernels for gfx950 (wave64).
//
// With vertices in degree-descending order, every triangle {a < b < c} is
// found exactly once: at source b, probing c in N>(b), element a in
// N(c) ∩ [0, b) that also lies in N<(b).  Its three edges receive one count
// each through three roles:
//   (b,c)  summed in a register per c, stored into the upper slot (b,c);
//   (b,a)  summed in an LDS counter per position of a in N<(b), flushed with
//          one global atomic per counter into the lower slot (b,a);
//   (c,a)  one global atomic into the lower slot (c,a), the probe element's
//          own slot.
// Upper slots therefore receive only plain stores and lower slots only
// atomics.  Scanning N(c) up to b also yields the reverse slot of (b,c)
// (the position of b in N(c)), which k_combine uses to add the two halves
// and write 2 + count into both slots.  The counts equal the reference's
// merge counts (scan-query/scan_xp.cpp:675-857) for every slot.

#include "acc_internal.h"

namespace {

using accfwd::kChunk;
using accfwd::kLdsCap;
using accfwd::kTableBits;
using accfwd::kWG;

__device__ __forceinline__ uint32_t hash_slot(int32_t a, int bits) {
  return (static_cast<uint32_t>(a) * 0x9E3779B1u) >> (32 - bits);
}

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_init(int32_t *__restrict__ cnt, uint32_t *__restrict__ rev,
                                              uint64_t m, unsigned long long *__restrict__ sum) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t stride = gridDim.x * 256ull;
  const uint64_t m4 = m / 4;
  const int4 z = make_int4(0, 0, 0, 0);
  const uint4 f = make_uint4(~0u, ~0u, ~0u, ~0u);
  for (uint64_t i = tid; i < m4; i += stride) {
    reinterpret_cast<int4 *>(cnt)[i] = z;
    reinterpret_cast<uint4 *>(rev)[i] = f;
  }
  for (uint64_t i = m4 * 4 + tid; i < m; i += stride) {
    cnt[i] = 0;
    rev[i] = ~0u;
  }
  if (tid == 0) *sum = 0;
}

// One workgroup per unit (source b, upper slots [j0, j1)).  kLds: N<(b) in an
// LDS hash table (|N<(b)| <= kLdsCap); otherwise binary search of N<(b) in
// global memory and direct atomics for the (b,a) role.
template <bool kLds>
__global__ __launch_bounds__(kWG) void k_forward_count(const uint32_t *__restrict__ off,
                                                       const int32_t *__restrict__ adj,
                                                       const FwdUnit *__restrict__ units,
                                                       int32_t *cnt, uint32_t *__restrict__ rev) {
  constexpr int kT = kLds ? (1 << kTableBits) : 1;
  __shared__ int32_t keys[kT];
  __shared__ uint16_t kpos[kT];
  __shared__ int32_t ctr[kLds ? kLdsCap : 1];

  const FwdUnit u = units[blockIdx.x];
  const int32_t b = u.b;
  const uint32_t sb = off[b];
  const uint32_t nl = u.nlow;
  const int tid = threadIdx.x;
  int bits = 6;
  while ((1u << bits) < 2 * nl) bits++;
  const uint32_t mask = (1u << bits) - 1;

  if (kLds) {
    for (uint32_t i = tid; i <= mask; i += kWG) keys[i] = -1;
    for (uint32_t i = tid; i < nl; i += kWG) ctr[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < nl; i += kWG) {
      const int32_t a = adj[sb + i];
      uint32_t h = hash_slot(a, bits);
      while (atomicCAS(&keys[h], -1, a) != -1) h = (h + 1) & mask;
      kpos[h] = static_cast<uint16_t>(i);
    }
    __syncthreads();
  }

  const int wave = tid >> 6;
  const int lane = tid & 63;
  for (uint32_t j = u.j0 + wave; j < u.j1; j += kWG / 64) {
    const int32_t c = adj[j];
    const uint32_t cb = off[c];
    const uint32_t ce = off[c + 1];
    int reg = 0;
    uint32_t p = ~0u;
    for (uint32_t base = cb;; base += 64) {
      const uint32_t idx = base + lane;
      const int32_t a = idx < ce ? adj[idx] : INT32_MAX;
      const bool lt = a < b;
      if (lt) {
        int pos = -1;
        if (kLds) {
          uint32_t h = hash_slot(a, bits);
          for (;;) {
            const int32_t k = keys[h];
            if (k == a) {
              pos = kpos[h];
              break;
            }
            if (k == -1) break;
            h = (h + 1) & mask;
          }
        } else {
          uint32_t lo = 0, hi = nl;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (adj[sb + mid] < a)
              lo = mid + 1;
            else
              hi = mid;
          }
          if (lo < nl && adj[sb + lo] == a) pos = static_cast<int>(lo);
        }
        if (pos >= 0) {
          reg++;
          atomicAdd(&cnt[idx], 1);
          if (kLds)
            atomicAdd(&ctr[pos], 1);
          else
            atomicAdd(&cnt[sb + pos], 1);
        }
      }
      const unsigned long long mlt = __ballot(lt);
      if (mlt != ~0ull) {
        p = base + static_cast<uint32_t>(__builtin_ctzll(~mlt));
        if (p >= ce) p = ~0u;  // b absent from N(c): only possible for an unvalidated, asymmetric CSR
        break;
      }
    }
    reg = wave_sum_i32(reg);
    if (lane == 0) {
      cnt[j] = reg;
      rev[j] = p;
    }
  }
  if (kLds) {
    __syncthreads();
    for (uint32_t i = tid; i < nl; i += kWG) {
      const int v = ctr[i];
      if (v) atomicAdd(&cnt[sb + i], v);
    }
  }
}

__global__ __launch_bounds__(256) void k_combine(int32_t *__restrict__ cnt, const uint32_t *__restrict__ rev,
                                                 uint64_t m, int plus, unsigned long long *__restrict__ sum) {
  __shared__ unsigned long long part[4];
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t stride = gridDim.x * 256ull;
  unsigned long long s = 0;
  for (uint64_t i = tid; i < m; i += stride) {
    const uint32_t p = rev[i];
    if (p < m) {
      int c = cnt[i] + cnt[p];
      s += static_cast<unsigned long long>(c);
      c += plus;
      cnt[i] = c;
      cnt[p] = c;
    }
  }
  s = wave_sum_u64(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(sum, t);
  }
}

// Flags: 1 id out of range, 2 self-loop, 4 unsorted, 8 asymmetric.
__global__ __launch_bounds__(256) void k_validate(const uint32_t *__restrict__ off,
                                                  const int32_t *__restrict__ adj, uint32_t n, uint64_t m,
                                                  int32_t *__restrict__ flags) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t stride = gridDim.x * 256ull;
  int32_t f = 0;
  for (uint64_t i = tid; i < m; i += stride) {
    uint32_t lo = 0, hi = n;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if (off[mid] <= i)
        lo = mid;
      else
        hi = mid;
    }
    const int32_t v = adj[i];
    if (v < 0 || static_cast<uint32_t>(v) >= n) {
      f |= 1;
      continue;
    }
    if (static_cast<uint32_t>(v) == lo) f |= 2;
    if (i > off[lo] && adj[i - 1] >= v) f |= 4;
    uint32_t a = off[v], e = off[v + 1];
    while (a < e) {
      const uint32_t mid = a + ((e - a) >> 1);
      if (adj[mid] < static_cast<int32_t>(lo))
        a = mid + 1;
      else
        e = mid;
    }
    if (a >= off[v + 1] || adj[a] != static_cast<int32_t>(lo)) f |= 8;
  }
  if (f) atomicOr(flags, f);
}

inline unsigned grid_for(uint64_t items, int cus) {
  const uint64_t want = (items + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(cus > 0 ? cus : 256) * 8;
  return static_cast<unsigned>(want < cap ? (want ? want : 1) : cap);
}

}  // namespace

namespace accfwd {

int launch_init(acc_graph *g) {
  hipLaunchKernelGGL(k_init, dim3(grid_for(g->m / 4 + 1, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream,
                     g->d_cnt, g->d_rev, g->m, g->d_sum);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_count(acc_graph *g) {
  if (g->n_fwd_small) {
    hipLaunchKernelGGL(k_forward_count<true>, dim3(g->n_fwd_small), dim3(kWG), 0, g->ctx->stream, g->d_off,
                       g->d_adj, g->d_fwd_small, g->d_cnt, g->d_rev);
    ACC_HIP_TRY(hipGetLastError());
  }
  if (g->n_fwd_big) {
    hipLaunchKernelGGL(k_forward_count<false>, dim3(g->n_fwd_big), dim3(kWG), 0, g->ctx->stream, g->d_off,
                       g->d_adj, g->d_fwd_big, g->d_cnt, g->d_rev);
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

int launch_combine(acc_graph *g, int plus) {
  hipLaunchKernelGGL(k_combine, dim3(grid_for(g->m, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream, g->d_cnt,
                     g->d_rev, g->m, plus, g->d_sum);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_validate(const uint32_t *d_off, const int32_t *d_adj, uint32_t n, uint64_t m, int32_t *d_flags,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_validate, dim3(grid_for(m, 256)), dim3(256), 0, s, d_off, d_adj, n, m, d_flags);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

}  // namespace accfwd
