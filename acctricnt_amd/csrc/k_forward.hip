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
// atomics.  The reverse slot of every upper slot (b,c), the position of b in
// N(c), is computed once per resident graph (at upload); it bounds the
// probe of N(c) to exactly |N(c) ∩ [0, b)| elements, which lets a wave or a
// workgroup flatten all probes of a source into one element range with no
// idle lane, and k_combine uses it to add the two halves and write
// 2 + count into both slots.  The counts equal the reference's merge counts
// (scan-query/scan_xp.cpp:675-857) for every slot.

#include "acc_internal.h"
#include "flat_window.h"

namespace {

using accfwd::kChunk;
using accfwd::kLdsCap;
using accfwd::kTableBits;
using accfwd::kBigWG;
using accfwd::kWG;
using accfwd::kWaves;

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

// Zeroes the counts, the triangle sum and the per-graph totals of a
// partitioned graph (parts of them; gtot may be NULL when parts is 0).
__global__ __launch_bounds__(256) void k_init(int32_t *__restrict__ cnt, uint64_t m,
                                              unsigned long long *__restrict__ sum,
                                              unsigned long long *__restrict__ gtot, uint32_t parts) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t stride = gridDim.x * 256ull;
  const uint64_t m4 = m / 4;
  const int4 z = make_int4(0, 0, 0, 0);
  for (uint64_t i = tid; i < m4; i += stride) reinterpret_cast<int4 *>(cnt)[i] = z;
  for (uint64_t i = m4 * 4 + tid; i < m; i += stride) cnt[i] = 0;
  for (uint64_t k = tid; k < parts; k += stride) gtot[k] = 0;
  if (tid == 0) *sum = 0;
}

// Forward work estimate of every source b: over its upper slots (b, c) the
// probe length |N(c) ∩ [0, b)| (the reverse slot minus the start of N(c))
// plus 4 for the slot's fixed cost plus twice the probe's hits (hits[j], the
// triangles with middle b and top c: the upper slot's value after a
// counting pass without the combine), and per unit of `chunk` upper slots
// the build of the N<(b) table plus 256.  The weights follow a least-squares
// fit of the kernel times of 8 shares of R-MAT scale 22 (a hit weighs about
// 1.8 probe elements, a unit's table build well under its two earlier
// elements per entry).  One wave per source.
__global__ __launch_bounds__(256) void k_source_work(const uint32_t *__restrict__ off, const uint32_t *__restrict__ hs,
                                                     const int32_t *__restrict__ adj,
                                                     const uint32_t *__restrict__ rev, const int32_t *__restrict__ hits,
                                                     uint32_t n, uint32_t chunk,
                                                     unsigned long long *__restrict__ work) {
  const uint32_t waves = gridDim.x * 4u;
  const int lane = threadIdx.x & 63;
  for (uint32_t b = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6)); b < n; b += waves) {
    unsigned long long w = 0;
    for (uint32_t j = hs[b] + lane; j < off[b + 1]; j += 64) {
      const uint32_t p = rev[j];
      w += 4 + (p != ~0u ? p - off[adj[j]] : 0) + 2ull * static_cast<uint32_t>(hits[j]);
    }
    w = wave_sum_u64(w);
    const uint32_t nup = off[b + 1] - hs[b], nl = hs[b] - off[b];
    if (lane == 0) work[b] = w + static_cast<unsigned long long>((nup + chunk - 1) / chunk) * (nl + 256u);
  }
}

// Plan inputs of every vertex v: hs[v], the first slot of N(v) holding an id
// above v (the start of its upper slots), and lo[v] = min(N(v) ∪ {v}).
// maxes (zeroed first) gets max |N<(v)| and the maximum degree.
__global__ __launch_bounds__(256) void k_vertex_meta(const uint32_t *__restrict__ off,
                                                     const int32_t *__restrict__ adj, uint32_t n,
                                                     uint32_t *__restrict__ hs, int32_t *__restrict__ lo,
                                                     uint32_t *__restrict__ maxes) {
  uint32_t max_low = 0, max_deg = 0;
  for (uint32_t v = blockIdx.x * 256u + threadIdx.x; v < n; v += gridDim.x * 256u) {
    const uint32_t s = off[v], e = off[v + 1];
    uint32_t a = s, b = e;
    while (a < b) {
      const uint32_t mid = a + ((b - a) >> 1);
      if (adj[mid] <= static_cast<int32_t>(v)) a = mid + 1; else b = mid;
    }
    hs[v] = a;
    lo[v] = e > s ? min(adj[s], static_cast<int32_t>(v)) : static_cast<int32_t>(v);
    max_low = max(max_low, a - s);
    max_deg = max(max_deg, e - s);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    max_low = max(max_low, static_cast<uint32_t>(__shfl_xor(static_cast<int>(max_low), o, 64)));
    max_deg = max(max_deg, static_cast<uint32_t>(__shfl_xor(static_cast<int>(max_deg), o, 64)));
  }
  if ((threadIdx.x & 63) == 0) {
    if (max_low) atomicMax(&maxes[0], max_low);
    if (max_deg) atomicMax(&maxes[1], max_deg);
  }
}

// Reverse slots by binary search, used at upload only for a CSR trusted
// without validation whose upper slots outnumber its lower ones (the sort
// of the upper slots by target does the rest): rev[j] = position of u in
// N(v) for every upper slot j = (u, v), u < v (binary search of N(v)), ~0
// for lower slots and for an upper slot whose reverse is absent (possible
// only in an unvalidated, asymmetric CSR).  One thread per slot, the source
// u of slot j read from src.  With flags (validation): flag 8 for an upper
// slot without reverse, and the number of upper slots added to *upper.  In a
// CSR with sorted, duplicate-free lists and no self-loop, every upper slot
// having a reverse maps the upper slots one-to-one into the lower ones, so
// the CSR is symmetric iff additionally 2 * upper == m.
__global__ __launch_bounds__(256) void k_build_rev(const uint32_t *__restrict__ off,
                                                   const int32_t *__restrict__ adj,
                                                   const uint32_t *__restrict__ src, uint32_t n, uint64_t m,
                                                   uint32_t *__restrict__ rev, int32_t *__restrict__ flags,
                                                   unsigned long long *__restrict__ upper) {
  int32_t f = 0;
  unsigned long long up = 0;
  for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < m; j += gridDim.x * 256ull) {
    const int32_t v = static_cast<int32_t>(src[j]);
    const int32_t c = adj[j];
    uint32_t p = ~0u;
    if (c > v && static_cast<uint32_t>(c) < n) {
      uint32_t lo = off[c];
      const uint32_t end = off[c + 1];
      uint32_t hi = end;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (adj[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < end && adj[lo] == v) p = lo;
      up++;
      if (p == ~0u) f |= 8;
    }
    rev[j] = p;
  }
  if (flags) {
    up = wave_sum_u64(up);
    if (f) atomicOr(flags, f);
    if ((threadIdx.x & 63) == 0 && up) atomicAdd(upper, up);
  }
}

// Makes the wave's earlier LDS writes visible to all of its lanes before
// later LDS reads (wave-private tables need no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int table_bits_for(uint32_t nl) {
  int bits = 6;
  while ((1u << bits) < 2 * nl) bits++;
  return bits;
}

// Inserts N<(b) = adj[sb, sb+nl) with a stride of `step` threads starting at `t`.
__device__ __forceinline__ void lds_insert(int32_t *keys, uint16_t *kpos, int bits, uint32_t mask,
                                           const int32_t *__restrict__ adj, uint32_t sb, uint32_t nl, int t,
                                           int step) {
  for (uint32_t i = t; i < nl; i += step) {
    const int32_t a = adj[sb + i];
    uint32_t h = hash_slot(a, bits);
    while (atomicCAS(&keys[h], -1, a) != -1) h = (h + 1) & mask;
    kpos[h] = static_cast<uint16_t>(i);
  }
}

// Membership of N<(b) as an open-addressing hash table in LDS: key = vertex
// id, value (kpos) = its position in N<(b).  A round's U elements (a[q] < 0:
// no element) are looked up together: the first probe of every row is read
// before any row is decided, so the LDS reads overlap, and only a row whose
// first probe meets another key walks on.
struct HashSet {
  const int32_t *keys;
  const uint16_t *kpos;
  int bits;
  uint32_t mask;
  template <int U>
  __device__ __forceinline__ void walk_rows(const int32_t (&a)[U], uint32_t (&h)[U], bool (&hit)[U]) const {
    int32_t k[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      h[q] = hash_slot(a[q], bits);
      k[q] = keys[h[q]];
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      bool found = a[q] >= 0 && k[q] == a[q];
      bool more = a[q] >= 0 && k[q] != a[q] && k[q] != -1;
      while (more) {
        h[q] = (h[q] + 1) & mask;
        const int32_t kk = keys[h[q]];
        found = kk == a[q];
        more = !found && kk != -1;
      }
      hit[q] = found;
    }
  }
  template <int U>
  __device__ __forceinline__ void contains_rows(const int32_t (&a)[U], bool (&hit)[U]) const {
    uint32_t h[U];
    walk_rows(a, h, hit);
  }
  // the position of one element a in N<(b), -1 when absent
  __device__ __forceinline__ int find(int32_t a) const {
    const int32_t a1[1] = {a};
    uint32_t h[1];
    bool hit[1];
    walk_rows(a1, h, hit);
    return hit[0] ? static_cast<int>(kpos[h[0]]) : -1;
  }
};

// Membership of N<(b) ⊂ [lo, b), lo = min N(b), as an exact LDS bitmap over
// [lo, b) with the position of a in N<(b) recovered from the index of the
// first element of a's 32-bit word (pre) plus a popcount.  A round's U
// elements (a[q] < 0: no element, and then r < 0 since lo >= 0) are looked up
// together with no branch, so the LDS reads overlap; a row without an element
// reads word 0.  Callers only ask for a < b.
struct BitSet {
  const uint32_t *bm;
  const uint16_t *pre;
  int32_t lo;
  template <int U>
  __device__ __forceinline__ void contains_rows(const int32_t (&a)[U], bool (&hit)[U]) const {
    uint32_t w[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const int32_t r = a[q] - lo;
      w[q] = bm[r >= 0 ? (r >> 5) : 0];
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      const int32_t r = a[q] - lo;
      hit[q] = r >= 0 && ((w[q] >> (r & 31)) & 1u);
    }
  }
  // the position of one element a in N<(b), -1 when absent (both words
  // read with no branch)
  __device__ __forceinline__ int find(int32_t a) const {
    const int32_t r = a - lo;
    const int32_t wi = r >= 0 ? (r >> 5) : 0;
    const uint32_t w = bm[wi], p = pre[wi];
    const uint32_t bit = 1u << (r & 31);
    return r >= 0 && (w & bit) ? static_cast<int>(p + __popc(w & (bit - 1))) : -1;
  }
};

// Exclusive prefix sum of v over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, int lane) {
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

// Atomic add on an LDS word through an LDS-qualified pointer, so that it
// is always a DS instruction: with two plain atomicAdds on either side of a
// branch (one on LDS, one on global memory) the compiler selects between
// the two pointers and issues one FLAT atomic for both.
typedef __attribute__((address_space(3))) int32_t lds_int32;
__device__ __forceinline__ void lds_atomic_add(int32_t *p, int v) {
  __hip_atomic_fetch_add((lds_int32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Records one triangle {a, b, c} found at element slot idx of N(c): the
// (c,a) slot itself and the (b,a) counter of a's position pos in N<(b)
// (LDS below ctr_cap, else an atomic on the lower slot sb+pos).
__device__ __forceinline__ void count_triangle(int32_t *cnt, int32_t *ctr, uint32_t ctr_cap, uint32_t sb,
                                               uint32_t idx, int pos) {
  atomicAdd(&cnt[idx], 1);
  if (static_cast<uint32_t>(pos) < ctr_cap)
    lds_atomic_add(&ctr[pos], 1);
  else
    atomicAdd(&cnt[sb + pos], 1);
}

// Adds the round's hits to the (b,c_k) counters reg[k] (LDS): one add of the
// ballot's popcount when the whole round lies in one probe, else one LDS
// atomic per hit.  (A per-segment add, one per probe of the round, measured
// no faster here: the (c,a) updates, not these counters, bound the pass.)
__device__ __forceinline__ void count_round(int32_t *reg, int k, bool hit, int lane) {
  const unsigned long long hits = __ballot(hit);
  if (!hits) return;
  const int k0 = __builtin_amdgcn_readfirstlane(k);
  if (k0 == __builtin_amdgcn_readlane(k, 63)) {
    if (lane == 0) atomicAdd(&reg[k0], static_cast<int>(__popcll(hits)));
  } else if (hit) {
    atomicAdd(&reg[k], 1);
  }
}

// Partial triangle totals of the triangle-only pass: kTriParts counters one
// 128-byte line apart, workgroup w adding into counter w mod kTriParts, so
// the atomics of many small units do not queue on one address.
constexpr int kTriParts = 256;
constexpr int kTriStride = 16;  // unsigned long longs per 128-byte line
__device__ __forceinline__ unsigned long long *tri_part(unsigned long long *parts) {
  return parts + (blockIdx.x % kTriParts) * kTriStride;
}

constexpr int kWT = 1 << accfwd::kWaveTableBits;
constexpr int kBT = 1 << kTableBits;
constexpr size_t kBlockLdsBytes = 4 * kBT + 2 * kBT + 4 * kLdsCap + 12 * kChunk + 4 * kWaves;
constexpr size_t kWaveLdsBytes = kWaves * (4 * kWT + 2 * kWT + 4 * accfwd::kWaveCap + 4 * 64);
constexpr size_t kCountLdsBytes = kBlockLdsBytes > kWaveLdsBytes ? kBlockLdsBytes : kWaveLdsBytes;

// Workgroup-flattened probing of a chunk's nup <= kChunk upper slots: s_P
// (kChunk entries, T beyond nup) holds the exclusive prefix of the probe
// lengths and s_cb the start of each N(c_k) less its prefix, so that element
// e of probe k is slot s_cb[k] + e (mod 2^32).  Wave w takes the rounds
// w, w + kWaves, ... of 64 * U consecutive elements, so the chunk's work is
// split evenly over the waves whatever the list lengths; k(e) comes from a
// per-round window search (flat_window.h).  Round indices are wave-uniform.
template <int U, bool kSlots, class Set>
__device__ __forceinline__ uint32_t block_probe_flat(const int32_t *__restrict__ adj, int32_t *cnt, const Set &set,
                                                     int32_t *ctr, uint32_t sb, const uint32_t *s_P,
                                                     const uint32_t *s_cb, uint32_t T, int32_t *s_reg, int wave,
                                                     int lane) {
  static_assert(kChunk == 64 * 8, "the window search covers 64 groups of 8 entries");
  constexpr uint32_t kStride = kWaves * 64 * U;
  uint32_t found = 0;
  uint32_t e0 = static_cast<uint32_t>(wave) * 64 * U;
  if constexpr (kSlots) {
    for (; e0 < T; e0 += kStride) {
      int32_t a[U];
      uint32_t idx[U];
      int kq[U];
      accflat::lds_search<U>(s_P, e0, accflat::round_last<U>(e0, T), lane, kq);
#pragma unroll
      for (int q = 0; q < U; q++) {
        const uint32_t e = e0 + 64 * q + lane;
        idx[q] = s_cb[kq[q]] + e;
        a[q] = e < T ? adj[idx[q]] : -1;
      }
      // one row at a time: looking the U rows up together, as the
      // triangle-only pass does, made the persistent grid 1.7x slower here
#pragma unroll
      for (int q = 0; q < U; q++) {
        const int pos = set.find(a[q]);
        const bool hit = pos >= 0;
        if (hit) count_triangle(cnt, ctr, kLdsCap, sb, idx[q], pos);
        count_round(s_reg, kq[q], hit, lane);
      }
    }
    return 0;
  } else {
    // Triangle-only: the next round's elements are loaded (unconditionally:
    // the last round re-reads its own, lanes past T read the round's last
    // element) before this round's lookups, so the loads of one round are
    // in flight while the previous round is decided.  (In the per-slot pass
    // the same pipelining was slower: its count updates, issued between the
    // loads, are waited for with them.)
    if (e0 >= T) return 0;
    struct Round {
      int32_t v[U];  // element id, -1 past T
    };
    auto fetch = [&](uint32_t f0, Round &r) {
      const uint32_t f1 = accflat::round_last<U>(f0, T);
      int k[U];
      accflat::lds_search<U>(s_P, f0, f1, lane, k);
#pragma unroll
      for (int q = 0; q < U; q++) {
        const uint32_t e = f0 + 64 * q + lane;
        const int32_t x = adj[s_cb[k[q]] + (e < T ? e : f1)];  // past T: probe k(f1), element f1
        r.v[q] = e < T ? x : -1;
      }
    };
    Round cur, nxt;
    fetch(e0, cur);
    for (;;) {
      const uint32_t e1 = e0 + kStride;
      const bool more = e1 < T;  // wave-uniform
      fetch(more ? e1 : e0, nxt);
      bool hit[U];
      set.contains_rows(cur.v, hit);
#pragma unroll
      for (int q = 0; q < U; q++) found += hit[q];
      if (!more) break;
      e0 = e1;
      cur = nxt;
    }
    return found;
  }
}

// Workgroup unit: one chunk of at most kChunk upper slots of a source b.
// N<(b) is an LDS bitmap when b - min N(b) <= kBitmapIds (|N<(b)| < 65536),
// otherwise an LDS hash table (|N<(b)| <= kLdsCap).  Both layouts share the
// counter and probe-metadata regions of smem.
// kSlots false: count the unit's triangles into *total instead of writing slots.
template <int U, bool kSlots>
__device__ __forceinline__ void fwd_block_unit(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                               const uint32_t *__restrict__ rev, const FwdUnit u, int32_t *cnt,
                                               unsigned long long *total, unsigned char *smem) {
  int32_t *keys = reinterpret_cast<int32_t *>(smem);  // hash keys, or bitmap words
  int32_t *ctr = keys + kBT;
  uint32_t *s_cb = reinterpret_cast<uint32_t *>(ctr + kLdsCap);
  uint32_t *s_P = s_cb + kChunk;
  int32_t *s_reg = reinterpret_cast<int32_t *>(s_P + kChunk);
  uint32_t *s_wsum = reinterpret_cast<uint32_t *>(s_reg + kChunk);
  uint16_t *kpos = reinterpret_cast<uint16_t *>(s_wsum + kWaves);  // hash positions, or word prefixes
  const int32_t b = u.b;
  const uint32_t sb = off[b];
  const uint32_t nl = u.nlow;
  const uint32_t nup = u.j1 - u.j0;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int32_t lo = nl ? adj[sb] : b;
  const uint32_t span = static_cast<uint32_t>(b - lo);
  const bool bitmap = span <= accfwd::kBitmapIds;
  const uint32_t nctr = nl < kLdsCap ? nl : kLdsCap;
  const int bits = bitmap ? 0 : table_bits_for(nl);
  const uint32_t mask = bitmap ? 0 : (1u << bits) - 1;
  const uint32_t nclear = bitmap ? (span + 31) / 32 : mask + 1;
  const int32_t fill = bitmap ? 0 : -1;
  for (uint32_t i = tid; i < nclear; i += kWG) keys[i] = fill;
  if (kSlots)
    for (uint32_t i = tid; i < nctr; i += kWG) ctr[i] = 0;
  // probe lengths L_t = |N(c_t) ∩ [0, b)| = rev - start of N(c_t), and their prefix
  uint32_t len = 0, cb = 0;
  if (static_cast<uint32_t>(tid) < nup) {
    const int32_t c = adj[u.j0 + tid];
    cb = off[c];
    const uint32_t p = rev[u.j0 + tid];
    len = p != ~0u ? p - cb : 0;
  }
  s_reg[tid] = 0;
  const uint32_t excl = wave_excl_scan(len, lane);
  if (lane == 63) s_wsum[wave] = excl + len;
  __syncthreads();
  uint32_t before = 0, T = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    const uint32_t ws = s_wsum[w];
    before += w < wave ? ws : 0;
    T += ws;
  }
  s_P[tid] = static_cast<uint32_t>(tid) < nup ? before + excl : T;
  if (static_cast<uint32_t>(tid) < nup) s_cb[tid] = cb - (before + excl);
  if (bitmap) {
    uint32_t *bm = reinterpret_cast<uint32_t *>(keys);
    for (uint32_t i = tid; i < nl; i += kWG) {
      const int32_t r = adj[sb + i] - lo;
      atomicOr(&bm[r >> 5], 1u << (r & 31));
      if (i == 0 || ((adj[sb + i - 1] - lo) >> 5) != (r >> 5)) kpos[r >> 5] = static_cast<uint16_t>(i);
    }
  } else {
    lds_insert(keys, kpos, bits, mask, adj, sb, nl, tid, kWG);
  }
  __syncthreads();
  T = __builtin_amdgcn_readfirstlane(T);
  uint32_t found;
  if (bitmap) {
    const BitSet set{reinterpret_cast<const uint32_t *>(keys), kpos, lo};
    found = block_probe_flat<U, kSlots>(adj, cnt, set, ctr, sb, s_P, s_cb, T, s_reg, wave, lane);
  } else {
    const HashSet set{keys, kpos, bits, mask};
    found = block_probe_flat<U, kSlots>(adj, cnt, set, ctr, sb, s_P, s_cb, T, s_reg, wave, lane);
  }
  if (!kSlots) {
    const unsigned long long w = wave_sum_u64(found);
    if (lane == 0 && w) atomicAdd(tri_part(total), w);
    return;
  }
  __syncthreads();
  if (static_cast<uint32_t>(tid) < nup) cnt[u.j0 + tid] = s_reg[tid];
  for (uint32_t i = tid; i < nctr; i += kWG) {
    const int v = ctr[i];
    if (v) atomicAdd(&cnt[sb + i], v);
  }
}

// Largest k in [0, 64) with X_k <= e, for X nondecreasing across the lanes
// and X_0 <= e (a 6-step search over the lanes; every lane must call it).
__device__ __forceinline__ int lane_search(uint32_t X, uint32_t e) {
  int k = 0;
#pragma unroll
  for (int st = 32; st; st >>= 1)
    if (static_cast<uint32_t>(__shfl(X, k + st)) <= e) k += st;
  return k;
}

// Hash table size of one source of a wave pack: the smallest power of two
// >= 2 |N<(b)|, at least 16 entries.
__device__ __forceinline__ int pack_bits(uint32_t nl) {
  int bits = 4;
  while ((1u << bits) < 2 * nl) bits++;
  return bits;
}

// Wave pack: the consecutive sources [b0, b1) (at most 64), each with at
// most wave_slots upper slots and |N<(b)| <= wave_low, together at most 64
// upper slots, 256 lower entries and 512 hash entries (the plan guarantees
// these bounds).  One wave builds one LDS hash table per source, then
// flattens the probes of all the pack's upper slots into one element range:
// element e belongs to probe i = k(e), whose source's table decides it.
// Packing several small sources into one wave shares the latency of the
// metadata loads and the fixed work of a unit between them.
template <int U, bool kSlots>
__device__ __forceinline__ void fwd_pack_unit(const uint32_t *__restrict__ off, const uint32_t *__restrict__ hs,
                                              const int32_t *__restrict__ adj, const uint32_t *__restrict__ rev,
                                              const FwdUnit u, int32_t *cnt, unsigned long long *total,
                                              unsigned char *smem, int wave) {
  int32_t *base = reinterpret_cast<int32_t *>(smem);
  int32_t *keys = base + wave * kWT;
  int32_t *ctr = base + kWaves * kWT + wave * accfwd::kWaveCap;
  int32_t *reg = base + kWaves * (kWT + accfwd::kWaveCap) + wave * 64;
  uint16_t *kpos = reinterpret_cast<uint16_t *>(base + kWaves * (kWT + accfwd::kWaveCap + 64)) + wave * kWT;
  const int lane = threadIdx.x & 63;
  const int32_t b0 = u.b;
  const uint32_t nv = u.j0 - static_cast<uint32_t>(b0);
  const uint32_t nup_tot = u.j1;
  const uint32_t nl_tot = u.nlow;
  // one source per lane: lower slots [o0, h), upper slots [h, o1)
  uint32_t o0 = 0, h = 0, nupv = 0, nlv = 0, tsz = 0;
  int bits = 0;
  if (static_cast<uint32_t>(lane) < nv) {
    const int32_t v = b0 + lane;
    o0 = off[v];
    h = hs[v];
    nupv = off[v + 1] - h;
    if (nupv) {
      nlv = h - o0;
      if (nlv) {
        bits = pack_bits(nlv);
        tsz = 1u << bits;
      }
    }
  }
  const uint32_t Q = wave_excl_scan(nupv, lane);   // first probe of each source
  const uint32_t NL = wave_excl_scan(nlv, lane);   // first counter of each source
  const uint32_t TB = wave_excl_scan(tsz, lane);   // first table entry of each source
  const uint32_t tbl_tot = __builtin_amdgcn_readfirstlane(__shfl(TB + tsz, 63));
  // one probe per lane: its source, upper slot, N(c) start and length
  const int pk = lane_search(Q, static_cast<uint32_t>(lane));
  const uint32_t j = static_cast<uint32_t>(__shfl(h, pk)) + (static_cast<uint32_t>(lane) - __shfl(Q, pk));
  const uint32_t p_tb = __shfl(TB, pk);
  const int p_bits = __shfl(bits, pk);  // 0: the probe's source has an empty N<(b) and no table
  const uint32_t p_nl = __shfl(NL, pk);
  const bool probe = static_cast<uint32_t>(lane) < nup_tot;
  uint32_t cb = 0, len = 0;
  if (probe) {
    const int32_t c = adj[j];
    const uint32_t p = rev[j];
    cb = off[c];
    // a source with an empty N<(b) is the middle of no triangle: nothing to probe
    len = p != ~0u && p_bits > 0 ? p - cb : 0;
  }
  // the pack's lower entries, four per lane, with their source
  int32_t low[4];
  uint32_t lbase[4], lpos[4];
  int lbits[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t e = lane + 64u * r;
    const int k = lane_search(NL, e);
    const uint32_t nlk = __shfl(NL, k);
    const uint32_t o0k = __shfl(o0, k);
    lbase[r] = __shfl(TB, k);
    lbits[r] = __shfl(bits, k);
    lpos[r] = e - nlk;
    low[r] = e < nl_tot ? adj[o0k + lpos[r]] : -1;
  }
  for (uint32_t i = lane; i < tbl_tot; i += 64) keys[i] = -1;
  if (kSlots) {
    for (uint32_t i = lane; i < nl_tot; i += 64) ctr[i] = 0;
    reg[lane] = 0;
  }
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 4; r++) {
    if (low[r] >= 0) {
      const uint32_t mask = (1u << lbits[r]) - 1;
      uint32_t hh = hash_slot(low[r], lbits[r]);
      while (atomicCAS(&keys[lbase[r] + hh], -1, low[r]) != -1) hh = (hh + 1) & mask;
      kpos[lbase[r] + hh] = static_cast<uint16_t>(lpos[r]);
    }
  }
  const uint32_t P = wave_excl_scan(len, lane);
  const uint32_t T = __builtin_amdgcn_readfirstlane(__shfl(P + len, 63));
  // per probe, for one lane exchange each per element: element e of the
  // probe is slot cbP + e (mod 2^32), and its source's table start (< 512),
  // table bits (<= 9) and first counter (<= 256) packed into one word
  static_assert(accfwd::kWaveTableBits <= 9 && accfwd::kWaveCap <= 512, "the packing below needs 9 + 4 + 10 bits");
  const uint32_t cbP = cb - P;
  const uint32_t meta = p_tb | (static_cast<uint32_t>(p_bits) << 9) | (p_nl << 13);
  wave_lds_sync();
  uint32_t found = 0;
  for (uint32_t e0 = 0; e0 < T; e0 += 64 * U) {
    int32_t a[U];
    uint32_t idx[U];
    int kq[U];
    uint32_t eq[U];
#pragma unroll
    for (int q = 0; q < U; q++) eq[q] = e0 + 64 * q + lane;
    accflat::lane_window<U>(P, e0, accflat::round_last<U>(e0, T), lane, kq);
#pragma unroll
    for (int q = 0; q < U; q++) {
      idx[q] = static_cast<uint32_t>(__shfl(cbP, kq[q])) + eq[q];
      a[q] = eq[q] < T ? adj[idx[q]] : -1;
    }
    // the first probe of every row's table is read before any row is
    // decided, so the LDS reads overlap; a row without an element, or whose
    // source has no table, reads entry 0 of the wave's tables and misses
    uint32_t pm[U], h[U];
    int32_t k[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      pm[q] = static_cast<uint32_t>(__shfl(meta, kq[q]));
      const int tbits = static_cast<int>((pm[q] >> 9) & 15u);
      const bool ok = a[q] >= 0 && tbits > 0;
      h[q] = ok ? hash_slot(a[q], tbits) : 0;
      k[q] = keys[(pm[q] & 511u) + h[q]];
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      const uint32_t tb = pm[q] & 511u;
      const int tbits = static_cast<int>((pm[q] >> 9) & 15u);
      const bool ok = a[q] >= 0 && tbits > 0;
      bool hit = ok && k[q] == a[q];
      bool more = ok && k[q] != a[q] && k[q] != -1;
      while (more) {
        h[q] = (h[q] + 1) & ((1u << tbits) - 1);
        const int32_t kk = keys[tb + h[q]];
        hit = kk == a[q];
        more = !hit && kk != -1;
      }
      if constexpr (kSlots) {
        if (hit) {
          atomicAdd(&cnt[idx[q]], 1);
          atomicAdd(&ctr[(pm[q] >> 13) + kpos[tb + h[q]]], 1);
        }
        count_round(reg, kq[q], hit, lane);
      } else {
        found += hit;
      }
    }
  }
  if (!kSlots) {
    const unsigned long long w = wave_sum_u64(found);
    if (lane == 0 && w) atomicAdd(tri_part(total), w);
    return;
  }
  wave_lds_sync();
  if (probe) cnt[j] = reg[lane];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t e = lane + 64u * r;
    const int k = lane_search(NL, e);
    const uint32_t dst = static_cast<uint32_t>(__shfl(o0, k)) + (e - static_cast<uint32_t>(__shfl(NL, k)));
    if (e < nl_tot) {
      const int v = ctr[e];
      if (v) atomicAdd(&cnt[dst], v);
    }
  }
}

// The forward counting launch.  Tasks [0, n_block) are one workgroup unit
// each, the following tasks kWaves wave packs each; workgroup w runs tasks
// w, w + gridDim.x, ... (one task per workgroup when the grid covers them, a
// persistent grid otherwise).  The task index is workgroup-uniform, so the
// barrier that separates consecutive workgroup units' use of LDS is reached
// by all.  Wave packs use only their wave's part of LDS, so once its
// workgroup units are done each wave goes through its packs on its own, with
// no workgroup barrier: a wave does not wait for the slowest pack of its
// workgroup before taking its next one.
// kSlots false: the triangle total only, accumulated into the partial sums at total.
template <int U, bool kSlots>
__global__ __launch_bounds__(kWG) void k_fwd_count(const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ hs,
                                                   const int32_t *__restrict__ adj,
                                                   const uint32_t *__restrict__ rev,
                                                   const FwdUnit *__restrict__ block_units, uint32_t n_block,
                                                   const FwdUnit *__restrict__ wave_units, uint32_t n_wave,
                                                   int32_t *cnt, unsigned long long *total) {
  __shared__ __align__(16) unsigned char smem[kCountLdsBytes];
  const uint32_t tasks = n_block + (n_wave + kWaves - 1) / kWaves;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint32_t t = blockIdx.x;
  for (; t < n_block; t += gridDim.x) {
    fwd_block_unit<U, kSlots>(off, adj, rev, block_units[t], cnt, total, smem);
    __syncthreads();
  }
  for (; t < tasks; t += gridDim.x) {
    const uint32_t ui = (t - n_block) * kWaves + wave;
    if (ui < n_wave) fwd_pack_unit<U, kSlots>(off, hs, adj, rev, wave_units[ui], cnt, total, smem, wave);
  }
}

// Sources whose N<(b) exceeds the workgroup table: membership by binary
// search of N<(b) in global memory, (b,a) counted by direct atomics.  One
// wave per upper slot; the probe stops at the reverse slot.
template <bool kSlots>
__global__ __launch_bounds__(accfwd::kBigWG) void k_fwd_big(const uint32_t *__restrict__ off,
                                                            const int32_t *__restrict__ adj,
                                                            const uint32_t *__restrict__ rev,
                                                            const FwdUnit *__restrict__ units, int32_t *cnt,
                                                            unsigned long long *total) {
  const FwdUnit u = units[blockIdx.x];
  const int32_t b = u.b;
  const uint32_t sb = off[b];
  const uint32_t nl = u.nlow;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  for (uint32_t j = u.j0 + wave; j < u.j1; j += accfwd::kBigWG / 64) {
    const int32_t c = adj[j];
    const uint32_t cb = off[c];
    const uint32_t p = rev[j];
    const uint32_t ce = p != ~0u ? p : cb;
    int reg = 0;
    for (uint32_t base = cb; base < ce; base += 64) {
      const uint32_t idx = base + lane;
      if (idx < ce) {
        const int32_t a = adj[idx];
        uint32_t lo = 0, hi = nl;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (adj[sb + mid] < a)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (lo < nl && adj[sb + lo] == a) {
          reg++;
          if (kSlots) {
            atomicAdd(&cnt[idx], 1);
            atomicAdd(&cnt[sb + lo], 1);
          }
        }
      }
    }
    reg = wave_sum_i32(reg);
    if (lane == 0) {
      if (kSlots)
        cnt[j] = reg;
      else if (reg)
        atomicAdd(tri_part(total), static_cast<unsigned long long>(reg));
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

__global__ __launch_bounds__(kTriParts) void k_tri_zero(unsigned long long *__restrict__ parts) {
  parts[threadIdx.x * kTriStride] = 0;
}

__global__ __launch_bounds__(kTriParts) void k_tri_sum(const unsigned long long *__restrict__ parts,
                                                       unsigned long long *__restrict__ total) {
  __shared__ unsigned long long w[kTriParts / 64];
  const unsigned long long v = wave_sum_u64(parts[threadIdx.x * kTriStride]);
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kTriParts / 64; i++) t += w[i];
    *total = t;
  }
}

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long *__restrict__ p, uint32_t count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < count) p[i] = 0;
}

// Per-graph sums of raw counts (cnt - plus) over a multi-graph CSR whose
// graph k owns slots [gsoff[k], gsoff[k+1]).  Workgroup w sums the slots
// [w * kTotSpan, (w + 1) * kTotSpan), split at graph boundaries, and adds
// each piece to its graph's total.
constexpr uint32_t kTotSpan = 1u << 16;
__global__ __launch_bounds__(256) void k_graph_totals(const uint32_t *__restrict__ gsoff, uint32_t parts,
                                                      uint32_t m, const int32_t *__restrict__ cnt, int plus,
                                                      unsigned long long *__restrict__ gtot) {
  __shared__ unsigned long long part[4];
  uint32_t s0 = blockIdx.x * kTotSpan;
  const uint32_t s1 = m - s0 < kTotSpan ? m : s0 + kTotSpan;
  // k: the last graph with gsoff[k] <= s0
  uint32_t lo = 0, hi = parts;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (gsoff[mid] <= s0) lo = mid; else hi = mid;
  }
  uint32_t k = __builtin_amdgcn_readfirstlane(lo);
  while (s0 < s1) {
    const uint32_t ge = __builtin_amdgcn_readfirstlane(gsoff[k + 1]);
    const uint32_t e = ge < s1 ? ge : s1;
    unsigned long long s = 0;
    for (uint32_t i = s0 + threadIdx.x; i < e; i += 256) s += static_cast<unsigned long long>(cnt[i] - plus);
    s = wave_sum_u64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned long long t = part[0] + part[1] + part[2] + part[3];
      if (t) atomicAdd(&gtot[k], t);
    }
    __syncthreads();
    s0 = e;
    k++;
  }
}

// Slots per workgroup of k_combine_parts: small enough that the grid is many
// times the resident workgroups (a short tail), large enough that the
// boundary search and the per-piece atomics stay negligible.
constexpr uint32_t kCombSpan = 1u << 12;

// Span of workgroup b in a grid of N: workgroups reach the 8 XCDs
// round-robin, so b is remapped to give each XCD runs of kXcdRun
// consecutive spans (4 MiB of counts, the size of an XCD's L2), the runs
// interleaved over the whole array: a small graph's reverse slots then lie
// in one XCD's L2.  A bijection on [0, N); the tail short of a round of 8
// runs keeps its order.  On the 512-graph mix, runs of 64 / 256 / 1024
// spans made the whole pass 0.45 / 0.7 / 0.7 ms shorter than the plain
// order (17.85 ms).
constexpr uint32_t kXcds = 8, kXcdRun = 256;
__device__ __forceinline__ uint32_t xcd_span(uint32_t b, uint32_t N) {
  const uint32_t full = N / (kXcds * kXcdRun) * (kXcds * kXcdRun);
  if (b >= full) return b;
  const uint32_t x = b % kXcds, local = b / kXcds;
  return (local / kXcdRun) * (kXcds * kXcdRun) + x * kXcdRun + local % kXcdRun;
}

// k_combine for a partitioned graph, with its per-graph totals: workgroup w
// takes the slots [w * kCombSpan, (w + 1) * kCombSpan), split at graph
// boundaries (a reverse slot lies in its upper slot's graph), combines every
// upper slot of a piece and adds twice the piece's raw edge totals (both
// slots of every edge: 6 T_k in all) to its graph's total.
__global__ __launch_bounds__(256) void k_combine_parts(int32_t *__restrict__ cnt, const uint32_t *__restrict__ rev,
                                                       uint32_t m, int plus, unsigned long long *__restrict__ sum,
                                                       const uint32_t *__restrict__ gsoff, uint32_t parts,
                                                       unsigned long long *__restrict__ gtot) {
  __shared__ unsigned long long part[4];
  uint32_t s0 = xcd_span(blockIdx.x, gridDim.x) * kCombSpan;
  const uint32_t s1 = m - s0 < kCombSpan ? m : s0 + kCombSpan;
  // k: the last graph with gsoff[k] <= s0
  uint32_t lo = 0, hi = parts;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (gsoff[mid] <= s0) lo = mid; else hi = mid;
  }
  uint32_t k = __builtin_amdgcn_readfirstlane(lo);
  unsigned long long all = 0;
  while (s0 < s1) {
    const uint32_t ge = __builtin_amdgcn_readfirstlane(gsoff[k + 1]);
    const uint32_t e = ge < s1 ? ge : s1;
    unsigned long long s = 0;
    for (uint32_t i = s0 + threadIdx.x; i < e; i += 256) {
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
      if (t) atomicAdd(&gtot[k], 2 * t);
      all += t;
    }
    __syncthreads();
    s0 = e;
    k++;
  }
  if (threadIdx.x == 0 && all) atomicAdd(sum, all);
}

// Flags: 1 id out of range, 2 self-loop, 4 unsorted (duplicates included).
// Symmetry is checked by the reverse-slot build.  The source of slot i is src[i]
// (off[src[i]] <= i < off[src[i] + 1]).
__global__ __launch_bounds__(256) void k_validate(const uint32_t *__restrict__ off,
                                                  const int32_t *__restrict__ adj,
                                                  const uint32_t *__restrict__ src, uint32_t n, uint64_t m,
                                                  int32_t *__restrict__ flags) {
  const uint64_t tid = blockIdx.x * 256ull + threadIdx.x;
  const uint64_t stride = gridDim.x * 256ull;
  int32_t f = 0;
  for (uint64_t i = tid; i < m; i += stride) {
    const uint32_t lo = src[i];
    const int32_t v = adj[i];
    if (v < 0 || static_cast<uint32_t>(v) >= n) {
      f |= 1;
      continue;
    }
    if (static_cast<uint32_t>(v) == lo) f |= 2;
    if (i > off[lo] && adj[i - 1] >= v) f |= 4;
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
                     g->d_cnt, g->m, g->d_sum, g->num_parts ? g->d_gtot : nullptr, g->num_parts);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_build_rev(acc_graph *g, bool validate) {
  if (!g->m) return ACC_OK;
  hipLaunchKernelGGL(k_build_rev, dim3(grid_for(g->m, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream, g->d_off,
                     g->d_adj, reinterpret_cast<const uint32_t *>(g->d_cnt), g->n, g->m, g->d_rev,
                     validate ? g->d_flags : nullptr,
                     reinterpret_cast<unsigned long long *>(g->d_flags + 2));
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_source_work(acc_graph *g, unsigned long long *work) {
  if (!g->n) return ACC_OK;
  const uint64_t want = (static_cast<uint64_t>(g->n) + 3) / 4;
  const uint64_t cap = static_cast<uint64_t>(g->ctx->num_cus > 0 ? g->ctx->num_cus : 256) * 16;
  hipLaunchKernelGGL(k_source_work, dim3(static_cast<unsigned>(want < cap ? want : cap)), dim3(256), 0, g->ctx->stream,
                     g->d_off, g->d_hs, g->d_adj, g->d_rev, g->d_cnt, g->n, g->chunk, work);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_vertex_meta(acc_graph *g, int32_t *lo, uint32_t *maxes) {
  ACC_HIP_TRY(hipMemsetAsync(maxes, 0, 2 * sizeof(uint32_t), g->ctx->stream));
  if (!g->n) return ACC_OK;
  hipLaunchKernelGGL(k_vertex_meta, dim3(grid_for(g->n, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream, g->d_off,
                     g->d_adj, g->n, g->d_hs, lo, maxes);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

template <bool kSlots>
int launch_count_t(acc_graph *g, unsigned long long *total) {
  hipStream_t st = g->ctx->stream;
  const uint64_t tasks = g->n_fwd_block + (g->n_fwd_wave + kWaves - 1) / kWaves;
  const uint64_t cap = static_cast<uint64_t>(g->persist) * static_cast<uint64_t>(g->ctx->num_cus);
  const unsigned grid = static_cast<unsigned>(g->persist && cap < tasks ? cap : tasks);
  if (grid) {
    switch (g->fwd_unroll) {
      case 1:
        hipLaunchKernelGGL((k_fwd_count<1, kSlots>), dim3(grid), dim3(kWG), 0, st, g->d_off, g->d_hs, g->d_adj, g->d_rev,
                           g->d_fwd_block, g->n_fwd_block, g->d_fwd_wave, g->n_fwd_wave, g->d_cnt, total);
        break;
      case 4:
        hipLaunchKernelGGL((k_fwd_count<4, kSlots>), dim3(grid), dim3(kWG), 0, st, g->d_off, g->d_hs, g->d_adj, g->d_rev,
                           g->d_fwd_block, g->n_fwd_block, g->d_fwd_wave, g->n_fwd_wave, g->d_cnt, total);
        break;
      default:
        hipLaunchKernelGGL((k_fwd_count<2, kSlots>), dim3(grid), dim3(kWG), 0, st, g->d_off, g->d_hs, g->d_adj, g->d_rev,
                           g->d_fwd_block, g->n_fwd_block, g->d_fwd_wave, g->n_fwd_wave, g->d_cnt, total);
        break;
    }
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

template <bool kSlots>
int launch_big_t(acc_graph *g, unsigned long long *total) {
  if (g->n_fwd_big) {
    hipLaunchKernelGGL(k_fwd_big<kSlots>, dim3(g->n_fwd_big), dim3(kBigWG), 0, g->ctx->stream, g->d_off, g->d_adj,
                       g->d_rev, g->d_fwd_big, g->d_cnt, total);
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

int launch_count(acc_graph *g) { return launch_count_t<true>(g, nullptr); }

int launch_count_big(acc_graph *g) { return launch_big_t<true>(g, nullptr); }

int launch_triangles(acc_graph *g) {
  if (!g->d_tri_parts) {
    if (hipMalloc(reinterpret_cast<void **>(&g->d_tri_parts), sizeof(unsigned long long) * kTriParts * kTriStride) !=
        hipSuccess) {
      g->d_tri_parts = nullptr;
      return acc_set_error(ACC_ENOMEM, "hipMalloc of the triangle partial sums failed");
    }
    g->device_bytes += sizeof(unsigned long long) * kTriParts * kTriStride;
  }
  hipStream_t st = g->ctx->stream;
  hipLaunchKernelGGL(k_tri_zero, dim3(1), dim3(kTriParts), 0, st, g->d_tri_parts);
  ACC_HIP_TRY(hipGetLastError());
  ACC_TRY(launch_count_t<false>(g, g->d_tri_parts));
  ACC_TRY(launch_big_t<false>(g, g->d_tri_parts));
  hipLaunchKernelGGL(k_tri_sum, dim3(1), dim3(kTriParts), 0, st, g->d_tri_parts, g->d_sum + 1);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_combine(acc_graph *g, int plus) {
  if (g->num_parts) {
    if (g->m) {
      const uint32_t blocks = static_cast<uint32_t>((g->m + kCombSpan - 1) / kCombSpan);
      hipLaunchKernelGGL(k_combine_parts, dim3(blocks), dim3(256), 0, g->ctx->stream, g->d_cnt, g->d_rev,
                         static_cast<uint32_t>(g->m), plus, g->d_sum, g->d_gsoff, g->num_parts, g->d_gtot);
    }
    ACC_HIP_TRY(hipGetLastError());
    return ACC_OK;
  }
  hipLaunchKernelGGL(k_combine, dim3(grid_for(g->m, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream, g->d_cnt,
                     g->d_rev, g->m, plus, g->d_sum);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

int launch_graph_totals(acc_graph *g, int plus) {
  if (!g->num_parts) return ACC_OK;
  hipLaunchKernelGGL(k_zero_u64, dim3((g->num_parts + 255) / 256), dim3(256), 0, g->ctx->stream, g->d_gtot,
                     g->num_parts);
  ACC_HIP_TRY(hipGetLastError());
  if (g->m) {
    const uint32_t blocks = static_cast<uint32_t>((g->m + kTotSpan - 1) / kTotSpan);
    hipLaunchKernelGGL(k_graph_totals, dim3(blocks), dim3(256), 0, g->ctx->stream, g->d_gsoff, g->num_parts,
                       static_cast<uint32_t>(g->m), g->d_cnt, plus, g->d_gtot);
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

int launch_validate(acc_graph *g) {
  if (!g->m) return ACC_OK;
  hipLaunchKernelGGL(k_validate, dim3(grid_for(g->m, g->ctx->num_cus)), dim3(256), 0, g->ctx->stream, g->d_off,
                     g->d_adj, reinterpret_cast<const uint32_t *>(g->d_cnt), g->n, g->m, g->d_flags);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

}  // namespace accfwd
