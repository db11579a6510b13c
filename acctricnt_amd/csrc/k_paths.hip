// The two per-edge counting paths of the reference, MI355X-native:
//
// ACC_PATH_BITMAP -- the reference's BMP algorithm
//   (set_intersection_GPU_bitmap_warp_per_vertex,
//    scan-query-cuda/scan_xp_multi_gpu_multi_pass_common.cu:308-407):
//   for each source u a membership set of N(u) is built once, then every
//   neighbor list N(v), v > u, is probed against it.  Small sets (up to
//   ACC_PARAM_BMP_HASH_MAX ids, 256 by default) live in an LDS hash table,
//   behind an LDS index of N(u) (one bit per 32 ids) in a workgroup unit;
//   larger sets spill to an n-bit bitmap in HBM,
//   built for a batch of hubs by one kernel and shared by every workgroup
//   that probes that hub, with an LDS index of one bit per 16 ids (per
//   larger group beyond 2^18 groups) checked before the HBM word, the
//   reference's BITMAP_SCALE_LOG index level (same file :320-382).  LDS hash
//   tables of 2^12, 2^14 and 2^15 keys (up to 16384 ids; gfx950 gives a
//   workgroup 160 KB of LDS) remain selectable for larger sets, though they
//   probe slower than the indexed HBM bitmaps (at R-MAT scale 22: 103 ms for
//   the whole path with tables up to 2048 ids, 80 ms with 256).
// ACC_PATH_MERGE -- the reference's MPS kernels
//   (set_intersection_GPU_shared_merge_filtered / _galloping_filtered, same
//    file :32-176): N(u) is staged in LDS in windows of sorted ids; every
//   element of a probed list N(v) finds its rank in the window by a binary
//   search there (the merge-path co-rank; a skewed pair costs log d_u per
//   element of the shorter list, a balanced pair log d_u too, with no
//   global-memory search), and a window's slice of N(v) is found by one
//   binary search per window boundary.
//
// Both paths share the probe layout of the forward path: a workgroup unit is
// one source u and at most kChunk of its upper slots (u, v); the lists N(v)
// (or their slices) are laid end to end, element e of the unit's range
// belongs to probe k(e) = max{k : P_k <= e} over the exclusive prefix P of
// the lengths (a window search, flat_window.h), each lane takes one element per
// step, so no lane idles on a short list, and per-probe hit counts go to
// LDS counters (one popcount when a 64-element round lies in one probe).
// Each unit then writes 2 + |N(u) ∩ N(v)| (or the raw count) into the upper
// slot (u, v) and its reverse slot (from the reverse-slot index built at
// upload), with no atomics on the counts, and adds the raw counts to the
// triangle sum.

#include "acc_internal.h"
#include "flat_window.h"

namespace {

constexpr int kWG = accfwd::kWG;
constexpr int kWaves = accfwd::kWaves;
constexpr uint32_t kChunk = accfwd::kChunk;
constexpr int kU = 4;  // 64-element rounds per wave step: loads in flight per lane
constexpr int kMergeU = 8;  // the same for the merge path's large windows (13-step LDS searches)

static_assert(kWG == kChunk, "one probe per thread of the unit");

__device__ __forceinline__ uint32_t hslot(int32_t a, int bits) {
  return (static_cast<uint32_t>(a) * 0x9E3779B1u) >> (32 - bits);
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, int lane) {
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

typedef __attribute__((address_space(3))) int32_t lds_i32;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ void lds_or(uint32_t *p, uint32_t v) {
  __hip_atomic_fetch_or((lds_u32 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Probe metadata of a unit in LDS.
struct UnitLds {
  uint32_t P[kChunk];    // exclusive prefix of the probe lengths; T beyond the probes
  uint32_t cb[kChunk];   // first slot of each probe less its prefix
  int32_t ctr[kChunk];   // hits per probe
  uint32_t wsum[kWaves];
};

// Lays the unit's probes end to end: thread t < nup passes its probe's first
// slot cb and length len (others len 0).  Returns the total T (uniform).
// Callers synchronise before reading P and cb.
__device__ __forceinline__ uint32_t flatten(UnitLds &L, uint32_t cb, uint32_t len, uint32_t nup) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t excl = wave_excl_scan(len, lane);
  if (lane == 63) L.wsum[wave] = excl + len;
  __syncthreads();
  uint32_t before = 0, T = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    const uint32_t ws = L.wsum[w];
    before += w < wave ? ws : 0;
    T += ws;
  }
  L.P[tid] = static_cast<uint32_t>(tid) < nup ? before + excl : T;
  if (static_cast<uint32_t>(tid) < nup) L.cb[tid] = cb - (before + excl);
  return __builtin_amdgcn_readfirstlane(T);
}

// Probes the T flattened elements against set.  Wave w takes the contiguous
// range [w R, (w + 1) R) of the elements (R a multiple of 64, the ranges
// equal but for the last), in steps of U * 64; each step's elements find
// their probes by the window search of flat_window.h.  Step indices are
// wave-uniform.
template <int U = kU, class Set>
__device__ __forceinline__ void probe_flat(const int32_t *__restrict__ adj, const Set &set, UnitLds &L, uint32_t T) {
  static_assert(kChunk == 64 * 8, "the window search covers 64 groups of 8 entries");
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t R = ((T + kWaves - 1) / kWaves + 63) & ~63u;
  const uint32_t lo = min(T, static_cast<uint32_t>(wave) * R), hi = min(T, lo + R);
  for (uint32_t e0 = lo; e0 < hi; e0 += 64 * U) {
    int32_t a[U];
    int kq[U];
    accflat::lds_search<U>(L.P, e0, accflat::round_last<U>(e0, hi), lane, kq);
#pragma unroll
    for (int q = 0; q < U; q++) {
      const uint32_t e = e0 + 64 * q + lane;
      a[q] = e < hi ? adj[L.cb[kq[q]] + e] : -1;
    }
    bool hit[U];
    set.template has_n<U>(a, hit);
#pragma unroll
    for (int q = 0; q < U; q++) accflat::add_row_hits(L.ctr, kq[q], hit[q], lane);
  }
}

// Writes each probe's count into its upper slot and reverse slot and adds
// the unit's raw counts to *sum.  Follows a __syncthreads.
__device__ __forceinline__ void finish_unit(const UnitLds &L, const uint32_t *__restrict__ rev, uint32_t j0,
                                            uint32_t nup, int32_t *cnt, int plus, unsigned long long *sum) {
  const int tid = threadIdx.x;
  int c = 0;
  if (static_cast<uint32_t>(tid) < nup) {
    c = L.ctr[tid];
    cnt[j0 + tid] = c + plus;
    const uint32_t p = rev[j0 + tid];
    if (p != ~0u) cnt[p] = c + plus;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((tid & 63) == 0 && c) atomicAdd(sum, static_cast<unsigned long long>(c));
}

// The probe of upper slot j: N(v) for v = adj[j] (empty for an id out of range).
__device__ __forceinline__ void probe_of(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj, uint32_t n,
                                         uint32_t j, uint32_t &cb, uint32_t &ce) {
  const int32_t v = adj[j];
  if (v >= 0 && static_cast<uint32_t>(v) < n) {
    cb = off[v];
    ce = off[v + 1];
  } else {
    cb = ce = 0;
  }
}

// ---------------------------------------------------------------- sets
// Open-addressing hash set of N(u) in LDS (keys only; -1 = empty).
struct LdsHash {
  const int32_t *keys;
  int bits;
  uint32_t mask;
  __device__ __forceinline__ bool has(int32_t a) const {
    uint32_t h = hslot(a, bits);
    for (;;) {
      const int32_t k = keys[h];
      if (k == a) return true;
      if (k == -1) return false;
      h = (h + 1) & mask;
    }
  }
  // U lookups (a < 0: none): the first probe of each together, then the
  // rare collision chains one by one.
  template <int U>
  __device__ __forceinline__ void has_n(const int32_t (&a)[U], bool (&hit)[U]) const {
    uint32_t h[U];
    int32_t k[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      h[q] = hslot(a[q] < 0 ? 0 : a[q], bits);
      k[q] = keys[h[q]];
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      hit[q] = a[q] >= 0 && k[q] == a[q];
      if (a[q] >= 0 && k[q] != a[q] && k[q] != -1) hit[q] = has(a[q]);
    }
  }
};

// The same table behind an LDS index with one bit per 2^shift words of 32
// ids (set when N(u) holds one of their ids): most probes of a small set miss
// the index and never walk the table.
struct IndexedHash {
  const uint32_t *idx;
  int shift;
  LdsHash h;
  template <int U>
  __device__ __forceinline__ void has_n(const int32_t (&a)[U], bool (&hit)[U]) const {
    int32_t b[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const uint32_t iw = (static_cast<uint32_t>(a[q] < 0 ? 0 : a[q]) >> 5) >> shift;
      b[q] = a[q] >= 0 && ((idx[iw >> 5] >> (iw & 31)) & 1u) ? a[q] : -1;
    }
    h.has_n<U>(b, hit);
  }
};

// N(u) of a hub as a bitmap in HBM, behind an LDS index with one bit per
// group of 16 << shift ids (set when N(u) holds one of them): finer than a
// bitmap word, so that fewer probe elements go on to read an HBM word (for a
// hub holding 2.5% of the ids, 33% of the elements pass a 16-id bit against
// 55% a 32-id one).
constexpr int kHubGroupLog = 4;
struct HubBits {
  const uint32_t *idx;
  const uint32_t *__restrict__ bm;
  int shift;
  template <int U>
  __device__ __forceinline__ void has_n(const int32_t (&a)[U], bool (&hit)[U]) const {
    bool in[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const uint32_t iw = (static_cast<uint32_t>(a[q] < 0 ? 0 : a[q]) >> kHubGroupLog) >> shift;
      in[q] = a[q] >= 0 && ((idx[iw >> 5] >> (iw & 31)) & 1u);
    }
    uint32_t w[U];
#pragma unroll
    for (int q = 0; q < U; q++) w[q] = in[q] ? bm[static_cast<uint32_t>(a[q]) >> 5] : 0u;
#pragma unroll
    for (int q = 0; q < U; q++) hit[q] = in[q] && ((w[q] >> (a[q] & 31)) & 1u);
  }
};

// A window of 2^LOG sorted ids of N(u) in LDS, padded past the list's end
// with INT_MAX (no valid id, above every id): membership by a branch-free
// binary search of LOG steps with no bounds to check.
template <int LOG>
struct PaddedWin {
  const int32_t *s;
  // U searches advanced together, one LDS read of each per step.
  template <int U>
  __device__ __forceinline__ void has_n(const int32_t (&a)[U], bool (&hit)[U]) const {
    uint32_t pos[U];
#pragma unroll
    for (int q = 0; q < U; q++) pos[q] = 0;
#pragma unroll
    for (uint32_t step = 1u << (LOG - 1); step; step >>= 1) {
      int32_t v[U];
#pragma unroll
      for (int q = 0; q < U; q++) v[q] = s[pos[q] + step - 1];
#pragma unroll
      for (int q = 0; q < U; q++)
        if (v[q] < a[q]) pos[q] += step;
    }
    // pos: the number of window ids below a, at most 2^LOG - 1
#pragma unroll
    for (int q = 0; q < U; q++) hit[q] = a[q] >= 0 && s[pos[q]] == a[q];
  }
};

// ---------------------------------------------------------------- wave units
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave units: one source u (d_u <= kEdgeWaveDeg, at most 64 upper slots) per
// wave, eight per workgroup, with no workgroup barrier: lane t holds probe t
// (upper slot j0 + t), the wave lays the probes end to end and finds each
// element's probe by a window search over the lanes' prefix.  N(u) is a
// wave-private LDS hash table (bitmap path) or sorted run (merge path).
template <bool MERGE>
__global__ __launch_bounds__(kWG) void k_edge_wave(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                   const uint32_t *__restrict__ rev, uint32_t n,
                                                   const FwdUnit *__restrict__ units, uint32_t n_units, int32_t *cnt,
                                                   int plus, unsigned long long *sum) {
  constexpr int kT = 2 * kEdgeWaveDeg;
  __shared__ int32_t tab[kWaves][kT];
  __shared__ int32_t ctrs[kWaves][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t ui = blockIdx.x * kWaves + static_cast<uint32_t>(wave);
  int c = 0;
  if (ui < n_units) {
    const FwdUnit un = units[ui];
    const int32_t u = un.b;
    const uint32_t su = off[u], du = off[u + 1] - su;
    const uint32_t nup = un.j1 - un.j0;
    int32_t *t = tab[wave];
    int32_t *ct = ctrs[wave];
    uint32_t cb = 0, ce = 0;
    if (static_cast<uint32_t>(lane) < nup) probe_of(off, adj, n, un.j0 + lane, cb, ce);
    const uint32_t len = ce - cb;
    const uint32_t P = wave_excl_scan(len, lane);
    const uint32_t T = __builtin_amdgcn_readfirstlane(__shfl(P + len, 63));
    ct[lane] = 0;
    int bits = 6;
    while ((1u << bits) < 2 * du) bits++;
    const uint32_t mask = (1u << bits) - 1;
    if (MERGE) {
      // the sorted run padded to kEdgeWaveDeg ids (du <= kEdgeWaveDeg)
      for (uint32_t i = lane; i < kEdgeWaveDeg; i += 64) t[i] = i < du ? adj[su + i] : INT32_MAX;
    } else {
      for (uint32_t i = lane; i <= mask; i += 64) t[i] = -1;
      wave_lds_sync();
      for (uint32_t i = lane; i < du; i += 64) {
        const int32_t a = adj[su + i];
        uint32_t h = hslot(a, bits);
        while (atomicCAS(&t[h], -1, a) != -1) h = (h + 1) & mask;
      }
    }
    wave_lds_sync();
    for (uint32_t e0 = 0; e0 < T; e0 += 64 * kU) {
      int32_t a[kU];
      int kq[kU];
      accflat::lane_window<kU>(P, e0, accflat::round_last<kU>(e0, T), lane, kq);
#pragma unroll
      for (int q = 0; q < kU; q++) {
        const uint32_t e = e0 + 64 * q + lane;
        const uint32_t pk = __shfl(P, kq[q]), ck = __shfl(cb, kq[q]);
        a[q] = e < T ? adj[ck + (e - pk)] : -1;
      }
      bool hit[kU];
      if (MERGE)
        PaddedWin<kEdgeWaveLog>{t}.has_n<kU>(a, hit);
      else
        LdsHash{t, bits, mask}.has_n<kU>(a, hit);
#pragma unroll
      for (int q = 0; q < kU; q++) accflat::add_row_hits(ct, kq[q], hit[q], lane);
    }
    wave_lds_sync();
    if (static_cast<uint32_t>(lane) < nup) {
      c = ct[lane];
      cnt[un.j0 + lane] = c + plus;
      const uint32_t p = rev[un.j0 + lane];
      if (p != ~0u) cnt[p] = c + plus;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if (lane == 0 && c) atomicAdd(sum, static_cast<unsigned long long>(c));
}

// ---------------------------------------------------------------- BITMAP path
// LDS-hash units: one workgroup per chunk of at most kChunk upper slots of a
// source u with d_u <= 2^(TBITS-1); the table has the smallest power of two
// >= 2 d_u entries (at least 64).
// With INDEX, an LDS index of N(u) (one bit per 2^shift words of 32 ids,
// kIdxWords words) filters the probes first.
constexpr uint32_t kIdxWords = 4096;  // LDS index of 2^17 bits
template <int TBITS, bool INDEX>
__global__ __launch_bounds__(kWG) void k_bmp_hash(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                  const uint32_t *__restrict__ rev, uint32_t n,
                                                  const FwdUnit *__restrict__ units, uint64_t words, int shift,
                                                  int32_t *cnt, int plus, unsigned long long *sum) {
  __shared__ int32_t keys[1 << TBITS];
  __shared__ uint32_t idx[INDEX ? kIdxWords : 1];
  __shared__ UnitLds L;
  const FwdUnit un = units[blockIdx.x];
  const int32_t u = un.b;
  const uint32_t su = off[u], du = off[u + 1] - su;
  const uint32_t nup = un.j1 - un.j0;
  const int tid = threadIdx.x;
  int bits = 6;
  while ((1u << bits) < 2 * du) bits++;
  const uint32_t mask = (1u << bits) - 1;
  for (uint32_t i = tid; i <= mask; i += kWG) keys[i] = -1;
  const uint32_t iwords = static_cast<uint32_t>((((words + (1ull << shift) - 1) >> shift) + 31) / 32);
  if (INDEX)
    for (uint32_t i = tid; i < iwords; i += kWG) idx[i] = 0;
  uint32_t cb = 0, ce = 0;
  if (static_cast<uint32_t>(tid) < nup) probe_of(off, adj, n, un.j0 + tid, cb, ce);
  L.ctr[tid] = 0;
  const uint32_t T = flatten(L, cb, ce - cb, nup);
  for (uint32_t i = tid; i < du; i += kWG) {
    const int32_t a = adj[su + i];
    uint32_t h = hslot(a, bits);
    while (atomicCAS(&keys[h], -1, a) != -1) h = (h + 1) & mask;
    if (INDEX) {
      const uint32_t iw = (static_cast<uint32_t>(a) >> 5) >> shift;
      lds_or(&idx[iw >> 5], 1u << (iw & 31));
    }
  }
  __syncthreads();
  if (INDEX)
    probe_flat(adj, IndexedHash{idx, shift, LdsHash{keys, bits, mask}}, L, T);
  else
    probe_flat(adj, LdsHash{keys, bits, mask}, L, T);
  __syncthreads();
  finish_unit(L, rev, un.j0, nup, cnt, plus, sum);
}

// Hub bitmaps of one batch: workgroup h sets the bits of N(hubs[h]) in
// bitmap h of the pool (zero before), or clears them again (set == false).
__global__ __launch_bounds__(256) void k_hub_bits(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                  const uint32_t *__restrict__ hubs, uint32_t *pool, uint64_t words,
                                                  bool set) {
  const uint32_t u = hubs[blockIdx.x];
  uint32_t *bm = pool + static_cast<uint64_t>(blockIdx.x) * words;
  for (uint32_t i = off[u] + threadIdx.x; i < off[u + 1]; i += 256) {
    const int32_t a = adj[i];
    if (set)
      atomicOr(&bm[static_cast<uint32_t>(a) >> 5], 1u << (a & 31));
    else
      bm[static_cast<uint32_t>(a) >> 5] = 0;
  }
}

// Hub units of one batch: the unit's source is hub slot un.nlow of the pool.
constexpr uint32_t kHubIdxWords = 8192;  // LDS index of 2^18 bits
__global__ __launch_bounds__(kWG) void k_bmp_hub(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                 const uint32_t *__restrict__ rev, uint32_t n,
                                                 const FwdUnit *__restrict__ units, const uint32_t *__restrict__ pool,
                                                 uint64_t words, int shift, int32_t *cnt, int plus,
                                                 unsigned long long *sum) {
  __shared__ uint32_t idx[kHubIdxWords];
  __shared__ UnitLds L;
  const FwdUnit un = units[blockIdx.x];
  const int32_t u = un.b;
  const uint32_t su = off[u], eu = off[u + 1];
  const uint32_t nup = un.j1 - un.j0;
  const int tid = threadIdx.x;
  const uint64_t groups = (static_cast<uint64_t>(n) + 15) >> kHubGroupLog;
  const uint32_t iwords = static_cast<uint32_t>((((groups + (1ull << shift) - 1) >> shift) + 31) / 32);
  for (uint32_t i = tid; i < iwords; i += kWG) idx[i] = 0;
  uint32_t cb = 0, ce = 0;
  if (static_cast<uint32_t>(tid) < nup) probe_of(off, adj, n, un.j0 + tid, cb, ce);
  L.ctr[tid] = 0;
  const uint32_t T = flatten(L, cb, ce - cb, nup);
  for (uint32_t i = su + tid; i < eu; i += kWG) {
    const uint32_t iw = (static_cast<uint32_t>(adj[i]) >> kHubGroupLog) >> shift;
    lds_or(&idx[iw >> 5], 1u << (iw & 31));
  }
  __syncthreads();
  probe_flat(adj, HubBits{idx, pool + static_cast<uint64_t>(un.nlow) * words, shift}, L, T);
  __syncthreads();
  finish_unit(L, rev, un.j0, nup, cnt, plus, sum);
}

// ---------------------------------------------------------------- MERGE path
// Windowed units: N(u) in windows of 2^WBITS sorted ids staged in LDS; per
// window w, probe k's slice of N(v) holds the ids in [N(u)[w W], N(u)[(w+1) W])
// (unbounded below for the first window and above for the last), found by a
// binary search of N(v) for the upper boundary, the lower being the previous
// window's.  Every element of N(v) lies in exactly one slice.
template <int WBITS, int U>
__global__ __launch_bounds__(kWG) void k_merge_win(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                   const uint32_t *__restrict__ rev, uint32_t n,
                                                   const FwdUnit *__restrict__ units, int32_t *cnt, int plus,
                                                   unsigned long long *sum) {
  constexpr uint32_t W = 1u << WBITS;
  __shared__ int32_t win[W];
  __shared__ UnitLds L;
  const FwdUnit un = units[blockIdx.x];
  const int32_t u = un.b;
  const uint32_t su = off[u], du = off[u + 1] - su;
  const uint32_t nup = un.j1 - un.j0;
  const int tid = threadIdx.x;
  uint32_t cur = 0, ce = 0;  // this thread's probe: the unprobed rest [cur, ce) of N(v)
  if (static_cast<uint32_t>(tid) < nup) probe_of(off, adj, n, un.j0 + tid, cur, ce);
  L.ctr[tid] = 0;
  const uint32_t nwin = (du + W - 1) / W;
  for (uint32_t w = 0; w < nwin; w++) {
    const uint32_t w0 = w * W, wl = du - w0 < W ? du - w0 : W;
    for (uint32_t i = tid; i < W; i += kWG) win[i] = i < wl ? adj[su + w0 + i] : INT32_MAX;
    // slice end: the first element of N(v) at or above the next window's first id
    uint32_t end = ce;
    if (w + 1 < nwin && cur < ce) {
      const int32_t bound = adj[su + w0 + W];
      uint32_t lo = cur, hi = ce;
      while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (adj[mid] < bound)
          lo = mid + 1;
        else
          hi = mid;
      }
      end = lo;
    }
    const uint32_t T = flatten(L, cur, end - cur, nup);
    __syncthreads();
    probe_flat<U>(adj, PaddedWin<WBITS>{win}, L, T);
    cur = end;
    __syncthreads();
  }
  finish_unit(L, rev, un.j0, nup, cnt, plus, sum);
}

__global__ void k_zero_sum(unsigned long long *sum) {
  if (threadIdx.x == 0) *sum = 0;
}

template <class K, class... A>
int launch_units(K kernel, uint32_t units, hipStream_t st, A... args) {
  if (!units) return ACC_OK;
  hipLaunchKernelGGL(kernel, dim3(units), dim3(kWG), 0, st, args...);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

}  // namespace

namespace accpaths {

int launch_bitmap(acc_graph *g, int plus) {
  hipStream_t st = g->ctx->stream;
  const EdgePlan &p = g->edge;
  hipLaunchKernelGGL(k_zero_sum, dim3(1), dim3(64), 0, st, g->d_sum);
  ACC_HIP_TRY(hipGetLastError());
  ACC_TRY(launch_units(k_edge_wave<false>, (p.n_units[kEdgeWave] + kWaves - 1) / kWaves, st, g->d_off, g->d_adj,
                       g->d_rev, g->n, p.d_units[kEdgeWave], p.n_units[kEdgeWave], g->d_cnt, plus, g->d_sum));
  ACC_TRY(launch_units(k_bmp_hash<12, true>, p.n_units[kEdgeHash12], st, g->d_off, g->d_adj, g->d_rev, g->n,
                       p.d_units[kEdgeHash12], p.pool_words, p.idx_shift, g->d_cnt, plus, g->d_sum));
  ACC_TRY(launch_units(k_bmp_hash<14, false>, p.n_units[kEdgeHash14], st, g->d_off, g->d_adj, g->d_rev, g->n,
                       p.d_units[kEdgeHash14], p.pool_words, p.idx_shift, g->d_cnt, plus, g->d_sum));
  ACC_TRY(launch_units(k_bmp_hash<15, false>, p.n_units[kEdgeHash15], st, g->d_off, g->d_adj, g->d_rev, g->n,
                       p.d_units[kEdgeHash15], p.pool_words, p.idx_shift, g->d_cnt, plus, g->d_sum));
  // hub batches: set the batch's bitmaps, probe its units, clear the bitmaps
  for (size_t k = 0; k + 1 < p.hub_batch.size(); k++) {
    const uint32_t h0 = p.hub_batch[k], h1 = p.hub_batch[k + 1];
    const uint32_t u0 = p.hub_unit_batch[k], u1 = p.hub_unit_batch[k + 1];
    if (h1 == h0) continue;
    hipLaunchKernelGGL(k_hub_bits, dim3(h1 - h0), dim3(256), 0, st, g->d_off, g->d_adj, p.d_hubs + h0, p.d_pool,
                       p.pool_words, true);
    ACC_HIP_TRY(hipGetLastError());
    ACC_TRY(launch_units(k_bmp_hub, u1 - u0, st, g->d_off, g->d_adj, g->d_rev, g->n, p.d_units[kEdgeHub] + u0,
                         p.d_pool, p.pool_words, p.hub_idx_shift, g->d_cnt, plus, g->d_sum));
    hipLaunchKernelGGL(k_hub_bits, dim3(h1 - h0), dim3(256), 0, st, g->d_off, g->d_adj, p.d_hubs + h0, p.d_pool,
                       p.pool_words, false);
    ACC_HIP_TRY(hipGetLastError());
  }
  return ACC_OK;
}

int launch_merge(acc_graph *g, int plus) {
  hipStream_t st = g->ctx->stream;
  const EdgePlan &p = g->edge;
  hipLaunchKernelGGL(k_zero_sum, dim3(1), dim3(64), 0, st, g->d_sum);
  ACC_HIP_TRY(hipGetLastError());
  ACC_TRY(launch_units(k_edge_wave<true>, (p.n_units[kEdgeWave] + kWaves - 1) / kWaves, st, g->d_off, g->d_adj,
                       g->d_rev, g->n, p.d_units[kEdgeWave], p.n_units[kEdgeWave], g->d_cnt, plus, g->d_sum));
  // d_u <= 2048: one window of 2048 ids; the rest: windows of 8192 ids (38 KB
  // of LDS: four 512-thread workgroups per CU)
  ACC_TRY(launch_units(k_merge_win<11, kU>, p.n_units[kEdgeHash12], st, g->d_off, g->d_adj, g->d_rev, g->n,
                       p.d_units[kEdgeHash12], g->d_cnt, plus, g->d_sum));
  for (int c : {kEdgeHash14, kEdgeHash15, kEdgeHub})
    ACC_TRY(launch_units(k_merge_win<13, kMergeU>, p.n_units[c], st, g->d_off, g->d_adj, g->d_rev, g->n, p.d_units[c],
                         g->d_cnt, plus, g->d_sum));
  return ACC_OK;
}

}  // namespace accpaths
