// Probe windows of flattened probing (device code shared by the forward and
// per-edge counting kernels).
//
// A unit lays its probes end to end: P is the exclusive prefix of the probe
// lengths (nondecreasing, P_0 = 0, T beyond the last probe) and element e of
// the flat range belongs to probe k(e) = max{k : P_k <= e}.  A wave takes a
// round of U rows of 64 consecutive elements, e = e0 + 64 q + lane, whose
// probes all lie in [klo, khi] = [k(e0), k(e1)], e1 the round's last element
// below T.  Both window ends are found once per round (wave-uniform values),
// and each element searches only its window: a first step of w + 1 - hp,
// then hp / 2, ..., 1, for w = khi - klo and hp the largest power of two
// <= w, which reaches every offset in [0, w] and reads nothing outside
// [klo, khi].  A round inside one probe (w = 0) needs no search.  Elements at
// or past T get khi, so a round's probe indices never decrease across lanes.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace accflat {

// Window ends for a prefix of 512 entries in LDS: the 64 lanes test every
// 8th entry (two ballots: e0 and e1), then lanes 0-7 and 8-15 test the 8
// entries of the group holding e0 and e1.  Every lane of the wave calls it.
__device__ __forceinline__ void lds_window(const uint32_t *P, uint32_t e0, uint32_t e1, int lane, uint32_t &klo,
                                           uint32_t &khi) {
  const uint32_t coarse = P[8 * lane];
  const uint32_t glo = 8 * (static_cast<uint32_t>(__popcll(__ballot(coarse <= e0))) - 1);
  const uint32_t ghi = 8 * (static_cast<uint32_t>(__popcll(__ballot(coarse <= e1))) - 1);
  const bool lo_end = lane < 8;
  const uint32_t fine = P[(lo_end ? glo : ghi) + (lane & 7)];
  const unsigned long long fb = __ballot(fine <= (lo_end ? e0 : e1));
  klo = glo + static_cast<uint32_t>(__popcll(fb & 0xFEull));
  khi = ghi + static_cast<uint32_t>(__popcll(fb & 0xFE00ull));
}

// k(e) of the round's U element rows from the LDS prefix P (512 entries),
// the rows starting at e0 (e1 = the round's last element below T).
template <int U>
__device__ __forceinline__ void lds_search(const uint32_t *P, uint32_t e0, uint32_t e1, int lane, int (&k)[U]) {
  uint32_t klo, khi;
  lds_window(P, e0, e1, lane, klo, khi);
#pragma unroll
  for (int q = 0; q < U; q++) k[q] = static_cast<int>(klo);
  const uint32_t w = khi - klo;
  if (!w) return;
  const uint32_t hp = 1u << (31 - __builtin_clz(w));
  const uint32_t d = klo + w + 1 - hp;
  const uint32_t pd = P[d];
#pragma unroll
  for (int q = 0; q < U; q++)
    if (pd <= e0 + 64 * q + lane) k[q] = static_cast<int>(d);
  // the U searches advance together, one LDS read of each per step
  for (uint32_t st = hp >> 1; st; st >>= 1) {
    uint32_t p[U];
#pragma unroll
    for (int q = 0; q < U; q++) p[q] = P[k[q] + st];
#pragma unroll
    for (int q = 0; q < U; q++)
      if (p[q] <= e0 + 64 * q + lane) k[q] += st;
  }
}

// The same over a prefix held one entry per lane (X_lane, at most 64
// probes): two ballots give the window ends, lane exchanges the steps.
template <int U>
__device__ __forceinline__ void lane_window(uint32_t X, uint32_t e0, uint32_t e1, int lane, int (&k)[U]) {
  const int klo = static_cast<int>(__popcll(__ballot(X <= e0))) - 1;
  const int khi = static_cast<int>(__popcll(__ballot(X <= e1))) - 1;
#pragma unroll
  for (int q = 0; q < U; q++) k[q] = klo;
  const uint32_t w = static_cast<uint32_t>(khi - klo);
  if (!w) return;
  const uint32_t hp = 1u << (31 - __builtin_clz(w));
  const int d = klo + static_cast<int>(w + 1 - hp);
  const uint32_t xd = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(X), d));
#pragma unroll
  for (int q = 0; q < U; q++)
    if (xd <= e0 + 64 * q + lane) k[q] = d;
  for (int st = static_cast<int>(hp >> 1); st; st >>= 1) {
    uint32_t v[U];
#pragma unroll
    for (int q = 0; q < U; q++) v[q] = static_cast<uint32_t>(__shfl(X, k[q] + st));
#pragma unroll
    for (int q = 0; q < U; q++)
      if (v[q] <= e0 + 64 * q + lane) k[q] += st;
  }
}

// Adds one row's hits to the per-probe counters ctr[k] in LDS.  The probe
// index k is nondecreasing over the lanes (the window search guarantees it,
// lanes past the end included), so the lanes fall into segments of equal k:
// the first lane of each segment adds the popcount of its segment's hits --
// one LDS atomic per probe of the row, never two lanes on one address (a
// per-hit atomic would serialise all the hits of a probe on one address).
__device__ __forceinline__ void add_row_hits(int32_t *ctr, int k, bool hit, int lane) {
  const unsigned long long hits = __ballot(hit);
  if (!hits) return;
  typedef __attribute__((address_space(3))) int32_t lds_i32;
  const int k0 = __builtin_amdgcn_readfirstlane(k);
  if (k0 == __builtin_amdgcn_readlane(k, 63)) {  // one probe: one add
    if (lane == 0)
      __hip_atomic_fetch_add((lds_i32 *)&ctr[k0], static_cast<int>(__popcll(hits)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
    return;
  }
  const bool head = lane == 0 || k != __shfl_up(k, 1, 64);
  const unsigned long long heads = __ballot(head);
  if (head) {
    const unsigned long long above = heads & ~((2ull << lane) - 1);  // segment heads after this lane
    const unsigned long long seg = (above ? (above & (0ull - above)) - 1 : ~0ull) & ~((1ull << lane) - 1);
    const int c = static_cast<int>(__popcll(hits & seg));
    if (c)
      __hip_atomic_fetch_add((lds_i32 *)&ctr[k], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// The round's last element below `end` (end > e0).
template <int U>
__device__ __forceinline__ uint32_t round_last(uint32_t e0, uint32_t end) {
  return e0 + 64 * U - 1 < end ? e0 + 64 * U - 1 : end - 1;
}

}  // namespace accflat
