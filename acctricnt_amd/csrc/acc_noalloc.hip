// Triangle total of a CSR that already lies in device memory, with no
// allocation at all (acc_triangle_total_device in include/acc_cn.h): for
// callers that cannot allocate device memory, such as code under HIP graph
// capture or a serving loop that must not call hipMalloc.
//
// Every triangle {a < b < c} is counted once at its middle vertex b: for each
// upper slot (b, c), the elements a of N(c) below b that also lie in N<(b).
// One wave per source b in a grid-stride loop.  Per round the wave takes 64
// upper slots of b, finds each probe's length |N(c) ∩ [0, b)| by a binary
// search of N(c), and lays the round's probes end to end, each lane taking one
// element per step (its probe by a 6-step search over the lanes' prefix).
// N<(b) is staged in LDS when it has at most kStage ids; membership is a
// binary search there, or in global memory beyond that.
//
// The only device memory beyond the caller's inputs and output is the
// file-scope array `scratch`: one partial sum per workgroup of the counting
// grid (each written with a plain store, so none needs clearing) and an error
// word, cleared by a kernel at the start of every call.  A list that leaves
// [0, m) or an id outside [0, n) sets the error word, the kernels never read
// outside the caller's arrays, and the total is then written as ~0.  Its
// device address is looked up once per device (hipGetSymbolAddress) when a
// context is created there; host code uses that address, never the array's
// name.  Calls on one device share the array, so calls that overlap in time
// on one device race.

#include "acc_internal.h"

namespace {

constexpr int kWG = 256;                  // threads per workgroup: 4 waves
constexpr int kWavesPerWG = kWG / 64;
constexpr uint32_t kParts = 2048;         // most workgroups of the counting grid, one partial sum each
constexpr uint32_t kErrWord = kParts;     // index of the error word
constexpr uint32_t kStage = 1024;         // N<(b) staged in LDS up to this many ids (per wave)
constexpr uint32_t kScratchWords = kParts + 32;  // partial sums, then the error word on its own 256 bytes

static __device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch_addr[kMaxDevices] = {};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t excl_scan(uint32_t v, int lane) {
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

// Largest lane k with X_k <= e, for X nondecreasing over the lanes and X_0 <= e.
__device__ __forceinline__ int lane_of(uint32_t X, uint32_t e) {
  int k = 0;
#pragma unroll
  for (int st = 32; st; st >>= 1)
    if (static_cast<uint32_t>(__shfl(X, k + st)) <= e) k += st;
  return k;
}

// First position in a[0, len) holding a value >= x (len when none).
template <class P>
__device__ __forceinline__ uint32_t lower_bound(P a, uint32_t len, int32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(64) void k_na_clear(unsigned long long *__restrict__ s) {
  if (threadIdx.x == 0) s[kErrWord] = 0;
}

__global__ __launch_bounds__(kWG) void k_na_count(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                  uint32_t n, uint64_t m, unsigned long long *__restrict__ s) {
  __shared__ int32_t stage[kWavesPerWG][kStage];
  __shared__ unsigned long long wsum[kWavesPerWG];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t waves = gridDim.x * kWavesPerWG;
  int32_t *st = stage[wave];
  unsigned long long hits = 0;
  bool bad = false;
  for (uint32_t b = blockIdx.x * kWavesPerWG + wave; b < n; b += waves) {
    const uint32_t s0 = off[b], e0 = off[b + 1];
    if (e0 < s0 || e0 > m) {
      bad = true;
      continue;
    }
    // N<(b) = adj[s0, s0 + nl): ids below b; upper slots [up, e0): ids above b
    const uint32_t nl = lower_bound(adj + s0, e0 - s0, static_cast<int32_t>(b));
    const uint32_t up = s0 + nl + (s0 + nl < e0 && adj[s0 + nl] == static_cast<int32_t>(b) ? 1u : 0u);
    if (nl == 0 || up == e0) continue;
    const bool staged = nl <= kStage;
    if (staged) {
      wave_lds_sync();  // the previous source's reads of the stage are done
      for (uint32_t i = lane; i < nl; i += 64) st[i] = adj[s0 + i];
      wave_lds_sync();
    }
    for (uint32_t j0 = up; j0 < e0; j0 += 64) {
      const uint32_t j = j0 + lane;
      uint32_t cb = 0, len = 0;
      if (j < e0) {
        const int32_t c = adj[j];
        if (c < 0 || static_cast<uint32_t>(c) >= n) {
          bad = true;
        } else {
          const uint32_t cs = off[c], ce = off[c + 1];
          if (ce < cs || ce > m) {
            bad = true;
          } else {
            cb = cs;
            len = lower_bound(adj + cs, ce - cs, static_cast<int32_t>(b));
          }
        }
      }
      const uint32_t P = excl_scan(len, lane);
      const uint32_t total = __builtin_amdgcn_readfirstlane(__shfl(P + len, 63));
      for (uint32_t e1 = 0; e1 < total; e1 += 64) {
        const uint32_t e = e1 + lane;
        const int k = lane_of(P, e < total ? e : total - 1);
        const uint32_t pk = __shfl(P, k), ck = __shfl(cb, k);
        const uint32_t idx = ck + (e - pk);
        if (e < total && idx >= m) bad = true;  // only a malformed CSR gets here
        if (e < total && idx < m) {
          const int32_t x = adj[idx];
          uint32_t pos;
          bool in;
          if (staged) {
            pos = lower_bound(st, nl, x);
            in = pos < nl && st[pos] == x;
          } else {
            pos = lower_bound(adj + s0, nl, x);
            in = pos < nl && adj[s0 + pos] == x;
          }
          hits += in;
        }
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) hits += __shfl_xor(hits, o, 64);
  if (lane == 0) wsum[wave] = hits;
  if (__ballot(bad) && lane == 0) atomicOr(&s[kErrWord], 1ull);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerWG; w++) t += wsum[w];
    s[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kWG) void k_na_reduce(const unsigned long long *__restrict__ s, uint32_t parts,
                                                   unsigned long long *__restrict__ out) {
  __shared__ unsigned long long wsum[kWavesPerWG];
  unsigned long long v = 0;
  for (uint32_t i = threadIdx.x; i < parts; i += kWG) v += s[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerWG; w++) t += wsum[w];
    *out = s[kErrWord] ? ~0ull : t;
  }
}

}  // namespace

namespace accnoalloc {

// Device address of `scratch` on the calling thread's current device, looked
// up once per device.
int bind_scratch(int device, unsigned long long **addr) {
  if (device < 0 || device >= kMaxDevices) return acc_set_error(ACC_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return acc_set_error(ACC_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[device] = static_cast<unsigned long long *>(p);
  }
  *addr = g_scratch_addr[device];
  return ACC_OK;
}

}  // namespace accnoalloc

extern "C" {

int acc_triangle_total_device(acc_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *d_node_off,
                              const int32_t *d_edge_dst, unsigned long long *d_triangles) {
  if (!ctx || !d_node_off || !d_triangles || (m && !d_edge_dst)) return acc_set_error(ACC_EINVAL, "null argument");
  if (n > 0x7fffffffu) return acc_set_error(ACC_EINVAL, "n must be below 2^31 (int32 vertex ids)");
  if (m > 0xffffff00ull) return acc_set_error(ACC_EINVAL, "m must be at most 2^32 - 256 (uint32 offsets)");
  if (!ctx->na_scratch) return acc_set_error(ACC_EDEVICE, "the context has no scratch address");
  ACC_HIP_TRY(hipSetDevice(ctx->device));
  unsigned long long *s = ctx->na_scratch;
  // one wave per source at least, at most kParts workgroups: the scratch holds one partial per workgroup
  const uint32_t want = (n + kWavesPerWG - 1) / kWavesPerWG;
  const uint32_t parts = want == 0 ? 1u : (want < kParts ? want : kParts);
  k_na_clear<<<1, 64, 0, ctx->stream>>>(s);
  k_na_count<<<parts, kWG, 0, ctx->stream>>>(d_node_off, d_edge_dst, n, m, s);
  k_na_reduce<<<1, kWG, 0, ctx->stream>>>(s, parts, d_triangles);
  ACC_HIP_TRY(hipGetLastError());
  return ACC_OK;
}

}  // extern "C"
