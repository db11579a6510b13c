// Upload-time preparation of a resident CSR on the device.
//
// Slot sources: every non-empty vertex o marks its first slot with o in a
// zeroed array and an inclusive max-scan spreads the marks, so slot j gets
// the vertex whose range holds it.  The work is per slot, not per vertex, so
// a hub costs no more than its slots; validation, the reverse-slot build and
// the relabel's fill are slot-parallel on top of it.
//
// Degree relabel of a resident CSR on the device: new ids by degree,
// descending, ties by ascending old id, every adjacency list mapped and
// sorted ascending.  This is the deterministic form of the reference's
// ReorderDegDescending + Reorder (scan-query/scan_xp.cpp:1279-1301 builds
// the order with an unstable parallel sort, :1145-1222 rebuilds and re-sorts
// the adjacency lists); the old id of every new vertex is kept on the graph
// (the reference's old_vid_dict) so cluster ids come out in old ids.
//
// Steps, all HBM-bound: degrees (n), a stable radix sort of (degree, id)
// pairs by degree descending (n), the inverse permutation and the new
// offsets (an exclusive scan of the sorted degrees), one thread per old slot
// writing its neighbor's new id at the new position (m), and a segmented
// radix sort of the new lists over ceil(log2 n) key bits (m).
//
// Reverse slots: the upper slots (v, c), c > v, taken in CSR order and
// sorted stably by target c, come out grouped by c with v ascending inside
// each group, which is the order of c's lower slots; so the k-th sorted
// entry's reverse is c's lower slot k - L[c], where L is the exclusive scan
// of the lower-slot counts.  This replaces the reference's per-edge binary
// search for the reverse offset (scan-query-cuda/experimental_scan_xp_multigpu_multipass.cu:6-23,
// scan-query/scan_xp.cpp:88-92) by HBM-streaming passes plus one scattered
// write per upper slot.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>

#include "acc_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_rl_keys(const uint32_t *__restrict__ off, uint32_t n,
                                                 uint32_t *__restrict__ deg, int32_t *__restrict__ id) {
  for (uint32_t v = blockIdx.x * 256u + threadIdx.x; v < n; v += gridDim.x * 256u) {
    deg[v] = off[v + 1] - off[v];
    id[v] = static_cast<int32_t>(v);
  }
}

// x is a new id; deg_sorted[x] is its degree.  len gets n + 1 entries, the
// last 0, so that its exclusive scan ends with m.
__global__ __launch_bounds__(256) void k_rl_inverse(const uint32_t *__restrict__ deg_sorted,
                                                    const int32_t *__restrict__ old_of_new, uint32_t n,
                                                    int32_t *__restrict__ new_of_old, uint32_t *__restrict__ len) {
  for (uint32_t x = blockIdx.x * 256u + threadIdx.x; x <= n; x += gridDim.x * 256u) {
    if (x == n) {
      len[n] = 0;
    } else {
      new_of_old[old_of_new[x]] = static_cast<int32_t>(x);
      len[x] = deg_sorted[x];
    }
  }
}

// Every non-empty vertex o marks its first slot with o (heads is zeroed
// first); an inclusive max-scan then gives the source of every slot.  Needs
// only monotone offsets ending at m.
__global__ __launch_bounds__(256) void k_rl_heads(const uint32_t *__restrict__ off, uint32_t n,
                                                  uint32_t *__restrict__ heads) {
  for (uint32_t o = blockIdx.x * 256u + threadIdx.x; o < n; o += gridDim.x * 256u) {
    const uint32_t b = off[o];
    if (off[o + 1] > b) heads[b] = o;
  }
}

// Old slot j of source o: its neighbor's new id goes to position
// j - off[o] of the new vertex's list (old order; sorted afterwards).
__global__ __launch_bounds__(256) void k_rl_fill(const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                 const uint32_t *__restrict__ src, uint64_t m,
                                                 const int32_t *__restrict__ new_of_old,
                                                 const uint32_t *__restrict__ new_off, uint32_t *__restrict__ out) {
  for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < m; j += gridDim.x * 256ull) {
    const uint32_t o = src[j];
    out[new_off[new_of_old[o]] + (j - off[o])] = static_cast<uint32_t>(new_of_old[adj[j]]);
  }
}

// low[v] = number of lower slots of v (neighbors below v), low[n] = 0.
__global__ __launch_bounds__(256) void k_rev_low(const uint32_t *__restrict__ off, const uint32_t *__restrict__ hs,
                                                 uint32_t n, uint32_t *__restrict__ low) {
  for (uint32_t v = blockIdx.x * 256u + threadIdx.x; v <= n; v += gridDim.x * 256u)
    low[v] = v < n ? hs[v] - off[v] : 0;
}

// Upper slot j of source v (j >= hs[v]) is the (j - L[v + 1])-th upper slot
// in CSR order (the slots before it, less the lower slots of vertices up to
// v); it is written there with its target as the key.
__global__ __launch_bounds__(256) void k_rev_keys(const int32_t *__restrict__ adj, const uint32_t *__restrict__ src,
                                                  const uint32_t *__restrict__ hs, const uint32_t *__restrict__ L,
                                                  uint64_t m, uint32_t n, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals) {
  for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < m; j += gridDim.x * 256ull) {
    const uint32_t v = src[j];
    if (j >= hs[v]) {
      const uint32_t k = static_cast<uint32_t>(j) - L[v + 1];
      const int32_t c = adj[j];
      // ids are checked unless the caller vouches for them; an id out of
      // range still keeps every access below in bounds
      keys[k] = c >= 0 && static_cast<uint32_t>(c) < n ? static_cast<uint32_t>(c) : n - 1;
      vals[k] = static_cast<uint32_t>(j);
    }
  }
}

// Sorted entry k = (target c, upper slot j): its reverse is c's lower slot
// t = off[c] + k - L[c], which must lie among c's lower slots and hold the
// source of j.  When every entry passes and the upper slots are half of all
// slots, the entries map one-to-one onto the lower slots, so the CSR is
// symmetric; a failing entry sets flag 8 and leaves rev[j] at ~0.
__global__ __launch_bounds__(256) void k_rev_fill(const uint32_t *__restrict__ keys,
                                                  const uint32_t *__restrict__ vals, uint32_t count,
                                                  const uint32_t *__restrict__ off, const int32_t *__restrict__ adj,
                                                  const uint32_t *__restrict__ L, uint32_t n,
                                                  uint32_t *__restrict__ rev, int32_t *__restrict__ flags) {
  int32_t f = 0;
  for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < count; k += gridDim.x * 256u) {
    const uint32_t c = keys[k], j = vals[k];
    const uint32_t r = k - L[c];
    bool ok = false;
    if (r < L[c + 1] - L[c]) {
      const uint32_t t = off[c] + r;
      const int32_t v = adj[t];
      if (v >= 0 && static_cast<uint32_t>(v) < n && off[v] <= j && j < off[v + 1]) {
        rev[j] = t;
        ok = true;
      }
    }
    if (!ok) f = 8;
  }
  if (flags && f) atomicOr(flags, f);
}

unsigned rl_grid(uint64_t items, int cus) {
  const uint64_t want = (items + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(cus > 0 ? cus : 256) * 8;
  return static_cast<unsigned>(want < cap ? (want ? want : 1) : cap);
}

int bit_width(uint32_t x) { return x ? 32 - __builtin_clz(x) : 1; }

}  // namespace

namespace accprep {

int slot_sources(acc_graph *g) {
  const uint64_t m = g->m;
  if (!m) return ACC_OK;
  hipStream_t st = g->ctx->stream;
  uint32_t *heads = g->d_rev;  // the reverse slots are built from the sources afterwards
  uint32_t *src = reinterpret_cast<uint32_t *>(g->d_cnt);
  size_t t_bytes = 0;
  void *temp = nullptr;
  hipError_t e = hipcub::DeviceScan::InclusiveScan(nullptr, t_bytes, heads, src, hipcub::Max(), m, st);
  if (e == hipSuccess) e = hipMalloc(&temp, t_bytes ? t_bytes : 1);
  if (e == hipSuccess) e = hipMemsetAsync(heads, 0, sizeof(uint32_t) * m, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rl_heads, dim3(rl_grid(g->n, g->ctx->num_cus)), dim3(256), 0, st, g->d_off, g->n, heads);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::InclusiveScan(temp, t_bytes, heads, src, hipcub::Max(), m, st);
  (void)hipStreamSynchronize(st);
  if (temp) (void)hipFree(temp);
  if (e != hipSuccess)
    return acc_set_error(e == hipErrorOutOfMemory ? ACC_ENOMEM : ACC_EDEVICE,
                         std::string("slot sources: ") + hipGetErrorString(e));
  return ACC_OK;
}

int reverse_slots(acc_graph *g, bool validate, uint64_t *upper) {
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  *upper = 0;
  if (!m) return ACC_OK;
  hipStream_t st = g->ctx->stream;
  const int cus = g->ctx->num_cus;
  uint32_t *low = nullptr, *L = nullptr;
  void *temp = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(st);
    for (void *p : {static_cast<void *>(low), static_cast<void *>(L), temp})
      if (p) (void)hipFree(p);
  };
  auto fail = [&](const char *what, hipError_t e) {
    cleanup();
    return acc_set_error(e == hipErrorOutOfMemory ? ACC_ENOMEM : ACC_EDEVICE,
                         std::string("reverse slots ") + what + ": " + hipGetErrorString(e));
  };
  const size_t nb1 = sizeof(uint32_t) * (static_cast<size_t>(n) + 1);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&low), nb1);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&L), nb1);
  size_t t_scan = 0;
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, low, L, static_cast<int>(n) + 1, st);
  if (e == hipSuccess) e = hipMalloc(&temp, t_scan ? t_scan : 1);
  if (e != hipSuccess) return fail("allocation", e);
  hipLaunchKernelGGL(k_rev_low, dim3(rl_grid(static_cast<uint64_t>(n) + 1, cus)), dim3(256), 0, st, g->d_off,
                     g->d_hs, n, low);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(temp, t_scan, low, L, static_cast<int>(n) + 1, st);
  uint32_t lower = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&lower, L + n, sizeof(lower), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail("lower-slot offsets", e);
  const uint64_t count = m - lower;
  *upper = count;
  if (validate && 2 * count != m) {  // not symmetric: the caller rejects the graph
    cleanup();
    return ACC_OK;
  }
  // The sort's input pairs go into the reverse-slot array and its output into
  // the count array (slot sources, read by k_rev_keys before the sort writes).
  // A graph too lopsided for that, trusted without validation, takes the
  // per-slot binary search instead.
  if (2 * count > m + 4 || count > static_cast<uint64_t>(INT32_MAX)) {
    cleanup();
    return accfwd::launch_build_rev(g, false);
  }
  uint32_t *keys_in = g->d_rev, *vals_in = g->d_rev + count;
  uint32_t *keys_out = reinterpret_cast<uint32_t *>(g->d_cnt), *vals_out = keys_out + count;
  const int bits = bit_width(n - 1);
  // double-buffered: the sort ping-pongs between the two arrays instead of
  // taking a copy of its input as temporary storage
  hipcub::DoubleBuffer<uint32_t> kb(keys_in, keys_out), vb(vals_in, vals_out);
  size_t t_sort = 0;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, t_sort, kb, vb, static_cast<int>(count), 0, bits, st);
  if (e == hipSuccess && t_sort > t_scan) {
    (void)hipFree(temp);
    temp = nullptr;
    e = hipMalloc(&temp, t_sort);
  }
  if (e != hipSuccess) return fail("sort storage", e);
  hipLaunchKernelGGL(k_rev_keys, dim3(rl_grid(m, cus)), dim3(256), 0, st, g->d_adj,
                     reinterpret_cast<const uint32_t *>(g->d_cnt), g->d_hs, L, m, n, keys_in, vals_in);
  e = hipGetLastError();
  kb = hipcub::DoubleBuffer<uint32_t>(keys_in, keys_out);  // as before the size query
  vb = hipcub::DoubleBuffer<uint32_t>(vals_in, vals_out);
  if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(temp, t_sort, kb, vb, static_cast<int>(count), 0, bits, st);
  // the fill rewrites the reverse-slot array: the sorted pairs must lie in the count array
  if (e == hipSuccess && kb.Current() != keys_out)
    e = hipMemcpyAsync(keys_out, kb.Current(), sizeof(uint32_t) * count, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess && vb.Current() != vals_out)
    e = hipMemcpyAsync(vals_out, vb.Current(), sizeof(uint32_t) * count, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(g->d_rev, 0xff, sizeof(uint32_t) * (m + 4), st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rev_fill, dim3(rl_grid(count, cus)), dim3(256), 0, st, keys_out, vals_out,
                       static_cast<uint32_t>(count), g->d_off, g->d_adj, L, n, g->d_rev,
                       validate ? g->d_flags : nullptr);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail("build", e);
  cleanup();
  return ACC_OK;
}

int relabel_degree(acc_graph *g) {
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  if (n == 0) return ACC_OK;
  if (m >= (1ULL << 31) || n >= (1u << 31) - 1)
    return acc_set_error(ACC_EINVAL, "the device relabel needs m < 2^31 and n < 2^31 - 1");
  hipStream_t st = g->ctx->stream;
  const int cus = g->ctx->num_cus;
  const int bits = bit_width(n - 1);  // ids and degrees are below n
  uint32_t *deg = nullptr, *deg_sorted = nullptr, *len = nullptr, *new_off = nullptr;
  int32_t *id = nullptr, *old_of_new = nullptr, *new_of_old = nullptr;
  void *temp = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(st);
    for (void *p : {static_cast<void *>(deg), static_cast<void *>(deg_sorted), static_cast<void *>(len),
                    static_cast<void *>(new_off), static_cast<void *>(id), static_cast<void *>(new_of_old), temp})
      if (p) (void)hipFree(p);
  };
  auto fail = [&](const char *what, hipError_t e) {
    cleanup();
    if (old_of_new) (void)hipFree(old_of_new);
    return acc_set_error(e == hipErrorOutOfMemory ? ACC_ENOMEM : ACC_EDEVICE,
                         std::string("relabel ") + what + ": " + hipGetErrorString(e));
  };
  const size_t nb = sizeof(uint32_t) * n, nb1 = sizeof(uint32_t) * (static_cast<size_t>(n) + 1);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&deg), nb);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&deg_sorted), nb);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&id), nb);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&old_of_new), nb);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&new_of_old), nb);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&len), nb1);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&new_off), nb1);
  if (e != hipSuccess) return fail("allocation", e);
  // cub temp storage: the largest of the three primitives' needs.  d_cnt
  // holds the slot sources (slot_sources); the reverse-slot array is free
  // until the reverse slots are built and takes the mapped lists.
  size_t t_sort = 0, t_scan = 0, t_seg = 0;
  uint32_t *mapped = g->d_rev;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(g->d_cnt);
  uint32_t *lists = reinterpret_cast<uint32_t *>(g->d_adj);
  e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, t_sort, deg, deg_sorted, id, old_of_new,
                                                   static_cast<int>(n), 0, bits, st);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, len, new_off, static_cast<int>(n) + 1, st);
  // the lists' sort is double-buffered between the mapped lists and the
  // adjacency array (the old lists are dead once mapped), so its temporary
  // storage holds no copy of the m keys
  hipcub::DoubleBuffer<uint32_t> lb(mapped, lists);
  if (e == hipSuccess && m)
    e = hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, t_seg, lb, static_cast<int>(m), static_cast<int>(n),
                                                   new_off, new_off + 1, 0, bits, st);
  size_t t_bytes = std::max(t_sort, std::max(t_scan, t_seg));
  if (e == hipSuccess) e = hipMalloc(&temp, t_bytes ? t_bytes : 1);
  if (e != hipSuccess) return fail("temporary storage", e);

  // device time of each step, read back with acc_cn_timings
  for (hipEvent_t ev : g->events) (void)hipEventDestroy(ev);
  g->events.clear();
  g->event_names.clear();
  g->timing_valid = false;
  int rc = record_event(g, "begin");
  if (rc) {
    cleanup();
    (void)hipFree(old_of_new);
    return rc;
  }
  hipLaunchKernelGGL(k_rl_keys, dim3(rl_grid(n, cus)), dim3(256), 0, st, g->d_off, n, deg, id);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(temp, t_bytes, deg, deg_sorted, id, old_of_new,
                                                     static_cast<int>(n), 0, bits, st);
  if (e != hipSuccess) return fail("degree order", e);
  (void)record_event(g, "rl_degree_order");
  hipLaunchKernelGGL(k_rl_inverse, dim3(rl_grid(static_cast<uint64_t>(n) + 1, cus)), dim3(256), 0, st, deg_sorted,
                     old_of_new, n, new_of_old, len);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(temp, t_bytes, len, new_off, static_cast<int>(n) + 1, st);
  if (e != hipSuccess) return fail("offsets", e);
  (void)record_event(g, "rl_offsets");
  if (m) {
    hipLaunchKernelGGL(k_rl_fill, dim3(rl_grid(m, cus)), dim3(256), 0, st, g->d_off, g->d_adj, src, m, new_of_old,
                       new_off, mapped);
    e = hipGetLastError();
    if (e == hipSuccess) (void)record_event(g, "rl_fill");
    lb = hipcub::DoubleBuffer<uint32_t>(mapped, lists);  // as before the size query
    if (e == hipSuccess)
      e = hipcub::DeviceSegmentedRadixSort::SortKeys(temp, t_bytes, lb, static_cast<int>(m), static_cast<int>(n),
                                                     new_off, new_off + 1, 0, bits, st);
    if (e == hipSuccess && lb.Current() != lists)
      e = hipMemcpyAsync(lists, lb.Current(), sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail("adjacency lists", e);
    (void)record_event(g, "rl_sort_lists");
  }
  e = hipMemcpyAsync(g->d_off, new_off, nb1, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return fail("offsets copy", e);
  (void)record_event(g, "rl_copy_offsets");
  cleanup();
  if (g->d_old) (void)hipFree(g->d_old);
  else g->device_bytes += nb;
  g->d_old = old_of_new;
  g->relabelled = true;
  return ACC_OK;
}

}  // namespace accprep
