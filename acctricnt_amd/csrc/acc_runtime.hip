// C ABI of libacc_cn.so (include/acc_cn.h): contexts, resident graphs,
// work plans and the counting pipeline.

#include <algorithm>
#include <cstring>
#include <numeric>
#include <thread>
#include <new>

#include "acc_internal.h"

namespace {
thread_local std::string g_err;

struct DeviceGuard {
  explicit DeviceGuard(int dev) { (void)hipSetDevice(dev); }
};

template <typename T>
int dev_alloc(acc_graph *g, T **p, uint64_t count) {
  const size_t bytes = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return acc_set_error(ACC_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                         hipGetErrorString(e));
  }
  g->device_bytes += bytes;
  return ACC_OK;
}

void free_graph(acc_graph *g) {
  if (!g) return;
  DeviceGuard guard(g->ctx->device);
  for (void *p : {static_cast<void *>(g->d_off), static_cast<void *>(g->d_adj), static_cast<void *>(g->d_cnt),
                  static_cast<void *>(g->d_rev), static_cast<void *>(g->d_sum), static_cast<void *>(g->d_flags),
                  static_cast<void *>(g->d_fwd_wave), static_cast<void *>(g->d_fwd_block),
                  static_cast<void *>(g->d_fwd_big), static_cast<void *>(g->d_src),
                  static_cast<void *>(g->d_old), static_cast<void *>(g->edge.d_units[kEdgeWave]),
                  static_cast<void *>(g->edge.d_units[kEdgeHash12]),
                  static_cast<void *>(g->edge.d_units[kEdgeHash14]), static_cast<void *>(g->edge.d_units[kEdgeHash15]),
                  static_cast<void *>(g->edge.d_units[kEdgeHub]), static_cast<void *>(g->edge.d_hubs),
                  static_cast<void *>(g->edge.d_pool), static_cast<void *>(g->d_skey),
                  static_cast<void *>(g->d_snbr), static_cast<void *>(g->d_sslot),
                  static_cast<void *>(g->d_tri_parts), static_cast<void *>(g->d_hs)})
    if (p) (void)hipFree(p);
  if (g->graph_exec) (void)hipGraphExecDestroy(g->graph_exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_gvoff) (void)hipFree(g->d_gvoff);
  if (g->d_gsoff) (void)hipFree(g->d_gsoff);
  if (g->d_gtot) (void)hipFree(g->d_gtot);
  for (hipEvent_t e : g->events) (void)hipEventDestroy(e);
  delete g;
}

// Per-edge plan: chunks of at most `chunk` upper slots of every source, by
// d_u: wave units (d_u <= 256 and at most 64 upper slots, one unit per
// source), LDS hash classes of 2^12, 2^14 and 2^15 keys for d_u up to 2048,
// 8192 and min(16384, hash_max); the rest are hubs.  Hubs go into batches of at
// most pool_hubs (one HBM bitmap each); a hub unit's nlow is its hub's
// bitmap within the batch.
constexpr uint64_t kHubPoolBytes = 8ull << 30;  // HBM bitmaps of one hub batch
constexpr uint32_t kHubBatch = 16384;           // most hubs of one batch (ACC_PARAM_BMP_BATCH)
constexpr uint32_t kHashIdxBits = 1u << 17;     // LDS index bits of a hash unit (k_bmp_hash's kIdxWords)
constexpr uint32_t kHubIdxBits = 1u << 18;      // LDS index bits of a hub unit (k_bmp_hub's kHubIdxWords)
void build_edge_plan(uint32_t n, const uint32_t *off, const uint32_t *hs, uint32_t chunk, uint32_t hash_max,
                     uint32_t pool_hubs, std::vector<FwdUnit> (&units)[kEdgeClasses], std::vector<uint32_t> &hubs,
                     std::vector<uint32_t> &hub_batch, std::vector<uint32_t> &hub_unit_batch) {
  hub_batch.assign(1, 0);
  hub_unit_batch.assign(1, 0);
  for (uint32_t u = 0; u < n; u++) {
    const uint32_t du = off[u + 1] - off[u];
    if (hs[u] == off[u + 1]) continue;
    int c = kEdgeHub;
    if (du <= kEdgeWaveDeg && off[u + 1] - hs[u] <= kEdgeWaveUp)
      c = kEdgeWave;
    else if (du <= hash_max)
      c = du <= 2048 ? kEdgeHash12 : du <= 8192 ? kEdgeHash14 : kEdgeHash15;
    uint32_t slot = 0;
    if (c == kEdgeHub) {
      if (hubs.size() - hub_batch.back() == pool_hubs) {
        hub_batch.push_back(static_cast<uint32_t>(hubs.size()));
        hub_unit_batch.push_back(static_cast<uint32_t>(units[kEdgeHub].size()));
      }
      slot = static_cast<uint32_t>(hubs.size()) - hub_batch.back();
      hubs.push_back(u);
    }
    for (uint32_t j0 = hs[u]; j0 < off[u + 1]; j0 += chunk)
      units[c].push_back(FwdUnit{static_cast<int32_t>(u), j0, std::min(j0 + chunk, off[u + 1]), c == kEdgeHub ? slot : du});
  }
  hub_batch.push_back(static_cast<uint32_t>(hubs.size()));
  hub_unit_batch.push_back(static_cast<uint32_t>(units[kEdgeHub].size()));
}

template <typename T>
void dev_free(acc_graph *g, T **p, uint64_t count) {
  if (*p) {
    (void)hipFree(*p);
    g->device_bytes -= sizeof(T) * (count ? count : 1);
  }
  *p = nullptr;
}

// Drops the captured counting pass: it holds the addresses and launch shapes
// of the plans and partition arrays, so any change to those invalidates it.
void drop_capture(acc_graph *g) {
  if (g->graph_exec) (void)hipGraphExecDestroy(g->graph_exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->graph_exec = nullptr;
  g->graph = nullptr;
}

// A host unit list and the device copy it replaces.
struct PlanList {
  std::vector<FwdUnit> *v;
  FwdUnit **d;
  uint32_t *count;
};

// Replaces the device copies of the given unit lists.
template <size_t N>
int upload_lists(acc_graph *g, PlanList (&lists)[N]) {
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  for (auto &l : lists) {
    dev_free(g, l.d, *l.count);
    *l.count = static_cast<uint32_t>(l.v->size());
    if (l.v->empty()) continue;
    ACC_TRY(dev_alloc(g, l.d, l.v->size()));
    ACC_HIP_TRY(hipMemcpyAsync(*l.d, l.v->data(), sizeof(FwdUnit) * l.v->size(), hipMemcpyHostToDevice,
                               g->ctx->stream));
  }
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  return ACC_OK;
}

// Rebuilds the forward plan from the host split points, the current bin
// parameters and the shard; the per-edge plans are rebuilt at their next use.
int upload_plans(acc_graph *g) {
  drop_capture(g);
  std::vector<FwdUnit> wave, block, big;
  const uint32_t device_waves = static_cast<uint32_t>(std::max(1, g->ctx->num_cus)) * 32u;
  build_forward_plan(g->shard_lo, g->shard_hi, g->h_off.data(), g->h_hs.data(), g->h_lo.data(), g->wave_slots,
                     g->wave_low, g->chunk, device_waves, wave, block, big);
  // Workgroup units in a stride order (the golden-ratio stride, coprime with
  // their number, so a permutation): the tasks that run at the same time
  // then come from distant sources.  At R-MAT s22 this took the
  // triangle-only pass from 6.7 to 6.1-6.4 ms and the per-slot pass 2%
  // faster; the wave packs keep source order (permuting them as well slowed
  // the 512-graph batch by 9%), and so do the units of a packed multi-graph
  // CSR, whose graphs' units then stay together (its replay 3% faster).
  // Counts do not depend on the order.
  if (block.size() > 1 && !g->num_parts) {
    const size_t n = block.size();
    size_t S = static_cast<size_t>(static_cast<double>(n) * 0.6180339887) | 1u;
    while (std::gcd(S, n) != 1) S += 2;
    std::vector<FwdUnit> p(n);
    for (size_t j = 0; j < n; j++) p[j] = block[(j * S) % n];
    block.swap(p);
  }
  PlanList lists[] = {{&wave, &g->d_fwd_wave, &g->n_fwd_wave},
                      {&block, &g->d_fwd_block, &g->n_fwd_block},
                      {&big, &g->d_fwd_big, &g->n_fwd_big}};
  g->edge.valid = false;
  return upload_lists(g, lists);
}

// The per-edge (bitmap and merge) plans, built at the first pass that needs
// them: most graphs only ever run the forward path.
int ensure_edge_plans(acc_graph *g) {
  EdgePlan &p = g->edge;
  if (p.valid) return ACC_OK;
  p.pool_words = std::max<uint64_t>(1, (static_cast<uint64_t>(g->n) + 31) / 32);
  p.idx_shift = 0;
  while (((p.pool_words + (1ull << p.idx_shift) - 1) >> p.idx_shift) > kHashIdxBits) p.idx_shift++;
  const uint64_t groups = (static_cast<uint64_t>(g->n) + 15) / 16;  // the hub index's 16-id groups
  p.hub_idx_shift = 0;
  while (((groups + (1ull << p.hub_idx_shift) - 1) >> p.hub_idx_shift) > kHubIdxBits) p.hub_idx_shift++;
  const uint32_t pool_hubs = static_cast<uint32_t>(
      std::min<uint64_t>(g->bmp_batch, std::max<uint64_t>(1, kHubPoolBytes / (sizeof(uint32_t) * p.pool_words))));
  std::vector<FwdUnit> units[kEdgeClasses];
  std::vector<uint32_t> hubs;
  build_edge_plan(g->n, g->h_off.data(), g->h_hs.data(), g->chunk, g->bmp_hash_max, pool_hubs, units, hubs,
                  p.hub_batch, p.hub_unit_batch);
  PlanList lists[] = {{&units[kEdgeWave], &p.d_units[kEdgeWave], &p.n_units[kEdgeWave]},
                      {&units[kEdgeHash12], &p.d_units[kEdgeHash12], &p.n_units[kEdgeHash12]},
                      {&units[kEdgeHash14], &p.d_units[kEdgeHash14], &p.n_units[kEdgeHash14]},
                      {&units[kEdgeHash15], &p.d_units[kEdgeHash15], &p.n_units[kEdgeHash15]},
                      {&units[kEdgeHub], &p.d_units[kEdgeHub], &p.n_units[kEdgeHub]}};
  ACC_TRY(upload_lists(g, lists));
  dev_free(g, &p.d_hubs, p.n_hubs);
  p.n_hubs = static_cast<uint32_t>(hubs.size());
  if (p.n_hubs) {
    ACC_TRY(dev_alloc(g, &p.d_hubs, p.n_hubs));
    ACC_HIP_TRY(hipMemcpyAsync(p.d_hubs, hubs.data(), sizeof(uint32_t) * p.n_hubs, hipMemcpyHostToDevice,
                               g->ctx->stream));
    ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  }
  p.valid = true;
  return ACC_OK;
}

// The hub bitmap pool (zeroed once: every bitmap pass clears what it sets).
int ensure_bitmap_pool(acc_graph *g) {
  EdgePlan &p = g->edge;
  uint32_t need = 0;
  for (size_t k = 0; k + 1 < p.hub_batch.size(); k++) need = std::max(need, p.hub_batch[k + 1] - p.hub_batch[k]);
  if (need <= p.pool_hubs) return ACC_OK;
  dev_free(g, &p.d_pool, static_cast<uint64_t>(p.pool_hubs) * p.pool_words);
  p.pool_hubs = 0;
  ACC_TRY(dev_alloc(g, &p.d_pool, static_cast<uint64_t>(need) * p.pool_words));
  ACC_HIP_TRY(hipMemsetAsync(p.d_pool, 0, sizeof(uint32_t) * need * p.pool_words, g->ctx->stream));
  p.pool_hubs = need;
  return ACC_OK;
}

int record(acc_graph *g, bool timing, const char *name) {
  if (!timing) return ACC_OK;
  hipEvent_t e;
  ACC_HIP_TRY(hipEventCreate(&e));
  g->events.push_back(e);
  g->event_names.emplace_back(name);
  ACC_HIP_TRY(hipEventRecord(e, g->ctx->stream));
  return ACC_OK;
}

// Copies between pageable host memory and the device near PCIe speed: the
// transfer is split into kCopyLanes contiguous parts, each moved by a host
// thread of its own through two pinned buffers of kStageBytes and a stream
// of its own, so copying one chunk into (or out of) pinned memory overlaps
// the DMA of the previous one.  Small transfers and host memory the device
// can address already (pinned) go through one hipMemcpyAsync.
constexpr int kCopyLanes = 4;
constexpr size_t kStageBytes = size_t(16) << 20;
constexpr size_t kStagedMin = size_t(8) << 20;

bool host_pinned(const void *p) {
  hipPointerAttribute_t attr;
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // pageable memory is not known to HIP: clear the error it leaves
    return false;
  }
  return attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int ensure_copy_lanes(acc_ctx *ctx) {
  if (!ctx->lanes.empty()) return ACC_OK;
  std::vector<CopyLane> lanes(kCopyLanes);
  hipError_t e = hipSuccess;
  for (CopyLane &l : lanes) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking);
    for (int b = 0; b < 2 && e == hipSuccess; b++) {
      e = hipHostMalloc(&l.buf[b], kStageBytes, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&l.ev[b], hipEventDisableTiming);
    }
  }
  if (e != hipSuccess) {
    for (CopyLane &l : lanes) {
      for (int b = 0; b < 2; b++) {
        if (l.buf[b]) (void)hipHostFree(l.buf[b]);
        if (l.ev[b]) (void)hipEventDestroy(l.ev[b]);
      }
      if (l.stream) (void)hipStreamDestroy(l.stream);
    }
    return acc_set_error(e == hipErrorOutOfMemory ? ACC_ENOMEM : ACC_EDEVICE,
                         std::string("copy lanes: ") + hipGetErrorString(e));
  }
  ctx->lanes = std::move(lanes);
  return ACC_OK;
}

// to_device: host src -> device dst; else device src -> host dst.  Ordered
// after the work already on ctx->stream; complete when it returns.
int staged_copy(acc_ctx *ctx, void *dst, const void *src, size_t bytes, bool to_device) {
  if (!bytes) return ACC_OK;
  const void *host = to_device ? src : static_cast<const void *>(dst);
  auto direct = [&]() {
    ACC_HIP_TRY(hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                               ctx->stream));
    ACC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ACC_OK;
  };
  if (bytes < kStagedMin || host_pinned(host)) return direct();
  std::lock_guard<std::mutex> lk(ctx->copy_mu);
  if (ensure_copy_lanes(ctx) != ACC_OK) return direct();  // no pinned memory to spare: copy directly
  hipEvent_t ready = nullptr;
  ACC_HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ready, ctx->stream);
  const size_t part = ((bytes + kCopyLanes - 1) / kCopyLanes + 4095) / 4096 * 4096;
  std::vector<hipError_t> errs(kCopyLanes, hipSuccess);
  auto lane_main = [&](int li) {
    CopyLane &L = ctx->lanes[li];
    const size_t lo = std::min(bytes, part * li), hi = std::min(bytes, part * (li + 1));
    hipError_t le = hipSetDevice(ctx->device);
    if (le == hipSuccess) le = hipStreamWaitEvent(L.stream, ready, 0);
    bool used[2] = {false, false};
    size_t pend_off = 0, pend_len = 0;
    int pend_b = -1;
    int k = 0;
    for (size_t o = lo; o < hi && le == hipSuccess; o += kStageBytes, k++) {
      const size_t len = std::min(kStageBytes, hi - o);
      const int b = k & 1;
      if (to_device) {
        if (used[b]) le = hipEventSynchronize(L.ev[b]);  // the chunk before last has left this buffer
        if (le != hipSuccess) break;
        std::memcpy(L.buf[b], static_cast<const char *>(src) + o, len);
        le = hipMemcpyAsync(static_cast<char *>(dst) + o, L.buf[b], len, hipMemcpyHostToDevice, L.stream);
        if (le == hipSuccess) le = hipEventRecord(L.ev[b], L.stream);
        used[b] = true;
      } else {
        le = hipMemcpyAsync(L.buf[b], static_cast<const char *>(src) + o, len, hipMemcpyDeviceToHost, L.stream);
        if (le == hipSuccess) le = hipEventRecord(L.ev[b], L.stream);
        if (le == hipSuccess && pend_b >= 0) {  // the previous chunk is out of its buffer once its event fires
          le = hipEventSynchronize(L.ev[pend_b]);
          if (le == hipSuccess) std::memcpy(static_cast<char *>(dst) + pend_off, L.buf[pend_b], pend_len);
        }
        pend_b = b;
        pend_off = o;
        pend_len = len;
      }
    }
    if (le == hipSuccess && !to_device && pend_b >= 0) {
      le = hipEventSynchronize(L.ev[pend_b]);
      if (le == hipSuccess) std::memcpy(static_cast<char *>(dst) + pend_off, L.buf[pend_b], pend_len);
    }
    const hipError_t se = hipStreamSynchronize(L.stream);
    errs[li] = le != hipSuccess ? le : se;
  };
  if (e == hipSuccess) {
    std::vector<std::thread> th;
    int spawned = 1;
    try {
      for (; spawned < kCopyLanes; spawned++) th.emplace_back(lane_main, spawned);
    } catch (...) {  // no thread to spare: the remaining lanes run on this one
    }
    lane_main(0);
    for (int li = spawned; li < kCopyLanes; li++) lane_main(li);
    for (auto &t : th) t.join();
    for (hipError_t le : errs)
      if (le != hipSuccess) e = le;
  }
  (void)hipEventDestroy(ready);
  if (e != hipSuccess) return acc_set_error(ACC_EDEVICE, std::string("staged copy: ") + hipGetErrorString(e));
  return ACC_OK;
}

// Plan inputs of every vertex, computed on the device (k_vertex_meta): d_hs
// stays resident, h_off, h_hs and h_lo are copied to the host for planning.
// host_off, when given, is the caller's offsets array (equal to d_off).
int load_vertex_meta(acc_graph *g, const uint32_t *host_off) {
  const uint32_t n = g->n;
  ACC_TRY(dev_alloc(g, &g->d_hs, static_cast<uint64_t>(n) + 1));
  int32_t *lo = nullptr;
  if (hipMalloc(reinterpret_cast<void **>(&lo), sizeof(int32_t) * (static_cast<size_t>(n) + 2)) != hipSuccess)
    return acc_set_error(ACC_ENOMEM, "hipMalloc of the vertex metadata failed");
  uint32_t *d_max = reinterpret_cast<uint32_t *>(lo + n);
  uint32_t h_max[2] = {0, 0};
  g->h_off.resize(static_cast<size_t>(n) + 1);
  g->h_hs.resize(n);
  g->h_lo.resize(n);
  int rc = accfwd::launch_vertex_meta(g, lo, d_max);
  hipStream_t st = g->ctx->stream;
  if (!rc) {
    if (host_off)
      std::copy(host_off, host_off + static_cast<size_t>(n) + 1, g->h_off.begin());
    else
      rc = staged_copy(g->ctx, g->h_off.data(), g->d_off, sizeof(uint32_t) * (static_cast<size_t>(n) + 1), false);
  }
  if (!rc) rc = staged_copy(g->ctx, g->h_hs.data(), g->d_hs, sizeof(uint32_t) * n, false);
  if (!rc) rc = staged_copy(g->ctx, g->h_lo.data(), lo, sizeof(int32_t) * n, false);
  if (!rc) rc = staged_copy(g->ctx, h_max, d_max, sizeof(h_max), false);
  (void)hipStreamSynchronize(st);
  (void)hipFree(lo);
  if (rc) return rc;
  g->max_low = h_max[0];
  g->max_deg = h_max[1];
  return ACC_OK;
}

}  // namespace

int record_event(acc_graph *g, const char *name) {
  if (g->events.empty()) g->timing_valid = false;
  const int rc = record(g, true, name);
  if (!rc && g->events.size() > 1) g->timing_valid = true;
  return rc;
}

int acc_set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

extern "C" {

const char *acc_last_error(void) { return g_err.c_str(); }

int acc_device_count(int *count) {
  if (!count) return acc_set_error(ACC_EINVAL, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return ACC_OK;
}

int acc_ctx_create(int device, acc_ctx **out) {
  if (!out) return acc_set_error(ACC_EINVAL, "null out");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return acc_set_error(ACC_EDEVICE, "no HIP device is visible");
  if (device < 0 || device >= count) return acc_set_error(ACC_EINVAL, "device index out of range");
  ACC_HIP_TRY(hipSetDevice(device));
  auto *ctx = new (std::nothrow) acc_ctx;
  if (!ctx) return acc_set_error(ACC_ENOMEM, "host allocation failed");
  ctx->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete ctx;
    return acc_set_error(ACC_EDEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  }
  ctx->num_cus = prop.multiProcessorCount;
  ctx->lds_per_block = prop.sharedMemPerBlock;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return acc_set_error(ACC_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  const int rc = accnoalloc::bind_scratch(device, &ctx->na_scratch);
  if (rc) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return ACC_OK;
}

void acc_ctx_destroy(acc_ctx *ctx) {
  if (!ctx) return;
  DeviceGuard guard(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  for (CopyLane &l : ctx->lanes) {
    if (l.stream) {
      (void)hipStreamSynchronize(l.stream);
      (void)hipStreamDestroy(l.stream);
    }
    for (int b = 0; b < 2; b++) {
      if (l.buf[b]) (void)hipHostFree(l.buf[b]);
      if (l.ev[b]) (void)hipEventDestroy(l.ev[b]);
    }
  }
  delete ctx;
}

int acc_ctx_synchronize(acc_ctx *ctx) {
  if (!ctx) return acc_set_error(ACC_EINVAL, "null context");
  DeviceGuard guard(ctx->device);
  ACC_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ACC_OK;
}

void *acc_ctx_stream(acc_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int acc_graph_upload(acc_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst,
                     uint32_t flags, acc_graph **out) {
  if (!ctx || !out || !node_off || (m && !edge_dst)) return acc_set_error(ACC_EINVAL, "null argument");
  *out = nullptr;
  if (flags & ~(ACC_UPLOAD_NO_VALIDATE | ACC_UPLOAD_RELABEL))
    return acc_set_error(ACC_EINVAL, "unknown upload flags " + std::to_string(flags));
  if (m >= (1ULL << 32)) return acc_set_error(ACC_EINVAL, "m must be below 2^32 (uint32 node_off)");
  if (n >= (1u << 31)) return acc_set_error(ACC_EINVAL, "n must be below 2^31 (int32 ids)");
  if (node_off[0] != 0 || node_off[n] != m) return acc_set_error(ACC_EFORMAT, "node_off[0] != 0 or node_off[n] != m");
  for (uint32_t u = 0; u < n; u++)
    if (node_off[u + 1] < node_off[u])
      return acc_set_error(ACC_EFORMAT, "node_off not monotone at vertex " + std::to_string(u));
  DeviceGuard guard(ctx->device);
  auto *g = new (std::nothrow) acc_graph;
  if (!g) return acc_set_error(ACC_ENOMEM, "host allocation failed");
  g->ctx = ctx;
  g->n = n;
  g->m = m;
  g->shard_hi = n;
  int rc = ACC_OK;
  auto bail = [&](int code) {
    free_graph(g);
    return code;
  };
  if ((rc = dev_alloc(g, &g->d_off, static_cast<uint64_t>(n) + 1))) return bail(rc);
  if ((rc = dev_alloc(g, &g->d_adj, m + 64))) return bail(rc);
  if ((rc = dev_alloc(g, &g->d_cnt, m + 4))) return bail(rc);
  if ((rc = dev_alloc(g, &g->d_rev, m + 4))) return bail(rc);
  if ((rc = dev_alloc(g, &g->d_sum, 2))) return bail(rc);
  if ((rc = dev_alloc(g, &g->d_flags, 4))) return bail(rc);
  if ((rc = staged_copy(ctx, g->d_off, node_off, sizeof(uint32_t) * (static_cast<size_t>(n) + 1), true)))
    return bail(rc);
  if ((rc = staged_copy(ctx, g->d_adj, edge_dst, sizeof(int32_t) * m, true))) return bail(rc);
  hipError_t e = hipMemsetAsync(g->d_flags, 0, sizeof(int32_t) * 4, ctx->stream);
  if (e != hipSuccess) return bail(acc_set_error(ACC_EDEVICE, std::string("upload: ") + hipGetErrorString(e)));
  const bool relabel = (flags & ACC_UPLOAD_RELABEL) != 0;
  const bool validate = !(flags & ACC_UPLOAD_NO_VALIDATE) || relabel;  // the relabel indexes by neighbor id
  int32_t hf[4] = {0, 0, 0, 0};
  auto read_flags = [&]() {
    hipError_t err = hipMemcpyAsync(hf, g->d_flags, sizeof(hf), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    return err == hipSuccess ? ACC_OK
                             : acc_set_error(ACC_EDEVICE, std::string("validate: ") + hipGetErrorString(err));
  };
  if ((rc = accprep::slot_sources(g))) return bail(rc);
  if (validate) {
    if ((rc = accfwd::launch_validate(g)) || (rc = read_flags())) return bail(rc);
    if (hf[0] & 1) return bail(acc_set_error(ACC_EFORMAT, "neighbor id out of range"));
    if (hf[0] & 2) return bail(acc_set_error(ACC_EFORMAT, "Self loop"));
    if (hf[0] & 4) return bail(acc_set_error(ACC_EFORMAT, "Edges not sorted in increasing id order"));
  }
  if (relabel && ((rc = accprep::relabel_degree(g)) || (rc = accprep::slot_sources(g)))) return bail(rc);
  if ((rc = load_vertex_meta(g, relabel ? nullptr : node_off))) return bail(rc);
  uint64_t upper = 0;
  if ((rc = accprep::reverse_slots(g, validate, &upper))) return bail(rc);
  if (validate) {
    if ((rc = read_flags())) return bail(rc);
    if ((hf[0] & 8) || 2 * upper != m) return bail(acc_set_error(ACC_EFORMAT, "CSR not symmetric"));
  }
  if ((rc = upload_plans(g))) return bail(rc);
  *out = g;
  return ACC_OK;
}

void acc_graph_release(acc_graph *g) { free_graph(g); }

int acc_graph_info(const acc_graph *g, uint32_t *n, uint64_t *m, uint64_t *device_bytes) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (device_bytes) *device_bytes = g->device_bytes;
  return ACC_OK;
}

int acc_graph_download(const acc_graph *g, uint32_t *node_off, int32_t *edge_dst, int32_t *old_id) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  DeviceGuard guard(g->ctx->device);
  hipStream_t st = g->ctx->stream;
  if (node_off) std::copy(g->h_off.begin(), g->h_off.end(), node_off);
  if (edge_dst && g->m) ACC_TRY(staged_copy(g->ctx, edge_dst, g->d_adj, sizeof(int32_t) * g->m, false));
  if (old_id && g->d_old)
    ACC_HIP_TRY(hipMemcpyAsync(old_id, g->d_old, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, st));
  else if (old_id)
    for (uint32_t v = 0; v < g->n; v++) old_id[v] = static_cast<int32_t>(v);
  ACC_HIP_TRY(hipStreamSynchronize(st));
  return ACC_OK;
}

int acc_cn_compute(acc_graph *g, int path, uint32_t flags) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (path == ACC_PATH_AUTO) path = g->auto_path;
  if (path != ACC_PATH_FORWARD && path != ACC_PATH_BITMAP && path != ACC_PATH_MERGE)
    return acc_set_error(ACC_EINVAL, "unknown counting path " + std::to_string(path));
  DeviceGuard guard(g->ctx->device);
  const bool timing = (flags & ACC_CN_TIMING) != 0;
  const int plus = (flags & ACC_CN_RAW) ? 0 : 2;
  for (hipEvent_t e : g->events) (void)hipEventDestroy(e);
  g->events.clear();
  g->event_names.clear();
  g->timing_valid = false;
  if (path == ACC_PATH_BITMAP || path == ACC_PATH_MERGE) ACC_TRY(ensure_edge_plans(g));
  if (path == ACC_PATH_BITMAP) ACC_TRY(ensure_bitmap_pool(g));
  ACC_TRY(record(g, timing, "begin"));
  if (path == ACC_PATH_BITMAP || path == ACC_PATH_MERGE) {
    if (path == ACC_PATH_BITMAP) {
      ACC_TRY(accpaths::launch_bitmap(g, plus));
      ACC_TRY(record(g, timing, "k_bmp"));
    } else {
      ACC_TRY(accpaths::launch_merge(g, plus));
      ACC_TRY(record(g, timing, "k_merge"));
    }
    if (g->num_parts) {
      ACC_TRY(accfwd::launch_graph_totals(g, plus));
      ACC_TRY(record(g, timing, "k_graph_totals"));
    }
    g->totals_valid = g->num_parts != 0;
    g->have_result = true;
    g->counts_share = false;  // the per-edge paths always count the whole graph
    g->cnt_plus2 = plus == 2;
    g->timing_valid = timing;
    return ACC_OK;
  }
  ACC_TRY(accfwd::launch_init(g));
  ACC_TRY(record(g, timing, "k_init"));
  ACC_TRY(accfwd::launch_count(g));
  ACC_TRY(record(g, timing, ("k_fwd_count<" + std::to_string(g->fwd_unroll) + ">").c_str()));
  if (g->n_fwd_big) {
    ACC_TRY(accfwd::launch_count_big(g));
    ACC_TRY(record(g, timing, "k_fwd_big"));
  }
  // for a partitioned graph the combine also reduces the per-graph totals
  ACC_TRY(accfwd::launch_combine(g, (flags & ACC_CN_RAW) ? 0 : 2));
  ACC_TRY(record(g, timing, "k_combine"));
  g->totals_valid = g->num_parts != 0;
  g->have_result = true;
  g->counts_share = g->shard_lo != 0 || g->shard_hi != g->n;
  g->cnt_plus2 = (flags & ACC_CN_RAW) == 0;
  g->timing_valid = timing;
  return ACC_OK;
}

int acc_cn_download(acc_graph *g, int32_t *cnt_host, uint64_t *triangles) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (!g->have_result) return acc_set_error(ACC_EINVAL, "no counts computed yet");
  DeviceGuard guard(g->ctx->device);
  if (cnt_host && g->m) ACC_TRY(staged_copy(g->ctx, cnt_host, g->d_cnt, sizeof(int32_t) * g->m, false));
  unsigned long long s = 0;
  ACC_HIP_TRY(hipMemcpyAsync(&s, g->d_sum, sizeof(s), hipMemcpyDeviceToHost, g->ctx->stream));
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  if (triangles) *triangles = s / 3;
  return ACC_OK;
}

int acc_cn_all_edges(acc_graph *g, int path, uint32_t flags, int32_t *cnt_host, uint64_t *triangles) {
  ACC_TRY(acc_cn_compute(g, path, flags));
  return acc_cn_download(g, cnt_host, triangles);
}

const int32_t *acc_cn_device_counts(const acc_graph *g) { return g ? g->d_cnt : nullptr; }

int acc_cn_timings(acc_graph *g, float *ms, int cap, char *names, int names_cap) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (!g->timing_valid) return 0;
  DeviceGuard guard(g->ctx->device);
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  std::string joined;
  int k = 0;
  for (size_t i = 1; i < g->events.size(); i++, k++) {
    float t = 0;
    ACC_HIP_TRY(hipEventElapsedTime(&t, g->events[i - 1], g->events[i]));
    if (ms && k < cap) ms[k] = t;
    if (!joined.empty()) joined += ";";
    joined += g->event_names[i];
  }
  if (names && names_cap > 0) {
    std::strncpy(names, joined.c_str(), static_cast<size_t>(names_cap) - 1);
    names[names_cap - 1] = '\0';
  }
  return k;
}

int acc_graph_set_param(acc_graph *g, int param, int64_t value) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  DeviceGuard guard(g->ctx->device);
  drop_capture(g);
  switch (param) {
    case ACC_PARAM_FWD_UNROLL:
      if (value != 1 && value != 2 && value != 4) return acc_set_error(ACC_EINVAL, "fwd_unroll must be 1, 2 or 4");
      g->fwd_unroll = static_cast<int>(value);
      return ACC_OK;
    case ACC_PARAM_WAVE_SLOTS:
      if (value < 0 || value > static_cast<int64_t>(accfwd::kWaveSlots))
        return acc_set_error(ACC_EINVAL, "wave_slots must lie in [0, 64]");
      g->wave_slots = static_cast<uint32_t>(value);
      return upload_plans(g);
    case ACC_PARAM_WAVE_LOW:
      if (value < 0 || value > static_cast<int64_t>(accfwd::kWaveCap))
        return acc_set_error(ACC_EINVAL, "wave_low must lie in [0, 256]");
      g->wave_low = static_cast<uint32_t>(value);
      return upload_plans(g);
    case ACC_PARAM_CHUNK:
      if (value < 64 || value > static_cast<int64_t>(accfwd::kChunk))
        return acc_set_error(ACC_EINVAL, "chunk must lie in [64, 512]");
      g->chunk = static_cast<uint32_t>(value);
      g->h_work.clear();  // the shard work estimate counts units of `chunk` slots
      return upload_plans(g);
    case ACC_PARAM_PERSIST:
      if (value < 0 || value > 256) return acc_set_error(ACC_EINVAL, "persist must lie in [0, 256]");
      g->persist = static_cast<uint32_t>(value);
      return ACC_OK;
    case ACC_PARAM_AUTO_PATH:
      if (value != ACC_PATH_FORWARD && value != ACC_PATH_BITMAP && value != ACC_PATH_MERGE)
        return acc_set_error(ACC_EINVAL, "auto_path must be a counting path");
      g->auto_path = static_cast<int>(value);
      return ACC_OK;
    case ACC_PARAM_BMP_HASH_MAX:
      if (value < 0 || value > static_cast<int64_t>(kEdgeHashMax))
        return acc_set_error(ACC_EINVAL, "bmp_hash_max must lie in [0, 16384]");
      if (static_cast<uint32_t>(value) != g->bmp_hash_max) g->edge.valid = false;
      g->bmp_hash_max = static_cast<uint32_t>(value);
      return ACC_OK;
    case ACC_PARAM_BMP_BATCH:
      if (value < 1 || value > static_cast<int64_t>(kHubBatch))
        return acc_set_error(ACC_EINVAL, "bmp_batch must lie in [1, 16384]");
      if (static_cast<uint32_t>(value) != g->bmp_batch) g->edge.valid = false;
      g->bmp_batch = static_cast<uint32_t>(value);
      return ACC_OK;
    default:
      return acc_set_error(ACC_EINVAL, "unknown parameter " + std::to_string(param));
  }
}

int acc_graph_get_param(const acc_graph *g, int param, int64_t *value) {
  if (!g || !value) return acc_set_error(ACC_EINVAL, "null argument");
  switch (param) {
    case ACC_PARAM_FWD_UNROLL: *value = g->fwd_unroll; return ACC_OK;
    case ACC_PARAM_WAVE_SLOTS: *value = g->wave_slots; return ACC_OK;
    case ACC_PARAM_WAVE_LOW: *value = g->wave_low; return ACC_OK;
    case ACC_PARAM_CHUNK: *value = g->chunk; return ACC_OK;
    case ACC_PARAM_AUTO_PATH: *value = g->auto_path; return ACC_OK;
    case ACC_PARAM_PERSIST: *value = g->persist; return ACC_OK;
    case ACC_PARAM_BMP_HASH_MAX: *value = g->bmp_hash_max; return ACC_OK;
    case ACC_PARAM_BMP_BATCH: *value = g->bmp_batch; return ACC_OK;
    default: return acc_set_error(ACC_EINVAL, "unknown parameter " + std::to_string(param));
  }
}

namespace {
// Device time (ms) of one counting pass with the current parameters: the
// median of `reps` passes after one warm-up pass.
int time_pass(acc_graph *g, int path, int reps, float *ms) {
  std::vector<float> t;
  for (int r = 0; r <= reps; r++) {
    ACC_TRY(acc_cn_compute(g, path, ACC_CN_TIMING));
    float k[32];
    const int nk = acc_cn_timings(g, k, 32, nullptr, 0);
    if (nk < 0) return nk;
    float tot = 0;
    for (int i = 0; i < nk && i < 32; i++) tot += k[i];
    if (r > 0) t.push_back(tot);
  }
  std::sort(t.begin(), t.end());
  *ms = t[t.size() / 2];
  return ACC_OK;
}
}  // namespace

int acc_tune(acc_graph *g, uint32_t flags, char *report, int report_cap) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  DeviceGuard guard(g->ctx->device);
  const bool quick = (flags & ACC_TUNE_QUICK) != 0;
  const int reps = quick ? 1 : 2;
  // Coordinate descent over the forward path's parameters, one at a time in
  // this order, each trial timed on the device; then the per-edge paths at
  // the chosen chunk size.
  struct Axis {
    int param;
    const char *name;
    std::vector<int64_t> values;
  };
  const std::vector<Axis> axes = {
      {ACC_PARAM_FWD_UNROLL, "fwd_unroll", quick ? std::vector<int64_t>{2, 4} : std::vector<int64_t>{1, 2, 4}},
      {ACC_PARAM_WAVE_SLOTS, "wave_slots", quick ? std::vector<int64_t>{16, 64} : std::vector<int64_t>{0, 16, 32, 64}},
      {ACC_PARAM_WAVE_LOW, "wave_low", quick ? std::vector<int64_t>{256} : std::vector<int64_t>{64, 128, 256}},
      {ACC_PARAM_CHUNK, "chunk", quick ? std::vector<int64_t>{256, 512} : std::vector<int64_t>{128, 256, 512}},
      {ACC_PARAM_PERSIST, "persist", quick ? std::vector<int64_t>{0, 24, 256} : std::vector<int64_t>{0, 8, 24, 64, 128, 256}},
  };
  std::string trials;
  auto add_trial = [&](const std::string &path, const char *param, int64_t value, float ms) {
    trials += (trials.empty() ? "" : ", ") + std::string("[\"") + path + "\", \"" + param + "\", " +
              std::to_string(value) + ", " + std::to_string(ms) + "]";
  };
  float best_ms = 0;
  ACC_TRY(time_pass(g, ACC_PATH_FORWARD, reps, &best_ms));
  add_trial("forward", "initial", 0, best_ms);
  for (const Axis &ax : axes) {
    int64_t keep = 0;
    ACC_TRY(acc_graph_get_param(g, ax.param, &keep));
    for (int64_t v : ax.values) {
      if (v == keep) continue;
      ACC_TRY(acc_graph_set_param(g, ax.param, v));
      float ms = 0;
      ACC_TRY(time_pass(g, ACC_PATH_FORWARD, reps, &ms));
      add_trial("forward", ax.name, v, ms);
      if (ms < best_ms) {
        best_ms = ms;
        keep = v;
      }
    }
    ACC_TRY(acc_graph_set_param(g, ax.param, keep));
  }
  int best_path = ACC_PATH_FORWARD;
  if (!(flags & ACC_TUNE_FORWARD_ONLY)) {
    // the bitmap path's hash/bitmap boundary, then both per-edge paths against the forward path
    float bmp_ms = 0;
    int64_t keep = g->bmp_hash_max;
    for (int64_t v : quick ? std::vector<int64_t>{256} : std::vector<int64_t>{256, 2048, kEdgeHashMax}) {
      ACC_TRY(acc_graph_set_param(g, ACC_PARAM_BMP_HASH_MAX, v));
      float ms = 0;
      ACC_TRY(time_pass(g, ACC_PATH_BITMAP, 1, &ms));
      add_trial("bitmap", "bmp_hash_max", v, ms);
      if (bmp_ms == 0 || ms < bmp_ms) {
        bmp_ms = ms;
        keep = v;
      }
    }
    ACC_TRY(acc_graph_set_param(g, ACC_PARAM_BMP_HASH_MAX, keep));
    if (bmp_ms < best_ms) {
      best_ms = bmp_ms;
      best_path = ACC_PATH_BITMAP;
    }
    float ms = 0;
    ACC_TRY(time_pass(g, ACC_PATH_MERGE, 1, &ms));
    add_trial("merge", "path", ACC_PATH_MERGE, ms);
    if (ms < best_ms) {
      best_ms = ms;
      best_path = ACC_PATH_MERGE;
    }
  }
  ACC_TRY(acc_graph_set_param(g, ACC_PARAM_AUTO_PATH, best_path));
  if (report && report_cap > 0) {
    std::string r = std::string("{\"path\": \"") +
                    (best_path == ACC_PATH_FORWARD ? "forward" : best_path == ACC_PATH_BITMAP ? "bitmap" : "merge") +
                    "\"";
    for (const Axis &ax : axes) {
      int64_t v = 0;
      ACC_TRY(acc_graph_get_param(g, ax.param, &v));
      r += ", \"" + std::string(ax.name) + "\": " + std::to_string(v);
    }
    r += ", \"ms\": " + std::to_string(best_ms) + ", \"trials\": [" + trials + "]}";
    std::strncpy(report, r.c_str(), static_cast<size_t>(report_cap) - 1);
    report[report_cap - 1] = '\0';
  }
  return ACC_OK;
}

int acc_graph_set_shard(acc_graph *g, uint32_t parts, uint32_t index, uint32_t *range_out) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (parts == 0 || index >= parts) return acc_set_error(ACC_EINVAL, "need 0 <= index < parts");
  DeviceGuard guard(g->ctx->device);
  const uint32_t n = g->n;
  uint32_t lo = 0, hi = n;
  if (parts > 1 && n) {
    if (g->h_work.empty()) {
      // one whole-graph counting pass without the combine leaves in every
      // upper slot (b, c) its probe's hits, which the work estimate weighs
      if (g->shard_lo != 0 || g->shard_hi != n) {
        g->shard_lo = 0;
        g->shard_hi = n;
        ACC_TRY(upload_plans(g));
      }
      g->have_result = false;
      g->totals_valid = false;
      ACC_TRY(accfwd::launch_init(g));
      ACC_TRY(accfwd::launch_count(g));
      if (g->n_fwd_big) ACC_TRY(accfwd::launch_count_big(g));
      unsigned long long *d_work = nullptr;
      ACC_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&d_work), sizeof(unsigned long long) * n));
      std::vector<unsigned long long> w(n);
      int rc = accfwd::launch_source_work(g, d_work);
      hipError_t e = hipSuccess;
      if (!rc) e = hipMemcpyAsync(w.data(), d_work, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost,
                                  g->ctx->stream);
      if (!rc && e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
      (void)hipStreamSynchronize(g->ctx->stream);
      (void)hipFree(d_work);
      if (rc) return rc;
      if (e != hipSuccess) return acc_set_error(ACC_EDEVICE, std::string("source work: ") + hipGetErrorString(e));
      g->h_work.assign(static_cast<size_t>(n) + 1, 0);
      for (uint32_t b = 0; b < n; b++) g->h_work[b + 1] = g->h_work[b] + w[b];
    }
    // share s = the sources whose work prefix lies in [s W / parts, (s + 1) W / parts)
    const uint64_t total = g->h_work[n];
    auto cut = [&](uint32_t s) -> uint32_t {
      if (s == 0) return 0;
      if (s == parts) return n;
      const unsigned __int128 t = static_cast<unsigned __int128>(total) * s / parts;
      const uint64_t target = static_cast<uint64_t>(t);
      return static_cast<uint32_t>(std::lower_bound(g->h_work.begin(), g->h_work.end(), target) - g->h_work.begin());
    };
    lo = std::min(cut(index), n);
    hi = std::max(lo, std::min(cut(index + 1), n));
  }
  if (range_out) {
    range_out[0] = lo;
    range_out[1] = hi;
  }
  if (lo == g->shard_lo && hi == g->shard_hi) return ACC_OK;
  g->shard_lo = lo;
  g->shard_hi = hi;
  return upload_plans(g);
}

int acc_graph_set_partition(acc_graph *g, uint32_t num_graphs, const uint32_t *vertex_off) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  DeviceGuard guard(g->ctx->device);
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  drop_capture(g);
  dev_free(g, &g->d_gvoff, g->num_parts + 1ULL);
  dev_free(g, &g->d_gsoff, g->num_parts + 1ULL);
  dev_free(g, &g->d_gtot, g->num_parts);
  g->num_parts = 0;
  g->totals_valid = false;
  if (num_graphs == 0) return ACC_OK;
  if (!vertex_off) return acc_set_error(ACC_EINVAL, "null vertex offsets");
  if (vertex_off[0] != 0 || vertex_off[num_graphs] != g->n)
    return acc_set_error(ACC_EINVAL, "vertex offsets must start at 0 and end at n");
  for (uint32_t k = 0; k < num_graphs; k++)
    if (vertex_off[k + 1] < vertex_off[k]) return acc_set_error(ACC_EINVAL, "vertex offsets must not decrease");
  std::vector<uint32_t> gsoff(num_graphs + 1ULL);
  for (uint32_t k = 0; k <= num_graphs; k++) gsoff[k] = g->h_off[vertex_off[k]];
  // No edge may join two graphs.  In a symmetric CSR an edge (u, v) with u in
  // graph k and v in a later graph puts u below v's range, so checking every
  // vertex's smallest neighbor against its own range's start suffices.
  for (uint32_t k = 0; k < num_graphs; k++)
    for (uint32_t v = vertex_off[k]; v < vertex_off[k + 1]; v++)
      if (static_cast<uint32_t>(g->h_lo[v]) < vertex_off[k])
        return acc_set_error(ACC_EINVAL, "an edge joins two graphs of the partition");
  // on any failure the partition arrays are freed again and none is set
  auto fail = [&](int rc) {
    (void)hipStreamSynchronize(g->ctx->stream);
    dev_free(g, &g->d_gvoff, num_graphs + 1ULL);
    dev_free(g, &g->d_gsoff, num_graphs + 1ULL);
    dev_free(g, &g->d_gtot, num_graphs);
    return rc;
  };
  int rc = dev_alloc(g, &g->d_gvoff, static_cast<uint64_t>(num_graphs) + 1);
  if (!rc) rc = dev_alloc(g, &g->d_gsoff, static_cast<uint64_t>(num_graphs) + 1);
  if (!rc) rc = dev_alloc(g, &g->d_gtot, num_graphs);
  if (rc) return fail(rc);
  hipError_t e = hipMemcpyAsync(g->d_gvoff, vertex_off, sizeof(uint32_t) * (num_graphs + 1ULL),
                                hipMemcpyHostToDevice, g->ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(g->d_gsoff, gsoff.data(), sizeof(uint32_t) * (num_graphs + 1ULL), hipMemcpyHostToDevice,
                       g->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
  if (e != hipSuccess) return fail(acc_set_error(ACC_EDEVICE, std::string("partition: ") + hipGetErrorString(e)));
  g->num_parts = num_graphs;
  return upload_plans(g);  // a packed batch keeps its workgroup units in source order
}

int acc_cn_graph_totals(acc_graph *g, uint64_t *totals) {
  if (!g || !totals) return acc_set_error(ACC_EINVAL, "null argument");
  if (!g->num_parts) return acc_set_error(ACC_EINVAL, "no partition set");
  if (!g->totals_valid) return acc_set_error(ACC_EINVAL, "no counting pass since the partition was set");
  DeviceGuard guard(g->ctx->device);
  std::vector<unsigned long long> t(g->num_parts);
  ACC_HIP_TRY(hipMemcpyAsync(t.data(), g->d_gtot, sizeof(unsigned long long) * g->num_parts, hipMemcpyDeviceToHost,
                             g->ctx->stream));
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  for (uint32_t k = 0; k < g->num_parts; k++) totals[k] = t[k] / 6;
  return ACC_OK;
}

int acc_cn_capture(acc_graph *g, int path, uint32_t flags) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (flags & ACC_CN_TIMING) return acc_set_error(ACC_EINVAL, "timing events cannot be captured");
  DeviceGuard guard(g->ctx->device);
  if (path == ACC_PATH_AUTO) path = g->auto_path;
  // plans and the bitmap pool are allocated and cleared outside the capture
  if (path == ACC_PATH_BITMAP || path == ACC_PATH_MERGE) ACC_TRY(ensure_edge_plans(g));
  if (path == ACC_PATH_BITMAP) ACC_TRY(ensure_bitmap_pool(g));
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  drop_capture(g);
  ACC_HIP_TRY(hipStreamBeginCapture(g->ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = acc_cn_compute(g, path, flags);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(g->ctx->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return acc_set_error(ACC_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  g->graph = graph;
  ACC_HIP_TRY(hipGraphInstantiate(&g->graph_exec, graph, nullptr, nullptr, 0));
  g->graph_plus2 = !(flags & ACC_CN_RAW);
  g->graph_share = g->counts_share;
  g->have_result = false;
  g->totals_valid = false;
  return ACC_OK;
}

int acc_cn_replay(acc_graph *g) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  if (!g->graph_exec) return acc_set_error(ACC_EINVAL, "no captured counting pass");
  DeviceGuard guard(g->ctx->device);
  ACC_HIP_TRY(hipGraphLaunch(g->graph_exec, g->ctx->stream));
  g->cnt_plus2 = g->graph_plus2;
  g->counts_share = g->graph_share;
  g->timing_valid = false;
  g->have_result = true;
  g->totals_valid = g->num_parts != 0;
  return ACC_OK;
}

int acc_triangle_total(acc_graph *g, int path, uint64_t *triangles) {
  if (!g || !triangles) return acc_set_error(ACC_EINVAL, "null argument");
  if (path == ACC_PATH_AUTO) path = g->auto_path;
  if (path != ACC_PATH_FORWARD) {
    ACC_TRY(acc_cn_compute(g, path, 0));
    return acc_cn_download(g, nullptr, triangles);
  }
  // forward lists without per-slot outputs: every triangle counted once at its middle vertex
  DeviceGuard guard(g->ctx->device);
  ACC_TRY(accfwd::launch_triangles(g));
  unsigned long long t = 0;
  ACC_HIP_TRY(hipMemcpyAsync(&t, g->d_sum + 1, sizeof(t), hipMemcpyDeviceToHost, g->ctx->stream));
  ACC_HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  *triangles = t;
  return ACC_OK;
}

int acc_triangle_launch(acc_graph *g) {
  if (!g) return acc_set_error(ACC_EINVAL, "null graph");
  DeviceGuard guard(g->ctx->device);
  return accfwd::launch_triangles(g);
}

}  // extern "C"
