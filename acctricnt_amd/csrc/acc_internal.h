// Internal declarations shared by the HIP translation units of libacc_cn.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "acc_cn.h"
#include "fwd_plan.h"

// Sets the thread's last error and returns code.
int acc_set_error(int code, const std::string &msg);

struct acc_graph;
// Records a named HIP event on the graph's stream; acc_cn_timings reports the
// intervals between consecutive events under the later event's name.
int record_event(acc_graph *g, const char *name);

#define ACC_HIP_TRY(expr)                                                                   \
  do {                                                                                      \
    hipError_t acc_e_ = (expr);                                                             \
    if (acc_e_ != hipSuccess)                                                               \
      return acc_set_error(ACC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(acc_e_)); \
  } while (0)

#define ACC_TRY(expr)        \
  do {                       \
    int acc_rc_ = (expr);    \
    if (acc_rc_) return acc_rc_; \
  } while (0)

// One lane of the staged host copies: a stream and two pinned buffers.
struct CopyLane {
  hipStream_t stream = nullptr;
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

struct acc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  size_t lds_per_block = 0;
  std::mutex copy_mu;             // the copy lanes serve one transfer at a time
  std::vector<CopyLane> lanes;    // allocated at the first large pageable transfer
  unsigned long long *na_scratch = nullptr;  // device address of the allocation-free entry point's scratch
};

namespace accnoalloc {
// Device address of acc_noalloc.hip's file-scope scratch array on `device`
// (the current device), looked up once per device.
int bind_scratch(int device, unsigned long long **addr);
}  // namespace accnoalloc



// Per-edge plans (bitmap and merge paths): chunks of at most kChunk upper
// slots of every source, by d_u.  Wave units: a source with d_u <= 256 and at
// most 64 upper slots, one wave each, N(u) in a wave-private LDS table or
// sorted run.  Block units: the bitmap path keeps N(u) in an LDS hash table
// of 2^12, 2^14 or 2^15 keys (d_u up to half of that) or, for hubs, in an HBM
// bitmap; the merge path stages N(u) in LDS windows of 2^11 ids (the first
// class) or 2^13 ids (the others).
enum EdgeClass { kEdgeWave = 0, kEdgeHash12, kEdgeHash14, kEdgeHash15, kEdgeHub, kEdgeClasses };
constexpr uint32_t kEdgeWaveDeg = 256;  // d_u of a wave unit
constexpr int kEdgeWaveLog = 8;         // log2 kEdgeWaveDeg
static_assert((1u << kEdgeWaveLog) == kEdgeWaveDeg, "wave-unit runs are searched in kEdgeWaveLog steps");
constexpr uint32_t kEdgeWaveUp = 64;    // upper slots of a wave unit
struct EdgePlan {
  FwdUnit *d_units[kEdgeClasses] = {};  // hub units: nlow = the hub's bitmap in its batch's pool
  uint32_t n_units[kEdgeClasses] = {};
  uint32_t *d_hubs = nullptr;           // hub sources in order
  uint32_t n_hubs = 0;
  std::vector<uint32_t> hub_batch;       // batch k: hubs [hub_batch[k], hub_batch[k+1])
  std::vector<uint32_t> hub_unit_batch;  // ... and hub units [hub_unit_batch[k], hub_unit_batch[k+1])
  uint32_t *d_pool = nullptr;            // bitmaps of one batch of hubs, zero between passes
  uint32_t pool_hubs = 0;
  uint64_t pool_words = 0;               // words per bitmap: ceil(n / 32)
  int idx_shift = 0;                     // LDS index of the hash units: one bit per 2^idx_shift bitmap words
  int hub_idx_shift = 0;                 // LDS index of the hub units: one bit per 16 << hub_idx_shift ids
  bool valid = false;
};
constexpr uint32_t kEdgeHashMax = 1u << 14;  // largest d_u an LDS hash table holds

struct acc_graph {
  acc_ctx *ctx = nullptr;
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t max_deg = 0;
  uint32_t max_low = 0;  // max |N<(b)|
  // resident CSR and outputs
  uint32_t *d_off = nullptr;
  int32_t *d_adj = nullptr;
  int32_t *d_cnt = nullptr;
  uint32_t *d_hs = nullptr;             // first upper slot of every vertex
  uint32_t *d_rev = nullptr;            // reverse slot of every upper slot, ~0 elsewhere (built at upload)
  unsigned long long *d_sum = nullptr;  // [0]: sum over upper slots of the raw counts (= 3T); [1]: T alone
  int32_t *d_flags = nullptr;           // validation flags
  unsigned long long *d_tri_parts = nullptr;  // partial sums of the triangle-only pass
  int32_t *d_src = nullptr;             // source vertex of every slot (built for SCAN)
  int32_t *d_old = nullptr;             // caller vertex ids for cluster ids (NULL: identity)
  // similarity index (built for serving): every vertex's slots sorted by a
  // monotone float key of cn^2 / ((d_u+1)(d_v+1)), descending
  float *d_skey = nullptr;     // sorted keys
  int32_t *d_snbr = nullptr;   // neighbor of every sorted position
  uint32_t *d_sslot = nullptr; // original slot of every sorted position
  bool cnt_plus2 = true;                // counts stored as 2 + |N(u) ∩ N(v)|
  bool relabelled = false;              // relabelled by degree at upload (d_old holds the old ids)
  // forward plan, binned by degree skew
  FwdUnit *d_fwd_wave = nullptr;  // wave packs of small sources: {b0, b1, upper slots, lower entries}
  uint32_t n_fwd_wave = 0;
  FwdUnit *d_fwd_block = nullptr;  // one workgroup per chunk of upper slots, N<(b) in LDS
  uint32_t n_fwd_block = 0;
  FwdUnit *d_fwd_big = nullptr;  // |N<(b)| beyond the LDS table: global-memory lookups
  uint32_t n_fwd_big = 0;
  EdgePlan edge;  // per-edge plans, built at the first bitmap or merge pass
  uint64_t device_bytes = 0;
  // tuning parameters (acc_graph_set_param)
  int fwd_unroll = 4;
  int auto_path = ACC_PATH_FORWARD;  // what ACC_PATH_AUTO runs
  uint32_t wave_slots = accfwd::kWaveSlots;  // wave-unit bin: at most this many upper slots
  uint32_t wave_low = accfwd::kWaveCap;      // ... and |N<(b)| at most this
  uint32_t chunk = accfwd::kChunk;           // upper slots per workgroup unit
  uint32_t persist = 256;                    // counting grid: 0 one workgroup per task, else at most this many
                                             // workgroups per CU looping over the tasks
  uint32_t bmp_hash_max = 256;               // bitmap path: d_u above this goes to the HBM bitmaps
  uint32_t bmp_batch = 4096;                 // bitmap path: hub bitmaps built and probed per batch
  uint32_t shard_lo = 0, shard_hi = 0;       // forward sources [shard_lo, shard_hi) (the whole graph: 0, n)
  std::vector<uint64_t> h_work;              // prefix of the forward probe work per source (n + 1), for shards
  // host copies kept for plan rebuilds: offsets, first upper slot and smallest
  // neighbor (the vertex itself when it has no lower neighbor) of every vertex
  std::vector<uint32_t> h_off;
  std::vector<uint32_t> h_hs;
  std::vector<int32_t> h_lo;
  // multi-graph partition: graph k owns vertices [gvoff[k], gvoff[k+1])
  uint32_t num_parts = 0;
  uint32_t *d_gvoff = nullptr;
  uint32_t *d_gsoff = nullptr;           // slot offsets: graph k owns slots [gsoff[k], gsoff[k+1])
  unsigned long long *d_gtot = nullptr;  // per-graph sum of raw counts (= 6 T_k)
  bool totals_valid = false;             // d_gtot holds the last pass's totals
  // captured counting pass
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_plus2 = true;  // count format the captured pass writes
  // last results / timing
  bool have_result = false;
  bool counts_share = false;  // the counts are one share's (a forward pass restricted by acc_graph_set_shard)
  bool graph_share = false;   // the captured pass counts one share
  std::vector<hipEvent_t> events;
  std::vector<std::string> event_names;
  bool timing_valid = false;
};

// Per-edge path launchers (k_paths.hip).
namespace accpaths {
int launch_bitmap(acc_graph *g, int plus);
int launch_merge(acc_graph *g, int plus);
}  // namespace accpaths

// Builds the slot -> source array (acc_scan.hip).
int acc_graph_ensure_src(acc_graph *g);

constexpr int kScanCounters = 16;
constexpr uint32_t kItem = 256;  // list positions per work item of the index path

// One SCAN query as the batched SCAN kernels see it (acc_scan.hip): the
// resident graph's arrays, the query's buffers and its exact eps = p/q, mu.
struct ScanJob {
  const uint32_t *off;
  const int32_t *adj;
  const int32_t *src;
  const int32_t *cnt;
  const int32_t *old;  // caller vertex ids, or NULL
  uint8_t *sim;
  const float *skey;     // similarity index (NULL: per-slot kernels)
  const int32_t *snbr;
  const uint32_t *sslot;
  uint32_t *send;        // per vertex: end of the possibly-similar prefix
  int2 *work;            // work items (vertex, first list position) of the index path
  int32_t *core_cnt;
  int32_t *label;
  int32_t *parent;
  int32_t *rmin;
  int32_t *rmax;
  int32_t *cl;
  int32_t *cid;
  unsigned long long *counters;  // core, cluster, hub, outlier, pairs, index-path sums (5..7), item counts (8, 9)
  int8_t *out_label;             // host-mapped copies of the results (NULL: none)
  int32_t *out_cid;
  unsigned long long *out_counters;
  int2 *pairs;
  unsigned long long pairs_cap;
  unsigned long long p2, q2;
  uint64_t m;
  uint32_t n;
  int32_t add;  // added to cnt to get 2 + |N(u) ∩ N(v)|
  int32_t mu;
  uint32_t vb0, vbn;  // workgroups of vertex-sized launches
  uint32_t sb0, sbn;  // workgroups of slot-sized launches
};

// A launch's job table, passed by value as a kernel argument (no copy to
// device memory before the kernels): at most kMaxJobs jobs per launch.
constexpr int kMaxJobs = 14;
struct JobTable {
  ScanJob jobs[kMaxJobs];
  int32_t nj;
};
static_assert(sizeof(JobTable) <= 3800, "kernel arguments are limited to 4 KiB");

// Per-query device buffers, grown on demand.
struct ScanBuffers {
  uint8_t *sim = nullptr;
  uint32_t *send = nullptr;
  int2 *work = nullptr;
  int32_t *core_cnt = nullptr;
  int32_t *label = nullptr;
  int32_t *parent = nullptr;
  int32_t *rmin = nullptr;
  int32_t *rmax = nullptr;
  int32_t *cl = nullptr;
  int32_t *cid = nullptr;
  unsigned long long *counters = nullptr;
  uint32_t n_cap = 0;
  uint64_t m_cap = 0;
  uint64_t bytes = 0;
};

namespace accscan {
void free_buffers(ScanBuffers *b);
int reserve_buffers(ScanBuffers *b, uint32_t n, uint64_t m);
uint64_t buffer_bytes(uint32_t n, uint64_t m);
int check_eps(int64_t eps_num, int64_t eps_den);
ScanJob make_job(const acc_graph *g, const ScanBuffers &b, int64_t eps_num, int64_t eps_den, int32_t mu);
void plan_jobs(ScanJob *jobs, int nj, int cus, uint32_t *vblocks, uint32_t *sblocks);
// Enqueues the eight per-slot SCAN kernels over the jobs of jt (planned).
int launch_jobs(const JobTable &jt, uint32_t vblocks, uint32_t sblocks, hipStream_t s);
int launch_pairs(const JobTable &jt, uint32_t sblocks, hipStream_t s);
// Builds the similarity index of a resident graph (counts computed, slot
// sources built); enqueued on the graph's context stream.
int build_index(acc_graph *g);
// Whether the index decides eps = p/q exactly (p, q < 2^26) on graph g.
bool index_usable(const acc_graph *g, int64_t p, int64_t q);
// The six index-path SCAN kernels over nj jobs whose skey is set.
int launch_index_jobs(const JobTable &jt, uint32_t vblocks, uint32_t sblocks, hipStream_t s);
}  // namespace accscan

// Kernel launchers (k_forward.hip).
namespace accfwd {
int launch_init(acc_graph *g);
// Slot-parallel over the sources in d_cnt (accprep::slot_sources).  With
// validate, d_flags[0] gets flag 8 for an upper slot without reverse and
// d_flags[2..3] (a uint64) the number of upper slots.
int launch_build_rev(acc_graph *g, bool validate);
// Triangle total only (forward path, no per-slot counts) into d_sum[1].
int launch_triangles(acc_graph *g);
int launch_count(acc_graph *g);
int launch_count_big(acc_graph *g);
// Adds each upper slot and its reverse; with a partition, also the per-graph
// totals (which launch_init zeroes).
int launch_combine(acc_graph *g, int plus);
int launch_graph_totals(acc_graph *g, int plus);
// Slot-parallel over the sources in d_cnt (accprep::slot_sources).
int launch_validate(acc_graph *g);
// Per-vertex plan inputs into g->d_hs (first upper slot) and lo (min(N(v) ∪ {v}));
// maxes[0] = max |N<(v)|, maxes[1] = max degree.
int launch_vertex_meta(acc_graph *g, int32_t *lo, uint32_t *maxes);
// Forward work estimate of every source b into work[b]: the sum over its
// upper slots (b, c) of |N(c) ∩ [0, b)| + 4 + 2 hits, plus |N<(b)| + 256 per
// unit, hits read from d_cnt's upper slots after a pass without the combine.
int launch_source_work(acc_graph *g, unsigned long long *work);
}  // namespace accfwd

// Upload-time preparation (k_prepare.hip).  slot_sources writes the source
// vertex of every slot into d_cnt (d_rev is its scratch); both arrays are
// free until the reverse slots are built and the first pass runs.
// relabel_degree, given those sources, rewrites d_off and d_adj in place by
// degree (d_rev is its scratch) and keeps the old id of every new vertex in
// d_old; the sources must be recomputed after it.  reverse_slots, given the
// sources and d_hs, fills d_rev (d_cnt is its scratch) and returns the number
// of upper slots; with validate, an upper slot without its reverse sets flag
// 8 in d_flags, and a count other than m / 2 leaves d_rev unbuilt.
namespace accprep {
int slot_sources(acc_graph *g);
int relabel_degree(acc_graph *g);
int reverse_slots(acc_graph *g, bool validate, uint64_t *upper);
}  // namespace accprep
