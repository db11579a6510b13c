/*
 * acc_cn.h -- C ABI of libacc_cn.so, the MI355X (gfx950) all-edge
 * common-neighbor library.
 *
 * This is the drop-in boundary for the reference's GPU plugin
 *   void SCAN_XP::CheckCore(Graph *g)
 *     (scan-query-cuda/scan_xp.h:36, defined at
 *      scan-query-cuda/experimental_scan_xp_multigpu_multipass.cu:3-166)
 * and for the triangle entry
 *   void SCAN_XP::TriCnt()        (scan-query/scan_xp.h:250,
 *                                   scan-query/scan_xp_triangle_cnt.cpp:62-116)
 * The C++ host classes in acctricnt_amd/csrc/host/ keep those signatures and
 * call the functions below.  Plain pointers and sizes only; no exit() on
 * error: every function returns 0 (ACC_OK) or a negative ACC_E* code from
 * acc_graph.h, and acc_last_error() gives the text of the calling thread's
 * last error.
 *
 * Threading: one acc_ctx owns one HIP stream on one device.  Calls on one
 * context must not overlap; different contexts may be used from different
 * threads.  All device work of a context runs on its stream.
 */
#ifndef ACC_CN_H
#define ACC_CN_H

#include <stdint.h>

#include "acc_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct acc_ctx acc_ctx;
typedef struct acc_graph acc_graph;

/* Counting paths.  All return identical counts. */
#define ACC_PATH_AUTO 0    /* the tuned choice for the graph (see acc_tune) */
#define ACC_PATH_FORWARD 1 /* degree-ordered forward lists: per source b an LDS hash of N<(b),
                              probe N(c) ∩ [0,b) for every c > b; one count per triangle role */
#define ACC_PATH_BITMAP 2  /* per source u an N(u) membership set (LDS, HBM bitmap for hubs),
                              probe N(v) for every v > u: the reference BMP kernel's algorithm
                              (scan-query-cuda/scan_xp_multi_gpu_multi_pass_common.cu:308-407) */
#define ACC_PATH_MERGE 4   /* merge path over both sorted lists, LDS-staged windows:
                              the reference MPS kernels' algorithm (same file :32-176) */

/* Flags for acc_cn_compute / acc_cn_all_edges. */
#define ACC_CN_RAW 0x1u     /* store |N(u) ∩ N(v)| instead of the reference's 2 + |N(u) ∩ N(v)| */
#define ACC_CN_TIMING 0x2u  /* record per-kernel HIP events; read them with acc_cn_timings */

/* Upload flags. */
#define ACC_UPLOAD_NO_VALIDATE 0x1u /* skip the CSR validation (trusted input: ids in range,
                                       lists sorted without self-loops, symmetric) */
#define ACC_UPLOAD_RELABEL 0x2u     /* relabel on the device by degree, descending, ties by
                                       ascending id; implies validation.  See acc_graph_download;
                                       acc_cn_timings then gives the relabel's device time per step
                                       until the next timed pass. */

/* Context -------------------------------------------------------------- */
int acc_device_count(int *count);
int acc_ctx_create(int device, acc_ctx **out);
void acc_ctx_destroy(acc_ctx *ctx);
int acc_ctx_synchronize(acc_ctx *ctx);
/* The context's HIP stream (a hipStream_t), for callers that order their own work after it. */
void *acc_ctx_stream(acc_ctx *ctx);
const char *acc_last_error(void);

/* Resident graphs ------------------------------------------------------- */
/* Copies a host CSR (uint32 node_off[n+1], int32 edge_dst[m], the layout of
 * struct Graph, scan-query-cuda/util/graph.h:22-61) into HBM, validates it as
 * Graph::CheckInputGraph does (scan-query-cuda/util/graph.cu:196-214) plus
 * symmetry and id range (ACC_EFORMAT on failure, instead of exit(1)), and
 * prepares the per-graph work plans.  Flags other than the ACC_UPLOAD_* ones
 * are rejected (ACC_EINVAL). */
int acc_graph_upload(acc_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *node_off,
                     const int32_t *edge_dst, uint32_t flags, acc_graph **out);
void acc_graph_release(acc_graph *g);
/* The resident CSR and the caller id of every vertex (any pointer may be NULL).
 * After an ACC_UPLOAD_RELABEL upload this is the relabelled graph and
 * old_id[new vertex] = original id: what SCAN_XP::Reorder with the "deg"
 * method leaves in the Graph and in old_vid_dict (scan-query/scan_xp.cpp:1145-1222,
 * 1279-1301; the reference's order among equal degrees is unspecified, here it is
 * ascending original id), and the graph reports cluster ids in original ids.
 * Otherwise the uploaded CSR and the ids set by acc_graph_set_vertex_ids
 * (the identity by default). */
int acc_graph_download(const acc_graph *g, uint32_t *node_off /* n+1 */, int32_t *edge_dst /* m */,
                       int32_t *old_id /* n */);
/* n, m and the device bytes the graph holds (CSR, counts, scratch, plans). */
int acc_graph_info(const acc_graph *g, uint32_t *n, uint64_t *m, uint64_t *device_bytes);

/* All-edge counting ------------------------------------------------------ */
/* Enqueue the count of every slot on the context stream.  On completion the
 * graph's device count array holds cnt[i] = 2 + |N(u) ∩ N(v)| for every slot
 * i = (u,v) (|N(u) ∩ N(v)| with ACC_CN_RAW), the postcondition of the
 * reference's CheckCore on common_node_num (scan-query-cuda/scan_xp_common.cu:220-237),
 * and the triangle total T = sum over u<v slots of |N(u) ∩ N(v)| / 3. */
int acc_cn_compute(acc_graph *g, int path, uint32_t flags);
/* Synchronise and copy the counts (int32[m], may be NULL) and T (may be NULL) to the host. */
int acc_cn_download(acc_graph *g, int32_t *cnt_host, uint64_t *triangles);
/* compute + download in one call. */
int acc_cn_all_edges(acc_graph *g, int path, uint32_t flags, int32_t *cnt_host, uint64_t *triangles);
/* Device address of the graph's int32[m] count array (valid until release). */
const int32_t *acc_cn_device_counts(const acc_graph *g);

/* Per-kernel times (ms) of the last acc_cn_compute issued with ACC_CN_TIMING:
 * up to cap entries, names written as a ';'-separated list into names
 * (names_cap bytes).  Returns the number of kernels. */
int acc_cn_timings(acc_graph *g, float *ms, int cap, char *names, int names_cap);

/* Tuning parameters of a resident graph (the knobs the autotuner sets). */
#define ACC_PARAM_FWD_UNROLL 1 /* 64-element chunks per scan round in the forward path: 1, 2 or 4 */
#define ACC_PARAM_WAVE_SLOTS 2 /* forward bins: sources with at most this many upper slots (0..64) ... */
#define ACC_PARAM_WAVE_LOW 3   /* ... and |N<(b)| at most this (0..256) run one source per wave */
#define ACC_PARAM_CHUNK 4      /* upper slots per workgroup unit (64..512), all paths */
#define ACC_PARAM_AUTO_PATH 5  /* the path ACC_PATH_AUTO runs */
#define ACC_PARAM_PERSIST 6    /* forward counting grid: 0 one workgroup per task, k in [1, 256] at most k per CU
                                  looping over the tasks (persistent; default 256) */
#define ACC_PARAM_BMP_HASH_MAX 7 /* bitmap path: sources with d_u at most this (0..16384, default 256) keep
                                    N(u) in an LDS hash table, larger ones in an HBM bitmap; both behind an
                                    LDS index of N(u) */
#define ACC_PARAM_BMP_BATCH 8  /* bitmap path: hub bitmaps built, probed and cleared per batch (1..16384;
                                  the pool of one batch holds at most 8 GiB) */
int acc_graph_set_param(acc_graph *g, int param, int64_t value);
int acc_graph_get_param(const acc_graph *g, int param, int64_t *value);

/* Autotuner: coordinate descent over the forward path's parameters (unroll,
 * wave-bin boundaries, chunk = upper slots per 8-wave workgroup, persistent
 * grid), each trial timed on the device, then the bitmap and merge paths;
 * keeps the fastest as the graph's ACC_PATH_AUTO configuration and writes a
 * JSON report of every trial (ms per full pass, HIP events) into report (may
 * be NULL). */
#define ACC_TUNE_QUICK 0x1u        /* fewer candidates and repetitions */
#define ACC_TUNE_FORWARD_ONLY 0x2u /* skip the bitmap and merge trials */
int acc_tune(acc_graph *g, uint32_t flags, char *report, int report_cap);

/* Shards ------------------------------------------------------------------ */
/* Restricts the forward path (ACC_PATH_FORWARD passes and the triangle-only
 * pass) to share `index` of `parts`: the sources cut into `parts`
 * contiguous vertex ranges of equal estimated work (an upper slot (b, c)
 * weighs |N(c) ∩ [0, b)| + 4 + twice the triangles with middle b and top c,
 * a work unit of up to `chunk` upper slots of b weighs |N<(b)| + 256), the share
 * one GPU of a `parts`-GPU run would count.  While counts are one share's,
 * acc_scan_run refuses to cluster on them (ACC_EINVAL) instead of counting
 * the whole graph; acc_triangle_total on the forward path returns the share's
 * triangles (those whose middle vertex lies in the range).  The first call with parts > 1 runs one counting pass to
 * measure the triangles (the counts are then undefined until the next pass).
 * A pass over a share zeroes every slot and counts the triangles whose middle
 * vertex lies in the range; with ACC_CN_RAW the counts of all shares add up
 * to the whole graph's, and their triangle totals to its total.  parts = 1
 * restores the whole graph.  range_out (may be NULL) receives the share's
 * [first, last + 1) source range.  Replaces the reference's per-GPU range
 * split, which balances by sum over v in N(u), v > u of (d_v + 2)
 * (ComputeVertexRangeAccumulatedFilteredDegree,
 * scan-query-cuda/scan_xp_multi_gpu_multi_pass_common.cu:516-554). */
int acc_graph_set_shard(acc_graph *g, uint32_t parts, uint32_t index, uint32_t *range_out);

/* Multi-graph batches ----------------------------------------------------- */
/* A packed multi-graph CSR is an ordinary CSR whose graphs occupy disjoint
 * vertex ranges: graph k owns vertices [vertex_off[k], vertex_off[k+1]) (int32
 * graph offsets, vertex_off[num_graphs] = n) and no edge joins two ranges.
 * With a partition set, every counting pass also reduces per-graph triangle
 * totals, read with acc_cn_graph_totals (num_graphs entries). */
int acc_graph_set_partition(acc_graph *g, uint32_t num_graphs, const uint32_t *vertex_off);
int acc_cn_graph_totals(acc_graph *g, uint64_t *totals);

/* Capture one counting pass (the same kernels acc_cn_compute enqueues) into a
 * HIP graph, and replay it with a single launch on the context stream.  The
 * captured pass contains no memset nodes (counters are cleared by kernels). */
int acc_cn_capture(acc_graph *g, int path, uint32_t flags);
int acc_cn_replay(acc_graph *g);

/* Triangle count (SCAN_XP::TriCnt / TriCntSortedArray,
 * scan-query/scan_xp_triangle_cnt.cpp:778-957).  On the forward path (and
 * ACC_PATH_AUTO resolving to it) a dedicated pass counts every triangle once
 * at its middle vertex and writes no per-slot counts; the bitmap and merge
 * paths run their full count and return its T. */
int acc_triangle_total(acc_graph *g, int path, uint64_t *triangles);
/* Enqueues that triangle-only pass on the context stream without waiting
 * (for timing; the total is what acc_triangle_total returns). */
int acc_triangle_launch(acc_graph *g);

/* Allocation-free triangle total.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops and harnesses that forbid hipMalloc).  Enqueues on the
 * context stream the triangle count T of a CSR that already lies in device
 * memory (uint32 d_node_off[n+1], int32 d_edge_dst[m]; sorted lists without
 * self-loops, symmetric) and writes T into *d_triangles (device memory) on
 * completion; it allocates nothing and does not wait.  Its only other device
 * memory is a fixed 16.6 KB array inside the library (one partial sum per
 * workgroup of a grid of at most 2048 workgroups, and an error word), one per
 * device, whatever the graph's size; calls that overlap in time on one device
 * share it and race.  An offset outside [0, m] or an id outside [0, n) is
 * detected without reading outside the arrays and makes *d_triangles ~0.
 * ACC_EINVAL for n >= 2^31 or m > 2^32 - 256.  The same T as
 * acc_triangle_total (SCAN_XP::TriCnt, scan-query/scan_xp_triangle_cnt.cpp:778-957). */
int acc_triangle_total_device(acc_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *d_node_off,
                              const int32_t *d_edge_dst, unsigned long long *d_triangles);

/* SCAN queries ---------------------------------------------------------- */
/* Structural clustering on the resident counts, replacing the reference's
 * post-processing after CheckCore: PostComputeCoreChecking / CheckCoreCompSimCore
 * (scan-query-cuda/scan_xp_common.cu:220-268, scan-query/scan_xp.cpp:112-141),
 * ClusterCore (:1915-1928), LabelNonCore (:1940-1955), PostProcess
 * (:1957-1981), MarkClusterMinEleAsId (:1983-2020) and the non-core part of
 * PrepareResultOutput (:2022-2049).  eps = eps_num / eps_den exactly; a slot is
 * similar iff eps_den^2 cnt^2 >= eps_num^2 (d_u+1)(d_v+1), the exact form of the
 * reference's long-double test.  A vertex is core iff it has at least mu similar
 * neighbors; cluster ids are the minimum vertex id (see acc_graph_set_vertex_ids)
 * of the cluster's cores. */
typedef struct acc_scan_ws acc_scan_ws; /* per-query buffers and a stream */

/* Ids reported as cluster ids and in non-core pairs: old_id[new vertex] (the
 * reference's old_vid_dict after a relabel).  NULL restores the identity. */
int acc_graph_set_vertex_ids(acc_graph *g, const int32_t *old_id);
int acc_scan_ws_create(acc_graph *g, acc_scan_ws **out);
void acc_scan_ws_destroy(acc_scan_ws *ws);
/* Device bytes held and the workspace's HIP stream (a hipStream_t). */
int acc_scan_ws_info(const acc_scan_ws *ws, uint64_t *device_bytes, void **stream);
/* Enqueue one query on the workspace stream, ordered after the graph's counts
 * (computed first with ACC_PATH_AUTO if the graph has none). */
int acc_scan_run(acc_scan_ws *ws, int64_t eps_num, int64_t eps_den, int32_t mu, uint32_t flags);
int acc_scan_synchronize(acc_scan_ws *ws);
/* Copy results: label[n] (0 outlier, 1 core, 2 hub; the reference's
 * UNCLASSIFIED/CORE/HUB), cluster_id[n] (-1 for non-cores), sim[m] (0/1),
 * core_count[n], counts[4] = {cores, clusters, hubs, outliers}.  Any may be NULL. */
int acc_scan_results(acc_scan_ws *ws, int32_t *label, int32_t *cluster_id, uint8_t *sim, int32_t *core_count,
                     int64_t *counts);
/* Non-core memberships (cluster id, vertex id) as int32 pairs in pairs[2*cap],
 * unordered; *count receives the total (which may exceed cap). */
int acc_scan_pairs(acc_scan_ws *ws, int32_t *pairs, uint64_t cap, uint64_t *count);

/* Serving ------------------------------------------------------------------
 * A server keeps graphs resident in HBM with their counts computed once at
 * admission, admits and evicts them by device footprint against a budget
 * (least recently used idle graphs go first; an evicted graph is admitted
 * again from the server's host copy when a query needs it), and answers
 * SCAN queries (graph, eps = eps_num/eps_den, mu) submitted from any thread.
 * A worker thread batches pending queries across graphs into one set of SCAN
 * launches per batch, on one of `streams` streams, while earlier batches are
 * still running (continuous batching).  At admission the server also builds
 * a similarity index per graph (every vertex's slots sorted by a monotone
 * key of cn^2 / ((d_u+1)(d_v+1))), with which a query's core check is O(1)
 * per vertex and its clustering visits only similar slots; ties with eps
 * take the exact integer test, so results do not change.  The last kernel
 * of a query writes its results straight into pinned host memory.  Each query's labels (0 unclassified,
 * 1 core, 2 hub, as acc_scan_results) and cluster ids (-1 for non-cores) are
 * the ones acc_scan_run gives for the same graph and parameters alone; they
 * are copied to pinned host memory, readable with acc_server_result until
 * acc_server_release.  No counterpart in the reference (its SCAN_XP runs one
 * query per process, scan-query/scan_xp.cpp:2083-2110). */
typedef struct acc_server acc_server;
typedef struct acc_server_config {
  uint64_t hbm_budget; /* device bytes for resident graphs plus result slots */
  int32_t max_batch;   /* queries per batched launch, 1..14 (the job table is a kernel argument) */
  int32_t streams;     /* batches in flight at once, 1..16 */
  int32_t slots;       /* queries holding results at once (launched, not released), 1..4096 */
  int32_t no_index;    /* 1: build no similarity index (every query takes the per-slot kernels) */
} acc_server_config;
typedef struct acc_server_stats {
  uint64_t submitted, completed, failed, cancelled;
  uint64_t batches, batched_queries, max_batch_seen;
  uint64_t admissions, evictions;
  uint64_t graph_bytes, slot_bytes, resident_graphs, pending;
  uint64_t index_queries; /* queries decided through the similarity index */
} acc_server_stats;

#define ACC_QUERY_PENDING 0
#define ACC_QUERY_RUNNING 1
#define ACC_QUERY_DONE 2
#define ACC_QUERY_FAILED 3

void acc_server_config_default(acc_server_config *cfg); /* 64 GiB, batch 14, 4 streams, 32 slots */
int acc_server_create(acc_ctx *ctx, const acc_server_config *cfg, acc_server **out);
void acc_server_destroy(acc_server *s);
/* Registers a CSR (copied) and admits it; the first admission validates it
 * unless flags has ACC_UPLOAD_NO_VALIDATE (the only flag taken here; ACC_EINVAL
 * for others).  ACC_EBUDGET if it cannot fit.
 * Idle result slots are grown for the graph here, within the budget, so
 * that serving queries allocates nothing. */
int acc_server_add_graph(acc_server *s, uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst,
                         uint32_t flags, int32_t *graph_id);
int acc_server_remove_graph(acc_server *s, int32_t graph_id);
int acc_server_graph_info(acc_server *s, int32_t graph_id, uint32_t *n, uint64_t *m, int *resident,
                          uint64_t *device_bytes);
int acc_server_submit(acc_server *s, int32_t graph_id, int64_t eps_num, int64_t eps_den, int32_t mu,
                      uint64_t *ticket);
/* Waits for a query: ACC_OK when done, ACC_ETIMEDOUT after timeout_ms (< 0: no limit),
 * the query's own error code when it failed. */
int acc_server_wait(acc_server *s, uint64_t ticket, int32_t timeout_ms);
int acc_server_poll(acc_server *s, uint64_t ticket, int *state);
/* Results of a completed query: labels (int8: 0 unclassified, 1 core, 2 hub)
 * and cluster ids (-1 for non-cores) in pinned memory that stays valid until
 * acc_server_release; counts = cores, clusters, hubs, outliers; latency from
 * submission to completion. */
int acc_server_result(acc_server *s, uint64_t ticket, const int8_t **label, const int32_t **cluster_id,
                      uint32_t *n, int64_t *counts, double *latency_ms);
/* Frees a query's result slot; a pending query is cancelled. */
int acc_server_release(acc_server *s, uint64_t ticket);
int acc_server_get_stats(acc_server *s, acc_server_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* ACC_CN_H */
