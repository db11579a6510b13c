/*
 * acc_graph.h -- host-side graph utilities of acctricnt_amd (C ABI, no GPU).
 *
 * Builds the sorted, symmetric int32 CSR that the all-edge common-neighbor
 * path consumes, reads and writes the reference's on-disk format, and
 * generates the seeded synthetic graphs used by tests and benchmarks.
 *
 * CSR contract (same as the reference's struct Graph,
 * scan-query/util/graph.h:36-77): node_off is uint32[n+1], edge_dst is
 * int32[m] with every adjacency list strictly increasing, no self-loops,
 * and (u,v) present iff (v,u) present.  m = number of directed slots.
 *
 * Every function returns 0 on success or a negative ACC_E* code; the text of
 * the last error on the calling thread is available from acc_graph_last_error().
 */
#ifndef ACC_GRAPH_H
#define ACC_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct acc_hgraph acc_hgraph; /* opaque host CSR */

/* Error codes shared with acc_cn.h. */
#define ACC_OK 0
#define ACC_EINVAL (-1)   /* bad argument */
#define ACC_EFORMAT (-2)  /* malformed CSR or file: self-loop, unsorted, asymmetric, out of range */
#define ACC_ENOMEM (-3)   /* host or device allocation failed */
#define ACC_EIO (-4)      /* file could not be opened, read or written */
#define ACC_EDEVICE (-5)  /* HIP runtime error, kernel launch failure or missing device */
#define ACC_EBUDGET (-6)  /* request exceeds a configured memory budget */
#define ACC_ENOTFOUND (-7) /* unknown graph id, ticket or file entry */
#define ACC_ETIMEDOUT (-8) /* a wait ended before the work completed */

const char *acc_graph_last_error(void);

/* Accessors.  Pointers stay valid until acc_hgraph_free or a relabel. */
uint32_t acc_hgraph_n(const acc_hgraph *g);
uint64_t acc_hgraph_m(const acc_hgraph *g);
const uint32_t *acc_hgraph_offsets(const acc_hgraph *g);
const int32_t *acc_hgraph_adj(const acc_hgraph *g);
void acc_hgraph_free(acc_hgraph *g);

/* Build from an undirected edge list: self-loops dropped, both directions
 * inserted, duplicates removed, lists sorted.  Ids must lie in [0, n). */
int acc_hgraph_from_edges(uint32_t n, uint64_t num_edges, const int32_t *src,
                          const int32_t *dst, acc_hgraph **out);

/* Build from caller CSR arrays (copied), after validating the CSR contract.
 * Validation follows Graph::CheckInputGraph (scan-query/util/graph.cpp:122-143)
 * and additionally checks symmetry and id range. */
int acc_hgraph_from_csr(uint32_t n, uint64_t m, const uint32_t *node_off,
                        const int32_t *edge_dst, acc_hgraph **out);

/* Validate a CSR without copying it: returns ACC_OK or ACC_EFORMAT. */
int acc_csr_validate(uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst);

/* Relabel vertices by degree, descending, ties by old id ascending (a stable
 * version of SCAN_XP::ReorderDegDescending, scan-query/scan_xp.cpp:1279-1301,
 * whose unstable sort leaves tie order unspecified).  If old_of_new is not
 * NULL it receives n entries: old_of_new[new_id] = old_id. */
int acc_hgraph_relabel_degree(acc_hgraph *g, int32_t *old_of_new);

/* Seeded generators.  All randomness is counter-based (SplitMix64 keyed by
 * seed and sample index), so results do not depend on the thread count. */

/* Erdos-Renyi G(n, M)-style: num_samples uniform pairs (u != v). */
int acc_gen_er(uint32_t n, uint64_t num_samples, uint64_t seed, acc_hgraph **out);

/* R-MAT with quadrant probabilities a, b, c (d = 1-a-b-c) over 2^scale ids;
 * samples whose endpoints fall at or beyond n are rejected and redrawn, so
 * n need not be a power of two.  num_samples accepted samples are drawn. */
int acc_gen_rmat(uint32_t n, int scale, uint64_t num_samples, double a, double b, double c,
                 uint64_t seed, acc_hgraph **out);

/* Chung-Lu: expected degree of vertex i proportional to (i+1)^(-1/(gamma-1)),
 * scaled so the mean expected degree is avg_deg with every expected degree
 * capped at max_deg; n*avg_deg/2 endpoint pairs drawn by weight. */
int acc_gen_chung_lu(uint32_t n, double gamma, double avg_deg, double max_deg, uint64_t seed,
                     acc_hgraph **out);

/* On-disk format of the reference (scan-query/util/graph.cpp:48-120,
 * scan-query/converter/EdgeListToBinsConverter.cpp:87-126):
 *   <dir>/b_degree.bin = int32 4, uint32 n, uint32 m, int32 deg[n]
 *   <dir>/b_adj.bin    = int32 adj[m], lists concatenated in vertex order. */
int acc_hgraph_write_bins(const acc_hgraph *g, const char *dir);
int acc_hgraph_read_bins(const char *dir, acc_hgraph **out);

/* Text edge list -> b_degree.bin/b_adj.bin, following
 * EdgeListToBinsConverter (scan-query/converter/EdgeListToBinsConverter.cpp:24-126):
 * '#' lines skipped, only lines with first < second kept, n = max id + 1. */
int acc_convert_edge_list(const char *txt_path, const char *out_dir);

/* SCAN result file of the reference (Graph::Output(..., old_vid_dict),
 * scan-query/util/graph.cpp:176-216; name built by SCAN_XP::PrepareResultOutput,
 * scan-query/scan_xp.cpp:2056-2069): "c/n vertex_id cluster_id", then
 * "c <id> <cluster id>" for every core (label 1) in ascending id order, then
 * "n <id> <cluster id>" for the sorted unique (cluster id, id) pairs.  ids are
 * old_id[v] when old_id is not NULL.  pairs holds npairs (cluster id, id)
 * int32 pairs in any order, duplicates allowed; it is sorted in place. */
int acc_write_scan_result(const char *path, uint32_t n, const int32_t *label, const int32_t *cluster_id,
                          const int32_t *old_id, int32_t *pairs, uint64_t npairs);

/* "<dir>/scanxp-result-<eps>-<mu>.txt" with eps formatted as to_string(eps)
 * with trailing '0' and 'u' characters removed (scan-query/scan_xp.cpp:2060-2061). */
int acc_scan_result_path(const char *dir, double eps, int mu, char *buf, int cap);

/* Reverse-slot index on the host: rev[i] = slot of (v,u) for slot i = (u,v). */
int acc_hgraph_reverse_slots(const acc_hgraph *g, uint32_t *rev_out);

/* CPU recount of a count array, the reference's Verify step after CheckCore
 * (scan-query-cuda/scan_xp_common.cu:270-323, BIT_VEC form): for every slot
 * i = (u, v) with u < v, cnt[i] must equal base + |N(u) ∩ N(v)| (base 2 for
 * the reference's layout, 0 for raw counts), recounted with a per-thread bit
 * vector of N(u) probed by N(v).  *mismatches receives the number of such
 * slots that differ and *first_bad the smallest of them (m when none).  The
 * CSR is trusted (validated at upload).  nthreads <= 0: all threads. */
int acc_csr_verify_counts(uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst,
                          const int32_t *cnt, int32_t base, int nthreads, uint64_t *mismatches,
                          uint64_t *first_bad);

#ifdef __cplusplus
}
#endif

#endif /* ACC_GRAPH_H */
