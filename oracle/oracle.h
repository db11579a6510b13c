/*
 * oracle.h -- CPU restatement of the reference's all-edge common-neighbor
 * path, in plain C with OpenMP.  TEST INFRASTRUCTURE ONLY: the tests, the
 * package's smoke check and the benchmark's CPU-baseline leg use it as the
 * checker and the timed CPU baseline.  The product path never calls it.
 *
 * Every function cites the reference function it restates.  CSR arguments
 * follow struct Graph (scan-query/util/graph.h:36-77): node_off uint32[n+1],
 * edge_dst int32[m], sorted symmetric lists without self-loops.
 */
#ifndef ACC_ORACLE_H
#define ACC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SCAN_XP::CheckCoreSortedArray with ComputeCNNaiveStdMerge
 * (scan-query/scan_xp.cpp:675-857, set-inter/csr_set_intersection_serial.cpp:108-135):
 * cnt[i] = 2 + |N(u) ∩ N(v)| for every slot.  Only u<v slots are merged; the
 * reverse slot is assigned through a binary search of u in N(v)
 * (scan_xp.cpp:828-829).  Slots are scheduled dynamic,6000 as in the reference. */
void oracle_cn_merge(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                     int nthreads);

/* The same computation restricted to source vertices [u_lo, u_hi): slots of
 * those sources with u<v and their reverse slots are written; other slots are
 * left untouched.  Used for bounded CPU-baseline samples. */
void oracle_cn_merge_range(uint32_t n, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                           uint32_t u_lo, uint32_t u_hi, int nthreads);

/* The same merge for an explicit list of u<v slots (slot index and its source
 * vertex), writing the slot and its reverse.  Used to time a uniform random
 * sample of the edges for the CPU baseline. */
void oracle_cn_merge_slots(uint32_t n, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                           const uint32_t *slots, const int32_t *srcs, uint64_t k, int nthreads);

/* SCAN_XP::CheckCoreBitmap without BIT_VEC_INDEX (scan-query/scan_xp.cpp:220-337):
 * per-thread bit vector of N(u), probe N(v) (ComputeCNHashBitVec,
 * set-inter/csr_set_intersection_serial.h:124-132) for pairs with
 * deg(u) > deg(v) or (equal degrees and u < v); reverse slot by binary search. */
void oracle_cn_bitmap(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, int32_t *cnt,
                      int nthreads);

/* SCAN_XP::TriCntSortedArray (scan-query/scan_xp_triangle_cnt.cpp:778-957) with
 * the default ComputeCNNaive merge: T = sum over u, v in N<(u) of
 * |N<(u) ∩ N<(v)| where N<(x) = neighbors with id < x. */
uint64_t oracle_triangle_count(uint32_t n, const uint32_t *off, const int32_t *adj, int nthreads);

/* SCAN_XP::TriCntBitmap without BIT_VEC_INDEX
 * (scan-query/scan_xp_triangle_cnt.cpp:655-721): per thread a bit vector of
 * N<(u), set once per u, every N<(v), v in N<(u), probed against it
 * (ComputeCNHashBitVec, set-inter/csr_set_intersection_serial.h:124-132),
 * then cleared.  The same T as oracle_triangle_count. */
uint64_t oracle_triangle_count_bitmap(uint32_t n, const uint32_t *off, const int32_t *adj, int nthreads);

/* Similarity decision exactly as SCAN_XP::CheckCoreCompSimCore
 * (scan-query/scan_xp.cpp:112-141): sim = (double)((long double)cnt /
 * sqrtl((d_u+1)(d_v+1))) >= eps.  Writes sim[m] (0/1), core_count[n] and
 * label[n] (1 = CORE when core_count >= mu, else 0).  core_count is computed
 * without the reference's unsynchronised increments. */
void oracle_scan_core_fp(uint32_t n, const uint32_t *off, const int32_t *adj, const int32_t *cnt,
                         double eps, int mu, uint8_t *sim, int32_t *core_count, int32_t *label);

/* The exact integer form of the same decision with eps = eps_num/eps_den:
 * similar iff eps_den^2 * cnt^2 >= eps_num^2 * (d_u+1)(d_v+1), in 128-bit. */
void oracle_scan_core_int(uint32_t n, const uint32_t *off, const int32_t *adj, const int32_t *cnt,
                          int64_t eps_num, int64_t eps_den, int mu, uint8_t *sim, int32_t *core_count,
                          int32_t *label);

/* SCAN_XP::ClusterCore + LabelNonCore + PostProcess + MarkClusterMinEleAsId
 * (scan-query/scan_xp.cpp:1915-2020, util/union_find.h:4-36).
 * In: label[n] (CORE marks from the core step; updated in place with HUB=2),
 * sim[m].  old_id may be NULL (identity).  Out: root[n] = union-find root
 * (minimum vertex id of the core's cluster; -1 for non-cores), cluster_id[n]
 * = minimum old id among the cores of that root (-1 for non-cores), and the
 * four counters of PostProcess.  Returns 0. */
int oracle_scan_cluster(uint32_t n, const uint32_t *off, const int32_t *adj, const uint8_t *sim,
                        int32_t *label, const int32_t *old_id, int32_t *root, int32_t *cluster_id,
                        int64_t *core_num, int64_t *cluster_num, int64_t *hub_num, int64_t *outlier_num);

/* SCAN_XP::PrepareResultOutput + Graph::Output(..., old_vid_dict)
 * (scan-query/scan_xp.cpp:2022-2074, scan-query/util/graph.cpp:176-216):
 * writes <path> with the header "c/n vertex_id cluster_id", one "c <old id>
 * <cluster id>" line per core sorted by old id, then the sorted unique
 * (cluster id, old id) pairs of non-core vertices similar to a core as
 * "n <old id> <cluster id>".  Returns 0, or -4 if the file cannot be written. */
int oracle_scan_write_result(const char *path, uint32_t n, const uint32_t *off, const int32_t *adj,
                             const uint8_t *sim, const int32_t *label, const int32_t *cluster_id,
                             const int32_t *old_id);

/* Threads an oracle call uses by default (omp_get_max_threads). */
int oracle_threads(void);

/* File name suffix of the result: to_string(eps) with trailing '0'/'u'
 * characters removed (scan-query/scan_xp.cpp:2060-2061).  buf >= 64 bytes. */
void oracle_eps_string(double eps, char *buf);

#ifdef __cplusplus
}
#endif

#endif
