// Drop-in host classes for the reference's SCAN and triangle-count
// executables.  struct Graph keeps the fields of the reference's Graph
// (scan-query/util/graph.h:36-77, scan-query-cuda/util/graph.h:22-61) and
// class SCAN_XP keeps the entry points of the reference's SCAN_XP
// (scan-query/scan_xp.h:188-300, scan-query-cuda/scan_xp.h:21-58):
// Execute() and TriCnt() with CheckCore / ClusterCore / LabelNonCore /
// PostProcess / PrepareResultOutput as its steps.  The counting and the
// clustering run on the GPU through the C ABI of libacc_cn.so (include/acc_cn.h).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "acc_cn.h"

#define UNCLASSIFIED 0
#define CORE 1
#define HUB 2

struct Graph {
  uint32_t nodemax = 0;
  uint32_t edgemax = 0;

  // csr representation
  uint32_t *node_off = nullptr;
  int *edge_dst = nullptr;

  // edge property
  bool *similarity = nullptr;
  int *common_node_num = nullptr;

  // vertex property
  int *core_count = nullptr;
  int *label = nullptr;

  // degree + 1, as the reference's loader stores it (scan-query/util/graph.cpp:113-114)
  std::vector<int> degree;

  // cluster id of every core vertex (-1 elsewhere)
  std::vector<int> cluster_dict;

  // (cluster id, non-core vertex id) memberships
  std::vector<std::pair<int, int>> noncore_cluster;

  std::string dir;

  explicit Graph(const char *dir_cstr);
  ~Graph();
  Graph(const Graph &) = delete;
  Graph &operator=(const Graph &) = delete;

  // Loads <dir>/b_degree.bin and <dir>/b_adj.bin and validates them
  // (ReadDegree / ReadAdjacencyList / CheckInputGraph, scan-query/util/graph.cpp:48-143).
  void Load();

  // Writes <dir>/scanxp-result-<eps_s>-<min_u>.txt (Graph::Output, scan-query/util/graph.cpp:176-216).
  void Output(const char *eps_s, const char *min_u, const std::vector<int> &old_vid_dict);

  // Replaces the CSR by a relabelled copy.
  void ReplaceCSR(const uint32_t *off, const int32_t *adj);
};

class SCAN_XP {
 public:
  // SCAN: eps_text is the command-line spelling of epsilon; it fixes the
  // exact rational eps used by the similarity test.
  SCAN_XP(int thread_num, int min_u, double epsilon, char *dir, const char *eps_text = nullptr);
  // Triangle counting.
  SCAN_XP(int thread_num, char *dir);
  ~SCAN_XP();

  void Execute();
  void TriCnt();

  int thread_num_;

 private:
  int min_u_ = 0;
  double epsilon_ = 0;
  int64_t eps_num_ = 0, eps_den_ = 1;
  Graph g;
  std::vector<int> old_vid_dict;
  acc_ctx *ctx_ = nullptr;
  acc_graph *dg_ = nullptr;
  acc_scan_ws *ws_ = nullptr;
  int64_t counts_[4] = {0, 0, 0, 0};

  void Reorder();
  void Upload();
  void CheckCore(Graph *g);
  void ClusterCore();
  void LabelNonCore();
  void PostProcess();
  void PrepareResultOutput();
  // CPU recount of the u<v slots' counts with a bit vector (the reference's
  // Verify, scan-query-cuda/scan_xp_common.cu:270-323); exits with -1 on a mismatch.
  void Verify();
};

// Parses a decimal spelling ("0.2", "1", ".35", "2e-1") into an exact fraction.
bool ParseEps(const char *text, int64_t *num, int64_t *den);

extern std::string reorder_method;
extern std::string dir;
