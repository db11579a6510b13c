// Drop-in for the reference's SCAN executables.  Accepted argument layouts:
//   [dir eps mu threads reorder log]            scan-query/main_scanxp.cpp:26-38 (CPU build)
//   [dir eps mu threads log]                    scan-query-cuda/main_scanxp.cpp:22-30 (GPU build)
//   [dir eps mu threads num_pass num_warps log] scan-query-cuda/experimental/main_scanxp.cpp:24-34
// num_pass and num_warps are accepted and unused: the graph stays resident in
// HBM (no destination-range passes) and launch shapes are chosen per graph.
#include <omp.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "log.h"
#include "scan_xp.h"

std::string reorder_method;
std::string dir;

static void Usage() {
  std::cout << "Usage: [1]exe [2]graph-dir [3]similarity-threshold "
               "[4]density-threshold [5 thread_num] [6 reorder_method] [7 log-file]\n";
}

int main(int argc, char *argv[]) {
  if (argc < 4) {
    Usage();
    return 0;
  }
  const double eps = std::strtod(argv[2], nullptr);
  const int mu = std::atoi(argv[3]);
  const int numt = argc > 4 ? std::atoi(argv[4]) : omp_get_max_threads();
  const char *log_path = nullptr;
  reorder_method = "";
  if (argc == 6) {
    log_path = argv[5];
  } else if (argc == 7) {
    reorder_method = argv[5];
    log_path = argv[6];
  } else if (argc >= 8) {
    log_path = argv[7];
  }
  if (numt > 0) omp_set_num_threads(numt);
  if (log_path) {
    FILE *f = std::fopen(log_path, "a+");
    if (f) log_set_fp(f);
  }
  dir = argv[1];
  log_info("graph dir: %s", argv[1]);
  SCAN_XP scanxp(numt, mu, eps, argv[1], argv[2]);
  scanxp.Execute();
  return 0;
}
