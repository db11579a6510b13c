// Drop-in for the reference's triangle-count executables:
//   [dir threads reorder log]   scan-query/main_triangle_count.cpp:13-32
#include <omp.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "log.h"
#include "scan_xp.h"

std::string reorder_method;
std::string dir;

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::cout << "Usage: [1]exe [2]graph-dir [3]thread_num [4]reorder_method [5]log-file\n";
    return 0;
  }
  const int numt = argc > 2 ? std::atoi(argv[2]) : omp_get_max_threads();
  if (numt > 0) omp_set_num_threads(numt);
  if (argc > 4) {
    FILE *f = std::fopen(argv[4], "a+");
    if (f) log_set_fp(f);
  }
  log_info("graph dir: %s", argv[1]);
  dir = std::string(argv[1]);
  SCAN_XP scan_xp(numt, argv[1]);
  reorder_method = argc > 3 ? std::string(argv[3]) : std::string();
  scan_xp.TriCnt();
  return 0;
}
