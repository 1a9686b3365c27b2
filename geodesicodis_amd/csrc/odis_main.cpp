// odis_main.cpp -- command-line entry with the reference's shape
// (src/main.cpp:46-68): run the case in the current (or given) directory.
//   odis [case_dir] [--grid-dir DIR] [--fp32]
#include <cstdio>
#include <cstring>
#include <string>

#include "odis.h"

int main(int argc, char **argv) {
  const char *dir = ".";
  const char *grid_dir = nullptr;
  int precision = ODIS_FP64;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--grid-dir") && i + 1 < argc) grid_dir = argv[++i];
    else if (!std::strcmp(argv[i], "--fp32")) precision = ODIS_FP32;
    else dir = argv[i];
  }
  const int rc = odis_run_case(dir, grid_dir, precision);
  if (rc < 0) std::fprintf(stderr, "odis: error %d (see %s/DATA/ERROR.txt)\n", rc, dir);
  return rc < 0 ? 1 : 0;
}
