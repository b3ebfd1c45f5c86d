// `process vlib_path gv_path sdf_path inter_file` — drop-in for the
// reference's src/process.cpp:178-194: parse the library, netlist and SDF,
// levelize, and write the intermediate file read by `simulate`.
#include <chrono>
#include <cstdio>

#include "gls.h"

int main(int argc, char** argv) {
  if (argc != 5) {
    std::printf("[USAGE] ./process vlib_path gv_path sdf_path [intermediate_file]\n");
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  gls_design* d = nullptr;
  int rc = gls_design_from_files(argv[1], argv[2], argv[3], &d);
  if (rc != GLS_OK) {
    std::fprintf(stderr, "process: %s (status %d)\n", gls_last_error(), rc);
    return 2;
  }
  gls_design_info info;
  rc = gls_design_info_get(d, &info);
  if (rc != GLS_OK) {
    std::fprintf(stderr, "process: %s (status %d)\n", gls_last_error(), rc);
    gls_design_free(d);
    return 2;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::printf("INFO: instances %llu nets %llu inputs %llu\n", (unsigned long long)info.n_instances,
              (unsigned long long)info.n_nets, (unsigned long long)info.n_inputs);
  std::printf("levels size: %llu\n", (unsigned long long)info.n_levels);
  std::printf("total time of parse + topology order: %lldms\n",
              (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
  rc = gls_design_save(d, argv[4]);
  gls_design_free(d);
  if (rc != GLS_OK) {
    std::fprintf(stderr, "process: %s (status %d)\n", gls_last_error(), rc);
    return 2;
  }
  return 0;
}
