// `glsgen config outdir [cells] [t_end] [mean_gap] [seed]` — writes the seeded
// synthetic workload of a benchmark configuration as lib.vlib, netlist.gv,
// netlist.sdf and stim.vcd.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gls.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::printf("usage: glsgen config(0|1|2) outdir [cells] [t_end_ps] [mean_gap_ps] [seed]\n");
    return 1;
  }
  int cfg = std::atoi(argv[1]);
  std::string dir = argv[2];
  int cells = argc > 3 ? std::atoi(argv[3]) : 0;
  static const unsigned t_ends[3] = {2000000u, 20000000u, 20000000u};
  static const double gaps[3] = {1000.0, 500.0, 200.0};
  if (cfg < 0 || cfg > 2) {
    std::fprintf(stderr, "glsgen: config must be 0, 1 or 2\n");
    return 1;
  }
  unsigned t_end = argc > 4 ? (unsigned)std::strtoul(argv[4], nullptr, 10) : t_ends[cfg];
  double gap = argc > 5 ? std::atof(argv[5]) : gaps[cfg];
  unsigned long long seed = argc > 6 ? std::strtoull(argv[6], nullptr, 0) : 0x5EED0000ull + cfg;
  gls_design* d = nullptr;
  int rc = gls_design_generate(cfg, seed, cells, 0, 0, &d);
  if (rc != GLS_OK) {
    std::fprintf(stderr, "glsgen: %s\n", gls_last_error());
    return 2;
  }
  rc = gls_design_write(d, (dir + "/lib.vlib").c_str(), (dir + "/netlist.gv").c_str(),
                        (dir + "/netlist.sdf").c_str());
  gls_stim* s = nullptr;
  if (rc == GLS_OK) rc = gls_stim_generate(d, seed, t_end, gap, 0.01, &s);
  if (rc == GLS_OK) rc = gls_stim_write_vcd(s, (dir + "/stim.vcd").c_str());
  if (rc != GLS_OK) {
    std::fprintf(stderr, "glsgen: %s\n", gls_last_error());
    return 2;
  }
  gls_stim_free(s);
  gls_design_free(d);
  return 0;
}
