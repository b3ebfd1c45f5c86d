// `simulate inter_file vcd_path dumpon_time dumpoff_time saif_out` — drop-in
// for the reference's GPU simulator CLI (src/simulate.cu:1786-1812): load the
// intermediate file, read the VCD, simulate every level on the GPU and write
// the SAIF report for the window [dumpon, dumpoff].
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gls.h"

static int fail(const char* what, int rc) {
  std::fprintf(stderr, "simulate: %s: %s (status %d)\n", what, gls_last_error(), rc);
  return 2;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::printf(
        "[USAGE] ./simulate [intermediate_file] vcd_path dumpon_time(ps) dumpoff_time(ps) "
        "saif_out\n");
    return 1;
  }
  unsigned long long on = std::strtoull(argv[3], nullptr, 10);
  unsigned long long off = std::strtoull(argv[4], nullptr, 10);
  auto t0 = std::chrono::steady_clock::now();
  gls_design* d = nullptr;
  int rc = gls_design_load(argv[1], &d);
  if (rc != GLS_OK) return fail("loading the intermediate file", rc);
  auto t1 = std::chrono::steady_clock::now();
  gls_stim* s = nullptr;
  rc = gls_stim_from_vcd(argv[2], &s);
  if (rc != GLS_OK) return fail("reading the VCD", rc);
  auto t2 = std::chrono::steady_clock::now();
  gls_ctx* c = nullptr;
  rc = gls_ctx_create(0, &c);
  if (rc != GLS_OK) return fail("opening the GPU", rc);
  gls_run* r = nullptr;
  rc = gls_simulate(c, d, s, on, off, nullptr, &r);
  if (rc != GLS_OK) return fail("simulating", rc);
  auto t3 = std::chrono::steady_clock::now();
  rc = gls_run_write_saif(r, argv[5]);
  if (rc != GLS_OK) return fail("writing the SAIF", rc);
  auto t4 = std::chrono::steady_clock::now();
  gls_run_stats st;
  gls_run_stats_get(r, &st);
  auto ms = [](auto a, auto b) {
    return (long long)std::chrono::duration_cast<std::chrono::milliseconds>(b - a).count();
  };
  std::printf("total time of loading inter file: %lldms\n", ms(t0, t1));
  std::printf("total time of vcd parse: %lldms\n", ms(t1, t2));
  std::printf("total time of GPU simulation: %lldms (device %.3f ms, %llu gate events, %.3f G events/s)\n",
              ms(t2, t3), st.kernel_ms, (unsigned long long)st.n_events,
              st.kernel_ms > 0 ? st.n_events / (st.kernel_ms * 1e6) : 0.0);
  std::printf("total time of saif write: %lldms\n", ms(t3, t4));
  gls_run_free(r);
  gls_ctx_destroy(c);
  gls_stim_free(s);
  gls_design_free(d);
  return 0;
}
