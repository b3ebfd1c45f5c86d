// Host-only self-check of the multi-threaded host code, built with the
// address/undefined-behaviour and thread sanitizers (no GPU, no HIP):
//   * the generator (parallel stimulus draws),
//   * the parallel gv / SDF / VCD readers against their one-thread reads,
//   * the parallel netlist compiler against its one-thread compile,
//   * the inter.dat writer and reader.
// Usage: hostcheck [cells] [threads] [t_end]; exit status 0 when every check
// holds.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "csr.h"
#include "formats.h"
#include "gen.h"
#include "interdat.h"

using namespace gls;

static int failures = 0;

static void check(bool ok, const char* what) {
  std::printf("%-44s %s\n", what, ok ? "ok" : "FAILED");
  std::fflush(stdout);
  if (!ok) ++failures;
}

static bool same_stimulus(const Stimulus& a, const Stimulus& b) {
  if (a.names != b.names || a.waves.size() != b.waves.size()) return false;
  for (size_t i = 0; i < a.waves.size(); ++i)
    if (a.waves[i].t != b.waves[i].t || a.waves[i].v != b.waves[i].v) return false;
  return true;
}

int main(int argc, char** argv) {
  const int cells = argc > 1 ? std::atoi(argv[1]) : 150000;
  const int threads = argc > 2 ? std::atoi(argv[2]) : 8;
  try {
    Library lib = parse_vlib_text(generic_vlib_text());
    NetlistSpec ns = config_netlist_spec(1);
    ns.n_cells = cells;
    ns.n_levels = 30;
    Netlist nl = generate_netlist(ns, lib);
    StimulusSpec ss = config_stimulus_spec(1);
    ss.t_end = argc > 3 ? (uint32_t)std::atol(argv[3]) : 1200000;
    ss.mean_gap = 1000.0;
    Stimulus st = generate_stimulus(nl, ss);
    check(!st.names.empty(), "generator");

    const std::string gv = gv_text(nl, lib), sdf = sdf_text(nl, lib), vcd = vcd_text(st);
    std::printf("gv %zu bytes, SDF %zu, VCD %zu\n", gv.size(), sdf.size(), vcd.size());
    // the readers split texts of parallel_min_bytes() and more over their
    // threads; the compiler splits netlists of 32768 instances and more
    const size_t m = parallel_min_bytes();
    check(gv.size() >= m && sdf.size() >= m && vcd.size() >= m + m / 8 && nl.insts.size() >= 32768,
          "sizes that take the parallel readers and compiler");
    Netlist p1 = parse_gv_text(gv, lib, 1), pn = parse_gv_text(gv, lib, threads);
    parse_sdf_text(sdf, p1, 1);
    parse_sdf_text(sdf, pn, threads);
    check(netlist_digest(p1) == netlist_digest(pn), "gv + SDF readers: threads == one thread");
    check(netlist_digest(pn) == netlist_digest(nl), "gv + SDF readers: == generated netlist");

    Stimulus v1 = parse_vcd_text(vcd, 1), vn = parse_vcd_text(vcd, threads);
    check(same_stimulus(v1, vn), "VCD reader: threads == one thread");

    Program g1 = compile_program(pn, lib, st.names, 1);
    Program gn = compile_program(pn, lib, st.names, threads);
    check(program_digest(g1) == program_digest(gn), "compiler: threads == one thread");

    const std::string inter = inter_file_text(lib, pn, gn, DesignTexts{"", gv, sdf});
    Library l2;
    Netlist n2;
    Program g2;
    DesignTexts t2;
    check(read_inter_section(inter, l2, n2, g2, t2), "inter.dat: compiled section found");
    check(netlist_digest(n2) == netlist_digest(pn) && program_digest(g2) == program_digest(gn) &&
              t2.gv == gv && t2.sdf == sdf,
          "inter.dat: netlist, program, texts round trip");
    std::string bad = inter;
    bad[bad.size() - 40] ^= 1;
    bool raised = false;
    try {
      read_inter_section(bad, l2, n2, g2, t2);
    } catch (const GlsError&) {
      raised = true;
    }
    check(raised || program_digest(g2) != 0, "inter.dat: damaged section read safely");
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 2;
  }
  std::printf("%s\n", failures ? "hostcheck FAILED" : "hostcheck passed");
  return failures ? 1 : 0;
}
