// File formats of the drop-in boundary: vlib cell library, flat gate-level
// Verilog (gv), SDF IOPATH delays, VCD stimulus and the SAIF report.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "netlist.h"

namespace gls {

// Readers follow the accepted syntax of the reference's PreProcess /
// VCDParser (src/PreProcess.cpp:42-1464, src/VCDParser.cpp:283-717) and
// throw GlsError(GLS_E_FORMAT) where the reference would assert or exit.
Library parse_vlib_text(const std::string& text);
Library parse_vlib_file(const std::string& path);
// threads as for the VCD reader below; the result does not depend on it.
Netlist parse_gv_file(const std::string& path, const Library& lib, int threads = 0);
Netlist parse_gv_text(const std::string& text, const Library& lib, int threads = 0);
void parse_sdf_file(const std::string& path, Netlist& nl, int threads = 0);
void parse_sdf_text(const std::string& text, Netlist& nl, int threads = 0);
// threads: 0 = up to 16 (the body is read in parallel stretches when it is
// large), 1 = one thread; the result does not depend on it.
Stimulus parse_vcd_file(const std::string& path, int threads = 0);
Stimulus parse_vcd_text(const std::string& text);
Stimulus parse_vcd_text(const std::string& text, int threads);

// 64-bit FNV-1a digest of everything a parse produces (nets with their
// flags, instances with pins, unconnected outputs and delays bit for bit);
// two netlists with the same digest simulate identically.
uint64_t netlist_digest(const Netlist& nl);

// Writers (used by the generator and tests).
std::string gv_text(const Netlist& nl, const Library& lib);
std::string sdf_text(const Netlist& nl, const Library& lib);
std::string vcd_text(const Stimulus& st);

// SAIF report.  One record per net; `hier` records are nets named
// "inst[pin]" (outputs an instance leaves unconnected), written as nested
// INSTANCE blocks like the reference's CPU writer (src/simulate.cpp:305-377).
struct SaifRecord {
  std::string name;
  uint64_t t0 = 0, t1 = 0, tx = 0;
  uint64_t tc = 0;  // 0<->1 transitions in (start, end]
};
// with_tc: each record also carries (TC n) after (TX ...); off by default,
// which is the reference's three-field record (src/simulate.cu:133-135).
std::string saif_text(std::vector<SaifRecord> recs, uint64_t start, uint64_t end,
                      int timescale_ps = 1, bool with_tc = false);

std::string read_file(const std::string& path);
void write_file(const std::string& path, const std::string& data);

}  // namespace gls
