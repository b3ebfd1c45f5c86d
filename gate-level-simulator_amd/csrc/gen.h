// Seeded synthetic workload generator: combinational netlists drawn from a
// generic cell library, SDF IOPATH delays and Poisson VCD stimulus.
//
// Every random draw goes through SplitMix64 and IEEE-754 basic arithmetic
// (no libm, no std::*_distribution), so the same seed gives the same bytes
// on every host.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "netlist.h"

namespace gls {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // Uniform in (0, 1].
  double u01() { return ((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
  // Uniform integer in [lo, hi].
  int64_t range(int64_t lo, int64_t hi) {
    return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1));
  }
};

// Deterministic natural log / exp built from + - * / only.
double det_log(double x);
double det_exp(double x);

struct NetlistSpec {
  uint64_t seed = 0x5EED0000ull;
  int n_cells = 2000;
  int n_pis = 64;
  int n_levels = 20;
  int n_types = 10;            // 10: INV..AOI21, 12: + AOI22, OAI22, 14: + AO221, AO222
  bool pareto_fanout = false;  // source choice weighted by Pareto(1.5) weights
  int fanout_cap = 256;        // realized fanout cap when pareto_fanout
  bool lognormal_widths = false;
  double lognormal_sigma = 2.0;
  int delay_min = 5, delay_max = 50;
};

struct StimulusSpec {
  uint64_t seed = 0x5EED0000ull;
  uint32_t t_end = 2000000;    // toggles are drawn in (0, t_end]
  double mean_gap = 1000.0;    // mean Poisson gap in ps
  double x_prob = 0.01;        // probability that a toggle goes to X
};

// The generic cell library text (module names and pins as in the
// reference's examples/GENERIC_STD_CELL.vlib) for the 14 generator cells.
const std::string& generic_vlib_text();

// Build a netlist (cells, nets, SDF delays) from the spec.  `lib` must hold
// the generic cells (parse generic_vlib_text()).
Netlist generate_netlist(const NetlistSpec& spec, const Library& lib,
                         std::vector<int>* level_of_cell = nullptr);

// Build stimulus for the netlist's primary inputs (nets declared `input`).
Stimulus generate_stimulus(const Netlist& nl, const StimulusSpec& spec);

// Named benchmark configurations (index 0..2 of the measurement plan; 3:
// configuration 1 with 5- and 6-input cells).
NetlistSpec config_netlist_spec(int config_index);
StimulusSpec config_stimulus_spec(int config_index);

}  // namespace gls
