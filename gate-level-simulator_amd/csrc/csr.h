// Levelizer and flat per-level CSR layout consumed by the device kernels.
//
// The reference packs per-level padded arrays of function chains and
// float delays (src/process.cpp:423-566) and re-evaluates the chain per
// event (src/simulate.cu:634-710).  Here every simulated output ("gate") is
// compiled once on the host into:
//   * a 4-state truth table over its distinct input nets (4^k 2-bit codes),
//   * per distinct input net, at most one IOPATH of each kind
//     (both edges / posedge / negedge) as integer rise/fall delays,
// and gates are ordered level by level.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "netlist.h"

namespace gls {

// Distinct input nets per gate: up to 6, as the reference's GPU path
// (MAX_IN_NUM, src/simulate.cu:25).
constexpr int kMaxK = 6;
constexpr uint16_t kNoDelay = 0xFFFF;

// Delay kinds within a gate's per-input record.
enum DelayKind { DK_BOTH = 0, DK_POS = 1, DK_NEG = 2 };

struct GateDelays {
  // For each distinct input j and kind: rise/fall in the simulation time
  // unit, or kNoDelay when the SDF has no such IOPATH.
  uint16_t v[kMaxK][3][2];
  uint8_t nrec[kMaxK];  // number of IOPATH records of input j (any kind)
  uint16_t dmax;        // max over all delays and the default 1
};

struct CompiledGate {
  int inst = -1;          // instance index
  int out_net = -1;       // net id written by this gate
  int k = 0;              // distinct inputs
  int in_net[kMaxK] = {-1, -1, -1, -1, -1, -1};
  int lut = -1;           // index into Program::luts
  // IOPATH records per distinct input, in SDF order: (kind, rise, fall)
  struct Rec { int kind; double rise, fall; };
  std::vector<Rec> recs[kMaxK];
};

// Net roles after resolution (precedence VCD > gate output > constant >
// assign, src/simulate.cpp:1617-1664).
enum NetRole : uint8_t { NR_UNDRIVEN = 0, NR_STIM = 1, NR_GATE = 2, NR_CONST = 3, NR_ALIAS = 4 };

struct Program {
  std::vector<std::string> net_names;   // netlist nets + "inst[pin]" pseudo nets
  std::vector<uint8_t> role;
  std::vector<int> const_val;           // for NR_CONST
  std::vector<uint8_t> in_saif;         // net gets a SAIF record
  std::vector<CompiledGate> gates;      // level order
  std::vector<uint32_t> level_start;    // gates of level l: [level_start[l], level_start[l+1])
  std::vector<std::vector<uint8_t>> luts;  // each 4^k entries, value codes
  std::vector<int> lut_k;
  int n_netlist_nets = 0;
  uint64_t n_pins = 0;                  // sum of distinct gate inputs
  int n_levels() const { return (int)level_start.size() - 1; }
};

// 4-state primitive evaluation with the symbolic semantics of the
// reference's calc() (src/simulate.cpp:382-870).
uint8_t eval_prim(int op, const uint8_t* in, int n);

// Compile netlist + library into a levelized program.  `stim_names` are the
// VCD signal names (they take precedence over every other driver).
// Throws GlsError(GLS_E_CYCLE) for combinational loops / undriven nets and
// GlsError(GLS_E_UNSUPPORTED) for gates outside the device domain.
Program compile_program(const Netlist& nl, const Library& lib,
                        const std::vector<std::string>& stim_names, int threads = 0);

// 64-bit FNV-1a digest of a compiled program: net names and roles, every
// gate (nets, truth-table number, IOPATH records bit for bit), the level
// boundaries and the truth tables.
uint64_t program_digest(const Program& pg);

// Integer delays for the given simulation time unit (ps per VCD tick):
// trunc(value_ps * factor) exactly as src/simulate.cpp:1165-1175.
GateDelays gate_delays(const CompiledGate& g, int timescale_ps);

}  // namespace gls
