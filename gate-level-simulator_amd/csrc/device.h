// Records shared between the host packer and the HIP kernels.
#pragma once
#include <cstdint>

namespace gls {

// Event word: (time << 2) | value, time < 2^30 time units.
constexpr uint32_t kTimeInf = 0x3FFFFFFFu;
constexpr uint32_t kWordInf = 0xFFFFFFFFu;

// One simulated output ("gate"), 80 bytes.  Delays are integer time units;
// 0xFFFF marks a missing IOPATH of that kind.
struct alignas(16) GateDev {
  uint32_t in_net[4];     // distinct input nets (resolved)
  uint32_t out_net;
  uint16_t lut;           // truth-table id (256 bytes each, index = sum v_j << 2j)
  uint8_t k;              // number of distinct inputs (1..4)
  uint8_t pad0;
  uint16_t dly[4][3][2];  // [input][both,pos,neg][rise,fall]
  uint8_t nrec[4];        // IOPATH records per input for this output
  uint16_t dmax;          // largest delay (>= 1)
  uint16_t pad1;
};
static_assert(sizeof(GateDev) == 80, "GateDev layout");

// Per time-segment summary written by the evaluation kernel.
struct SegSum {
  uint64_t scr_off;   // scratch offset of the segment's kept candidates
  uint32_t n_raw;     // kept candidates (after preemption, before dedup)
  uint32_t n_dd;      // after in-segment dedup (first entry always counted)
  uint32_t last_o;    // time of the last kept candidate
  uint8_t first_v, last_v, pad[2];
};

// Per-segment carry computed by the gate scan.
struct SegCarry {
  uint32_t pos;       // output index of the segment's first emitted entry
  uint32_t carry_t;   // time of the last kept candidate before the segment
  uint8_t carry_v;    // its value (4 = none: first entry always emitted)
  uint8_t pad[3];
};

}  // namespace gls
