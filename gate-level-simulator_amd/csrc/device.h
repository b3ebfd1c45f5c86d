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

// Gate record of the fused level kernel, 160 bytes.  Delays are folded into
// two lookup tables per input j, indexed by the output edge class s (1..4)
// and by whether input j fell at this time:
//   tmin[j][s-1][fall]: contribution of j to the minimum taken when several
//                       inputs have an entry (0xFFFF: no related IOPATH),
//   tone[j][s-1][fall]: the delay when j is the only input with an entry.
struct alignas(16) GateRec {
  uint32_t in_net[4];
  uint32_t out_net;
  uint16_t lut;
  uint8_t k;
  uint8_t pad0;
  uint16_t dmax;
  uint16_t pad1[3];
  uint16_t tmin[4][4][2];
  uint16_t tone[4][4][2];
};
static_assert(sizeof(GateRec) == 160, "GateRec layout");

// Per-gate segmentation of one level: segment c covers the leader input's
// events [c*stride, (c+1)*stride).
struct GatePlan {
  uint32_t nseg, stride, lead, pad;
};

// Per-tile input ranges: L[j] = first event of input j at or after the
// tile's start time, E[j] = end of the range including the look-ahead halo.
struct TileDesc {
  uint32_t g, ti;
  uint32_t L[4], E[4];
  uint32_t pad[2];
};
static_assert(sizeof(TileDesc) == 48, "TileDesc layout");

// Arguments of the fused level kernel.
struct LevelArgs {
  const GateRec* gates;       // gates of this level
  const GatePlan* plan;       // per gate
  const TileDesc* tiles;      // per tile
  const uint32_t* lutp;       // truth tables packed 2 bits per entry, 16 words each
  const uint64_t* out_base;   // per gate: output region (pool word offset)
  uint64_t ntiles;
  uint64_t* woff;
  uint32_t* wlen;
  uint32_t* pool;
  const uint8_t* luts;
  unsigned long long* tstate;  // 2 words per tile: aggregate, inclusive prefix
  unsigned int* ticket;
  uint32_t* ovf;               // stacks of tiles that do not fit LDS
  unsigned long long ovf_cap;
  unsigned long long* ovf_used;
  unsigned long long* saif;
  unsigned long long* ev_count;
  unsigned int* err;
  unsigned long long start, end;
  uint32_t cap;                // LDS input capacity per tile (words)
  uint32_t pad;
};

// LDS bytes of one fused-level workgroup for an input capacity of `cap`
// words: inputs (cap + 64), stacks (cap + 64), truth table, delay tables.
inline size_t level_lds_bytes(uint32_t cap) { return (2ull * (cap + 64)) * 4 + 256 + 128; }

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
