// Records shared between the host packer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gls {

// Event word: (time << 2) | value, time < 2^30 time units of the chunk's
// relative time (see ChunkWin).  Resident stimulus words are 64-bit:
// (absolute time << 2) | value.
constexpr uint32_t kTimeInf = 0x3FFFFFFFu;
constexpr uint32_t kWordInf = 0xFFFFFFFFu;
// Largest span of one chunk in relative time: every event and candidate
// time of a chunk stays below it (room is kept for one more delay).
constexpr uint64_t kRelSpan = (1ull << 30) - (1ull << 18);

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

// Gate record of the fused level path, 48 bytes.  Fan-in 1..4: the delay of
// an output event is one lookup in the gate's delay table (u16 entries at
// dtab + dt_off) indexed by dt_index(): the output edge class s (1..4) and,
// per input j, whether it has an event at this time (bit 2j) and whether
// that event fell (bit 2j+1).  Fan-in 5..6 ("wide" gates): the table at
// dt_off holds per-input delays instead (wide_dt_index), and the truth
// table of 4^k entries spans 4^k / 256 consecutive 256-byte blocks from
// block `lut`.
struct alignas(16) GateRec {
  uint32_t in_net[6];
  uint32_t out_net;
  uint32_t dt_off;
  uint16_t lut;   // first 256-byte truth-table block
  uint8_t k;
  uint8_t pad0;
  uint16_t dmax;
  uint16_t job;  // instance of a batch (SAIF window)
  uint32_t pad1[2];
};
static_assert(sizeof(GateRec) == 48, "GateRec layout");

// Largest fan-in of the lane-parallel walk; gates above it (up to 6
// distinct inputs, the reference's MAX_IN_NUM, src/simulate.cu:25) are
// walked by the wide-gate kernel.
constexpr uint32_t kTileK = 4;
constexpr uint32_t kWideK = 6;

// Delay-table entries of a gate with k inputs, and the index of one entry
// (mask: inputs with an event; fell: the subset of them that fell).
inline constexpr uint32_t dt_size(uint32_t k) { return 4u << (2 * k); }

// Wide gates: per input j, output edge class s (1..4) and whether the
// input's own event fell: the smallest delay among its edge-independent
// IOPATH and the one of its edge kind (tmin, 0xFFFF: none), and the delay
// used when it is the only input with an event (tone).
constexpr uint32_t kWideDtSize = kWideK * 4 * 2 * 2;
__host__ __device__ inline uint32_t wide_dt_index(uint32_t j, uint32_t s, uint32_t fell,
                                                  uint32_t one) {
  return (((j * 4 + (s - 1)) * 2 + fell) << 1) | one;
}
__host__ __device__ inline uint32_t dt_index(uint32_t k, uint32_t s, uint32_t mask,
                                             uint32_t fell) {
  uint32_t bits = 0;
  for (uint32_t j = 0; j < k; ++j)
    bits |= (((mask >> j) & 1u) << (2 * j)) | (((fell >> j) & 1u) << (2 * j + 1));
  return ((s - 1) << (2 * k)) | bits;
}

// Per-gate segmentation of one level: segment c covers the leader input's
// events [c*stride, (c+1)*stride).
struct GatePlan {
  uint32_t nseg, stride, lead, pad;
};

// Everything one tile of the fused level kernel needs, 128 bytes, written by
// k_tile_desc.  Input j's range covers events [B, E) of its waveform with
// B = L - 1 (L > 0) or 0: L = first event at or after the tile's start time,
// E = end of the range including the look-ahead halo; in_off = pool word of
// event B.
struct alignas(16) TileDesc {
  uint64_t in_off[4];
  uint32_t L[4], E[4];
  uint32_t lutw[4];                // packed truth table words (fan-in <= 3)
  uint32_t g, ti, nseg, stride;    // gate (index in the level), tile of the gate, segments
  uint32_t lead, k, dt_off, lut;   // leader input, fan-in, delay table, truth table id
  uint32_t dmax, out_net, job;
  uint32_t out_cap;  // words of the tile's output region in the level scratch
};
static_assert(sizeof(TileDesc) == 128, "TileDesc layout");

// What the walk kernel leaves for one tile, 32 bytes: its output entries
// (deduplicated inside the tile) sit in the level scratch; the per-gate
// carry pass joins consecutive tiles.
struct alignas(16) TileOut {
  uint32_t cnt;         // entries after dedup inside the tile
  uint32_t ft, lt;      // times of the first and last entries
  uint32_t flags;       // first value | last value << 2 | has entries << 4
  uint32_t t0, t1, tx;  // SAIF durations between the tile's own entries (clipped)
  uint32_t ev;          // gate events of the tile (input-event times from the chunk start on)
};
static_assert(sizeof(TileOut) == 32, "TileOut layout");

// SAIF counters per net: T0, T1, TX durations and TC, the number of direct
// 0<->1 transitions at times in (dumpon, dumpoff].
constexpr uint32_t kSaifWords = 4;

// A value change from p to v is a toggle when it is 0 -> 1 or 1 -> 0
// (codes 0 and 3; X <-> Z, codes 1 and 2, also differ by 3 and is excluded).
__host__ __device__ inline bool is_toggle(uint32_t p, uint32_t v) {
  return (p ^ v) == 3u && (v == 0u || v == 3u);
}

// Time chunks.  A simulation runs as a sequence of time chunks
// [F_k, F_k+1); every chunk runs all levels.  Inside chunk k, times are
// relative: rel = abs - (lb_k - 1) with lb_k = F_k - W (W = the largest gate
// delay; lb_0 = 0), so every net's chunk waveform starts with a state entry
// at rel 0 (its value before lb_k), followed by its entries at abs times in
// [lb_k, F_k+1).  Entries at abs times in [lb_k, F_k) are the look-back
// window: they were final after chunk k-1 and are walked again only so that
// the candidates they make with output times >= F_k are known.
//
// Per instance and chunk: SAIF bounds in the chunk's relative time.
// Durations are summed over [lo, hi) (the chunk's part of the SAIF window);
// toggles are counted at entry times in (tc_lo, tc_hi].
struct ChunkWin {
  unsigned long long lo, hi, tc_lo, tc_hi;
};

// Tails of the gate outputs, the first words of their chunk waveforms: per
// net, the state entry and the look-back entries, kept from the previous
// chunk (in its relative time; `shift` = this chunk's time base minus the
// previous one).  In chunk 0 every tail is one X state entry.
struct TailArgs {
  const uint32_t* tail_len;  // per net
  const uint64_t* tail_off;  // per net: first word in `tails`
  const uint32_t* tails;
  uint32_t shift;
  uint32_t first;            // 1 in chunk 0
};

// Arguments of the fused level (walk) kernel.
struct LevelArgs {
  const TileDesc* tiles;      // per tile
  const uint32_t* pool;       // input waveforms
  const uint16_t* dtab;       // delay tables (GateRec::dt_off)
  const uint8_t* luts;        // full truth tables (fan-in 4)
  const uint64_t* ntiles;     // tiles of the level (device memory)
  const uint64_t* scr_base;   // per gate: scratch word of its first tile's region
  uint32_t* scratch;          // tile outputs
  TileOut* tout;              // per tile
  uint32_t* ovf;              // stacks of tiles that do not fit LDS
  unsigned long long ovf_cap;
  unsigned long long* ovf_used;
  unsigned long long* ticket;   // next tile to hand out, minus the grid size (0 per level)
  unsigned long long* tiles_walked;  // [0] tiles walked, [1] of them on the HBM path (per run)
  unsigned long long* ev_count;
  unsigned int* err;
  const ChunkWin* cwin;        // SAIF bounds of this chunk per instance
  uint32_t cap;                // LDS input capacity per tile (words)
  uint32_t prefetch;           // 1: stage the next tile before storing this one
  uint32_t* stk;               // null: stacks in LDS; else (cap + 64) words per workgroup
  uint32_t lo_rel;             // F_k in chunk time: candidates below it only pre-empt
  uint32_t hi_rel;             // F_k+1 in chunk time (0xFFFFFFFF: last chunk): candidates
                               // at or above it belong to the next chunk and are dropped
  uint32_t first;              // 1 in chunk 0: outputs start with (0, X), f before t = 0 is X
};

// Arguments of the wide-gate kernel (gates of fan-in 5..6 of one level).
struct WideArgs {
  const GateRec* gates;        // the level's gates
  const uint32_t* wide;        // program gate indices of the level's wide gates
  uint32_t nwide;
  uint32_t g0;                 // program index of the level's first gate
  const uint64_t* tile_base;   // per gate of the level: its (single) tile
  const TileDesc* tiles;
  const uint64_t* woff;
  const uint32_t* wlen;
  const uint32_t* pool;
  const uint8_t* luts;
  const uint16_t* dtab;
  const uint64_t* scr_base;
  uint32_t* scratch;
  TileOut* tout;
  unsigned long long* ev_count;
  unsigned int* err;
  const ChunkWin* cwin;
  uint32_t lo_rel, hi_rel, first;
};

// LDS bytes of one fused-level workgroup for an input capacity of `cap`
// words: inputs (cap + 128: one sentinel per input range), stacks (cap + 64,
// when they live in LDS), truth table (256 B), delay table (2 KiB).
inline size_t level_lds_bytes(uint32_t cap, bool stacks_in_lds) {
  return ((uint64_t)cap + 128 + (stacks_in_lds ? cap + 64 : 0)) * 4 + 256 + 2 * dt_size(4);
}

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
