// Fused per-level gate evaluation (gfx950): walk, join, compact.
//
// Work unit: a tile = 64 consecutive time segments of one gate, one segment
// per lane of a wavefront.  Segment c of a gate covers the events
// [c*stride, (c+1)*stride) of the gate's longest input (the "leader"), so
// segment boundaries are exact event times.
//
// k_level (walk), one tile at a time per wavefront, tiles independent and
// handed out by ticket:
//   1. the tile's input ranges (one event before the tile for the starting
//      values, a D_max look-ahead halo after it), its delay table and its
//      truth table are copied into LDS with direct-to-LDS loads, each range
//      followed by an INF sentinel (tiles too long for LDS read HBM instead
//      and keep their stacks in the overflow buffer);
//   2. every lane merges its segment's input events in time order (branch
//      free), evaluates the 4-state truth table (packed in registers for
//      fan-in <= 3), classifies the output edge, picks the SDF delay with one
//      lookup in the gate's delay table (AddDelay folded on the host) and
//      applies the preemption rule on a stack (LDS, or a per-workgroup HBM
//      scratch), counting value changes as it goes; events in the halo only
//      pre-empt;
//   3. a wavefront scan (DPP) combines the lanes' kept lists (dedup of equal
//      adjacent values across lanes is an associative monoid), the lanes
//      store the tile's entries into the level scratch and lane 0 writes the
//      tile summary (count, first/last value and time, SAIF durations).
//   The next tile's loads are issued before the current tile is stored.
// k_gate_carry: per gate, joins its tiles in order (a tile whose first entry
//   repeats the value before it drops that entry), places them, and finishes
//   the gate's SAIF record.
// k_compact: copies each tile's joined entries into the gate's dense
//   waveform in the pool (coalesced), so the pool holds exactly the outputs,
//   and counts SAIF toggles inside tiles when asked.
//
// Event rules: src/simulate.cpp:919-1201 (Calculate, AddDelay, CheckDelay),
// src/simulate.cpp:1365-1506 (Process, GetValueAt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include <type_traits>

#include "device.h"

namespace gls {

constexpr int kLanes = 64;

// ERR_ORDER: a bound that holds for time-ordered inputs was hit (a lane's
// stack region or a tile's output region would have been overrun); the
// write was dropped.
enum : unsigned { ERR_OVERFLOW = 2u, ERR_RANGE = 4u, ERR_CAPACITY = 8u, ERR_POOL = 16u,
                  ERR_ORDER = 32u };

__device__ __forceinline__ uint32_t edge_cls(uint32_t pre, uint32_t cur) {
  // CheckDelay (src/simulate.cpp:1178-1201); codes 0=0 1=X 2=Z 3=1
  const uint64_t T = 0x0222ull << 48 | 0x1042ull << 32 | 0x1302ull << 16 | 0x1110ull;
  return (uint32_t)(T >> ((pre * 4 + cur) * 4)) & 0xF;
}

// Aggregate of a run of kept candidates (for dedup + SAIF carry):
// cnt = entries after dedup, fv/lv = first/last value, lt = last time.
struct Agg {
  uint32_t cnt, lt;
  uint32_t fv, lv, has;
};

__device__ __forceinline__ Agg agg_combine(const Agg& a, const Agg& b) {
  if (!a.has) return b;
  if (!b.has) return a;
  Agg r;
  r.cnt = a.cnt + b.cnt - (a.lv == b.fv ? 1u : 0u);
  r.fv = a.fv;
  r.lv = b.lv;
  r.lt = b.lt;
  r.has = 1;
  return r;
}

// Cross-lane moves with DPP (no LDS round trip).  Lanes without a source
// lane (outside the row, or rows left out by ROW_MASK) receive 0, which is
// the identity of the sums below and, for all three words, the empty Agg.
// Control codes: row_shr:d 0x110+d, row_shl:d 0x100+d, wave_shr:1 0x138,
// row_bcast:15 0x142, row_bcast:31 0x143.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, false);
}

__device__ __forceinline__ uint32_t agg_flags(const Agg& a) {
  return a.fv | (a.lv << 2) | (a.has << 4);
}
__device__ __forceinline__ Agg agg_from(uint32_t cnt, uint32_t lt, uint32_t f) {
  Agg r;
  r.cnt = cnt;
  r.lt = lt;
  r.fv = f & 3u;
  r.lv = (f >> 2) & 3u;
  r.has = (f >> 4) & 1u;
  return r;
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ Agg agg_dpp(const Agg& a) {
  return agg_from(dpp<CTRL, ROW_MASK>(a.cnt), dpp<CTRL, ROW_MASK>(a.lt),
                  dpp<CTRL, ROW_MASK>(agg_flags(a)));
}

__device__ __forceinline__ Agg agg_readlane(const Agg& a, int lane) {
  return agg_from(__builtin_amdgcn_readlane(a.cnt, lane), __builtin_amdgcn_readlane(a.lt, lane),
                  __builtin_amdgcn_readlane(agg_flags(a), lane));
}

// Inclusive scan of the lanes' aggregates in lane order: within rows of 16
// by row shifts, then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then lane 31's prefix into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ Agg agg_wave_scan(Agg x) {
  x = agg_combine(agg_dpp<0x111>(x), x);
  x = agg_combine(agg_dpp<0x112>(x), x);
  x = agg_combine(agg_dpp<0x114>(x), x);
  x = agg_combine(agg_dpp<0x118>(x), x);
  x = agg_combine(agg_dpp<0x142, 0xA>(x), x);
  x = agg_combine(agg_dpp<0x143, 0xC>(x), x);
  return x;
}

// Sum over the wavefront (uniform result).
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
  x += dpp<0x111>(x);
  x += dpp<0x112>(x);
  x += dpp<0x114>(x);
  x += dpp<0x118>(x);
  x += dpp<0x142, 0xA>(x);
  x += dpp<0x143, 0xC>(x);
  return __builtin_amdgcn_readlane(x, 63);
}

__device__ __forceinline__ void saif_acc(unsigned long long* acc, uint32_t v, unsigned long long a,
                                         unsigned long long b, unsigned long long start,
                                         unsigned long long end) {
  unsigned long long lo = a > start ? a : start, hi = b < end ? b : end;
  const unsigned long long len = hi > lo ? hi - lo : 0ull;
  acc[0] += v == 0 ? len : 0ull;
  acc[1] += v == 3 ? len : 0ull;
  acc[2] += (v == 1 || v == 2) ? len : 0ull;
}

template <typename P>
__device__ __forceinline__ uint32_t lb_words(P a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

typedef __attribute__((address_space(3))) uint32_t lds_u32;

// Pointer to a tile's input ranges: LDS (staged) or HBM (too long for LDS).
// The LDS form is a 32-bit local pointer, so its reads are ds_read.
template <bool LDS> struct InPtr { typedef const uint32_t* T; };
template <> struct InPtr<true> { typedef const lds_u32* T; };

// Uniform (per tile) data of one gate.
template <int K, bool LDS>
struct TileCtx {
  typename InPtr<LDS>::T src[K];  // input ranges, index relative to B
  uint32_t B[K], E[K], L[K];
  uint32_t nseg, stride, lead, dmax;
  uint32_t span;            // the tile's own input words (out_cap - 1)
  uint32_t lut[4];          // packed truth table, fan-in <= 3
};

template <int K, bool LDS>
__device__ __forceinline__ uint32_t lut_eval(const TileCtx<K, LDS>& X, const uint8_t* s_lut,
                                             uint32_t idx) {
  // fan-in <= 2: 16 entries of 2 bits in one register; else the staged table
  if (K <= 2) return __builtin_amdgcn_ubfe(X.lut[0], 2 * idx, 2);
  return s_lut[idx];
}

// Delay-table key bits of one event (dt_index): per input j, bit 2j = it has
// an event (from `ev`, already at even positions) and bit 2j+1 = it fell
// (CheckDelay == 2: to 0 from X/Z/1, or from 1 to X/Z) between the packed
// input values p and c (2 bits per input).  At bit 2j+1 of c, c << 1, p and
// p << 1 sit the input's c1, c0, p1, p0.
__device__ __forceinline__ uint32_t event_bits(uint32_t p, uint32_t c, uint32_t ev) {
  const uint32_t cs = c << 1, ps = p << 1;
  const uint32_t to0 = ~c & ~cs & (p | ps);        // now 0, was not 0
  const uint32_t from1 = (p & ps) & (c ^ cs);      // was 1, now X or Z
  return ((to0 | from1) & 0xAAu) | ev;
}

// What one lane's walk over its segment leaves behind: its kept candidates
// (st[0..sp), in time order), the last of them, the number of value changes
// among them, and the number of gate events it owned.  SP: pointer to the
// lane's region (LDS or HBM).
template <typename SP>
struct Lane {
  SP st;
  uint32_t sp, top, cnt, nown;
};

typedef __attribute__((address_space(3))) void lds_void;

// One 4-byte direct-to-LDS load per lane: lane i's word lands at l + i.
__device__ __forceinline__ void dma_word(const uint32_t* g, uint32_t* l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)l, 4, 0, 0);
}

// One 16-byte direct-to-LDS load per lane (gfx950): lane i's four words
// land at l + 4 i.  g is 16-byte aligned.
__device__ __forceinline__ void dma_x4(const uint32_t* g, uint32_t* l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)l, 16, 0, 0);
}

// Wait until this wavefront has no vector-memory access in flight (direct-to-
// LDS loads included): s_waitcnt vmcnt(0).
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ uint32_t range_start(uint32_t L) { return L > 0 ? L - 1 : 0u; }

__device__ __forceinline__ uint32_t tile_words(const TileDesc& D) {
  // loops over the fan-in run to 4 with constant indices so a descriptor
  // held in registers is never indexed at run time
  uint32_t total = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j)
    if (j < D.k) total += D.E[j] - range_start(D.L[j]);
  return total;
}

// LDS layout of a tile's input ranges: range j (n_j words from pool word
// G_j) starts at P_j = R_j + G_j mod 4 with R_j a multiple of 4 words, so it
// is copied with 16-byte loads from the aligned word G_j - G_j mod 4;
// R_{j+1} = the multiple of 4 at or above P_j + n_j + 1 keeps the sentinel
// word after each range.  The layout needs at most sum(n_j) + 7 K words.
__device__ __forceinline__ void range_layout(const TileDesc& D, uint32_t P[4]) {
  uint32_t R = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    const uint32_t n = D.E[j] - range_start(D.L[j]);
    P[j] = R + ((uint32_t)D.in_off[j] & 3u);
    if (j < D.k) R = (P[j] + n + 4) & ~3u;
  }
}

// Issue the loads of tile D into LDS: its input ranges when they fit (laid
// out by range_layout; the 16-byte copies may also bring up to 3 words
// before and after a range, inside its slot), its delay table and (fan-in
// 3 and 4) its truth table.  Completion: wait_vm().
__device__ __forceinline__ void stage_tile(const LevelArgs& A, const TileDesc& D, bool fits,
                                           uint32_t* s_in, uint32_t* s_lut, uint32_t* s_dt) {
  const uint32_t lane = threadIdx.x;
  if (D.k > kTileK) return;  // wide gate: walked by k_wide
  if (fits) {
    uint32_t P[4];
    range_layout(D, P);
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      if (j >= D.k) break;
      const uint32_t n = D.E[j] - range_start(D.L[j]);
      const uint32_t a = (uint32_t)D.in_off[j] & 3u;
      const uint32_t m = (a + n + 3) & ~3u;  // words copied, a multiple of 4
      const uint32_t* src = A.pool + (D.in_off[j] - a);
      uint32_t* l = s_in + (P[j] - a);
      for (uint32_t i = 0; i < m; i += 4 * kLanes)
        if (i + 4 * lane < m) dma_x4(src + i + 4 * lane, l + i);
    }
  }
  const uint32_t nw = dt_size(D.k) / 2;
  const uint32_t* dsrc = (const uint32_t*)(A.dtab + D.dt_off);
  for (uint32_t i = 0; i < nw; i += kLanes)
    if (i + lane < nw) dma_word(dsrc + i + lane, s_dt + i);
  if (D.k >= 3) dma_word((const uint32_t*)(A.luts + D.lut * 256u) + lane, s_lut);
}

// One lane's walk over its time segment, in two passes over the lane's
// region of the stack buffer.
//
// Pass 1 merges the inputs in time order and evaluates every event without
// branches: 4-state truth table, edge class (CheckDelay,
// src/simulate.cpp:1178-1201), delay (one lookup in the gate's table, which
// folds AddDelay, src/simulate.cpp:1012-1176) and candidate output time
// o = t + d; each candidate of an owned event is appended as (o << 2 | f).
// Events in the halo (the next segment's first D_max) only record the
// smallest output time they make.  Look-back events, owned events and halo
// events come in that order in time, so each has a loop of its own, tested
// at the bottom: a step carries only its own kind's bookkeeping.
//
// Pass 2 applies the preemption rule of src/simulate.cpp:990-1009 in its
// closed form: a candidate is kept iff its output time is strictly below
// that of every later candidate (those of the lane and of the halo).  It runs
// from the last candidate to the first, keeping the running minimum,
// compacting kept candidates towards the end of the region (in time order)
// and counting value changes among them.  Chunks: candidates below lo_rel
// (made by look-back events) are never kept but still pre-empt; candidates at
// or above hi_rel belong to the next chunk and are not kept (they pre-empt
// only entries that are not kept either).  Only events at or after lo_rel
// count as this chunk's gate events.
//
// Index 0 of every waveform is the net's state entry (its value before the
// chunk's look-back window), so segment 0 starts at index 1 and takes its
// starting values from index 0.
//
// stk_words: size of the tile's stack buffer at stk_base.  A lane's region
// runs from its offset to the next lane's and holds its owned events + 1
// words, enough for every candidate (and, in segment 0 of chunk 0, the
// initial (0, X) entry).  Regions in HBM are guarded anyway: a write that
// would leave the buffer (inputs out of order) lands on the buffer's last
// word instead and sets bad.  LDS regions are not guarded (an LDS access
// outside the workgroup's allocation is discarded by the hardware; all
// inputs are time-ordered by construction).
template <int K, bool LDS, typename SP>
__device__ __forceinline__ Lane<SP> lane_walk(const TileCtx<K, LDS>& X, uint32_t ti,
                                              const uint8_t* s_lut, const uint16_t* s_dt,
                                              SP stk_base, uint32_t stk_words, uint32_t& bad,
                                              uint32_t lo_rel, uint32_t hi_rel, bool first) {
  const int lane = threadIdx.x;
  const uint32_t c0 = ti * kLanes;                          // the tile's first segment
  const uint32_t m = min((uint32_t)kLanes, X.nseg - c0);   // its segments (>= 1)
  const uint32_t kX = 1u;
  const uint32_t INF = 0xFFFFFFFFu;
  uint32_t r[K], h[K], bound[K];
  typename InPtr<LDS>::T srcL = X.src[0];
  uint32_t BL = X.B[0], EL = X.E[0];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    bound[j] = X.E[j] - X.B[j];
    if ((uint32_t)j == X.lead) {
      srcL = X.src[j];
      BL = X.B[j];
      EL = X.E[j];
    }
  }
  constexpr bool kGuard = !std::is_same<SP, lds_u32*>::value;
  Lane<SP> ln;
  ln.st = stk_base;
  ln.sp = ln.top = ln.cnt = ln.nown = 0;
  // Lane bounds.  The tile's nominal segments (every `stride` leader events)
  // carry uneven numbers of input events; each lane instead takes an equal
  // share of the tile's input words.  Lane q < m first finds the word rank
  // of nominal boundary q; lane c then places its start at the leader event
  // interpolated between the nominal boundaries around rank c * span / 64.
  // Lane 0 keeps the tile's start; segment 0 of the gate starts at index 1
  // (index 0 is the state entry) at time 0.
  const bool gate_start = c0 == 0 && lane == 0;
  uint32_t rq = 0xFFFFFFFFu;  // rank of nominal boundary `lane` (lanes < m)
  if ((uint32_t)lane < m) {
    const uint32_t pq = (c0 + lane) * X.stride;
    if (c0 + lane == 0) {
      rq = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) rq += 1u - X.L[j];
    } else {
      const uint32_t tq = srcL[pq - BL] >> 2;
      rq = 0;
#pragma unroll
      for (int j = 0; j < K; ++j)
        rq += ((uint32_t)j == X.lead ? pq : X.B[j] + lb_words(X.src[j], bound[j], tq << 2)) - X.L[j];
    }
  }
  const uint32_t r0 = __shfl(rq, 0);
  const uint32_t target = r0 + (uint32_t)(((uint64_t)(X.span - r0) * (uint32_t)lane) >> 6);
  uint32_t q = 0;  // last nominal boundary at or below the target
#pragma unroll
  for (uint32_t step = 32; step; step >>= 1) {
    const uint32_t cand = q + step;
    const uint32_t rc = __shfl(rq, (int)(cand & 63u));
    q = (cand < m && rc <= target) ? cand : q;
  }
  const uint32_t r_lo = __shfl(rq, (int)q);
  const uint32_t r_nx = __shfl(rq, (int)min(q + 1, 63u));
  const uint32_t r_hi = q + 1 < m ? r_nx : X.span;
  const uint32_t p_lo = (c0 + q) * X.stride;
  const uint32_t p_hi = min((c0 + q + 1) * X.stride, EL);
  uint32_t pi = c0 * X.stride;  // lane 0: the tile's start
  if (lane > 0) {
    pi = p_lo;
    if (r_hi > r_lo && p_hi > p_lo) {
      pi = p_lo + (uint32_t)(((uint64_t)(target - r_lo) * (p_hi - p_lo)) / (r_hi - r_lo));
      pi = min(pi, p_hi - 1);
    }
  }
  pi = max(pi, 1u);
  const uint32_t tau_c = gate_start ? 0u : srcL[pi - BL] >> 2;
  // the segment ends where the next lane's starts; the tile's last lane ends
  // at the next tile's first leader event (none after the gate's last tile)
  const uint32_t tau_next = __shfl_down(tau_c, 1);
  uint32_t tau_n = 0x3FFFFFFFu, halo_end = 0x3FFFFFFFu;
  if (lane < kLanes - 1) {
    tau_n = tau_next;
  } else if (c0 + m < X.nseg) {
    tau_n = srcL[(c0 + m) * X.stride - BL] >> 2;
  }
  if (tau_n != 0x3FFFFFFFu) halo_end = tau_n + X.dmax;
  uint32_t soff = lane, idx = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    // pj >= 1: index 0 is the state entry, and every event time is >= 1
    uint32_t pj;
    if (gate_start) pj = 1;
    else if ((uint32_t)j == X.lead) pj = pi;
    else pj = X.B[j] + lb_words(X.src[j], bound[j], tau_c << 2);
    soff += pj - X.L[j];
    r[j] = pj - X.B[j];
    idx |= (X.src[j][r[j] - 1] & 3u) << (2 * j);
    // LDS ranges end with an INF sentinel
    h[j] = (LDS || r[j] < bound[j]) ? X.src[j][r[j]] : INF;
  }
  // guarded regions: the lane's words stay below the buffer end (a lane
  // whose offset is already past it keeps the buffer's last word)
  const uint32_t s0 = kGuard ? min(soff, stk_words - 1) : soff;
  const uint32_t stk_lim = stk_words - s0;
  SP st = stk_base + s0;
  const bool init = gate_start && first;
  // f before the first event: X at t = 0 (src/simulate.cpp:933), else the
  // function of the starting values
  uint32_t fprev = init ? kX : lut_eval<K>(X, s_lut, idx);
  uint32_t nc = 0, nown = 0, hmin = INF;
  if (init) {  // every output starts with (0, X) (src/simulate.cpp:882-887); t = 0 is rel 1
    st[0] = (1u << 2) | kX;
    nc = 1;
  }
  // pass 1: merge and evaluate.  One step takes every input event at the
  // next event time t (w: its smallest word) and returns the gate's value
  // after it, its output time o and whether it is a candidate (a change).
  auto step = [&](uint32_t w, uint32_t t, uint32_t& o, bool& cand) -> uint32_t {
    const uint32_t lim = w | 3u;
    uint32_t ev = 0, nidx = idx;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const bool take = h[j] <= lim;
      ev |= take ? (1u << (2 * j)) : 0u;
      nidx = take ? ((nidx & ~(3u << (2 * j))) | ((h[j] & 3u) << (2 * j))) : nidx;
      r[j] += take ? 1u : 0u;
    }
#pragma unroll
    for (int j = 0; j < K; ++j) h[j] = (LDS || r[j] < bound[j]) ? X.src[j][r[j]] : INF;
    const uint32_t f = lut_eval<K>(X, s_lut, nidx);
    const uint32_t sc = edge_cls(fprev, f);
    const uint32_t key = (((sc - 1u) & 3u) << (2 * K)) | event_bits(idx, nidx, ev);
    idx = nidx;
    fprev = f;
    o = t + s_dt[key];  // sc = 0 reads a valid entry that is not used
    cand = sc != 0u;
    return f;
  };
  auto next_word = [&]() -> uint32_t {
    uint32_t w = h[0];
#pragma unroll
    for (int j = 1; j < K; ++j) w = min(w, h[j]);
    return w;
  };
  // the lane's own events (t < tau_n): candidates go to the stack.  The
  // halo comes after them in time, so it gets a loop of its own and the
  // owned steps carry none of its bookkeeping.
  // (Both loops test at the bottom: one compare a step decides both the
  // lanes that go on and the loop's exit.)
  auto push = [&](uint32_t o, uint32_t f, bool cand) {
    const bool room = !kGuard || nc < stk_lim;
    bad |= room ? 0u : 1u;
    st[room ? nc : stk_lim - 1] = (o << 2) | f;
    nc += (cand && room) ? 1u : 0u;
  };
  uint32_t w = next_word();
  uint32_t t = w >> 2;
  // look-back events (t < lo_rel, chunks after the first) come first: they
  // make candidates but are not this chunk's gate events
  const uint32_t t_own = min(lo_rel, tau_n);
  while (t < t_own) {
    uint32_t o;
    bool cand;
    const uint32_t f = step(w, t, o, cand);
    push(o, f, cand);
    w = next_word();
    t = w >> 2;
  }
  while (t < tau_n) {  // the end of every input (INF) is never below tau_n
    uint32_t o;
    bool cand;
    const uint32_t f = step(w, t, o, cand);
    ++nown;
    push(o, f, cand);
    w = next_word();
    t = w >> 2;
  }
  // the halo (tau_n <= t < halo_end): only the smallest output time of its
  // candidates is kept
  while (t < halo_end) {
    uint32_t o;
    bool cand;
    step(w, t, o, cand);
    hmin = cand ? min(hmin, o) : hmin;
    w = next_word();
    t = w >> 2;
  }
  // pass 2: keep the candidates strictly earlier than every later one,
  // compacted (in time order) to the end of the region
  uint32_t runmin = min(hmin, hi_rel), j = nc, cnt = 0, lastv = 4u;
  for (uint32_t i = nc; i-- > 0;) {
    const uint32_t e = st[i];
    const uint32_t o = e >> 2;
    const bool keep = o >= lo_rel && o < runmin;
    runmin = min(runmin, o);
    st[keep ? j - 1 : i] = e;
    j -= keep ? 1u : 0u;
    const uint32_t v = e & 3u;
    cnt += (keep && v != lastv) ? 1u : 0u;
    lastv = keep ? v : lastv;
  }
  ln.st = st + j;
  ln.sp = nc - j;
  // the last kept candidate: the first one kept going backwards, now at the
  // region's last slot
  ln.top = j < nc ? st[nc - 1] : 0u;
  ln.cnt = cnt;
  ln.nown = nown;
  return ln;
}

// Tile context and walk for fan-in K, then `finish` with the lanes' stacks.
// Tiles that fit LDS read their inputs there (LDS pointers) and keep their
// stacks in LDS or in this workgroup's slice of the stack scratch; tiles that
// do not fit read their inputs from HBM and keep their stacks in the
// overflow buffer.
template <int K, typename F>
__device__ __forceinline__ void walk_k(const LevelArgs& A, const TileDesc& D, bool fits,
                                       uint32_t* s_in, uint32_t* s_st, const uint8_t* s_lut,
                                       const uint16_t* s_dt, F&& finish) {
  const int lane = threadIdx.x;
  uint32_t B[K], E[K], L[K], pos[4], total = 0;
  range_layout(D, pos);
#pragma unroll
  for (int j = 0; j < K; ++j) {
    L[j] = D.L[j];
    E[j] = D.E[j];
    B[j] = range_start(L[j]);
    total += E[j] - B[j];
  }
  auto fill = [&](auto& X) {
    X.nseg = D.nseg;
    X.stride = D.stride;
    X.lead = D.lead;
    X.dmax = D.dmax;
    X.span = D.out_cap - 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) X.lut[q] = D.lutw[q];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      X.L[j] = L[j];
      X.E[j] = E[j];
      X.B[j] = B[j];
    }
  };
  if (fits) {
    TileCtx<K, true> X;
    fill(X);
    const lds_u32* in = (const lds_u32*)s_in;
#pragma unroll
    for (int j = 0; j < K; ++j) X.src[j] = in + pos[j];
    uint32_t bad = 0;
    if (A.stk) {
      finish(lane_walk<K, true>(X, D.ti, s_lut, s_dt, A.stk + (uint64_t)blockIdx.x * (A.cap + kLanes),
                                A.cap + kLanes, bad, A.lo_rel, A.hi_rel, A.first != 0));
    } else {
      finish(lane_walk<K, true>(X, D.ti, s_lut, s_dt, (lds_u32*)s_st, A.cap + kLanes, bad, A.lo_rel,
                                A.hi_rel, A.first != 0));
    }
    if (bad) atomicOr(A.err, ERR_ORDER);
    return;
  }
  TileCtx<K, false> X;
  fill(X);
#pragma unroll
  for (int j = 0; j < K; ++j) X.src[j] = A.pool + D.in_off[j];
  unsigned long long base = 0;
  const unsigned long long need = (unsigned long long)total + kLanes;
  if (lane == 0) base = atomicAdd(A.ovf_used, need);
  base = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
         (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)base);  // lane 0's value
  if (base + need > A.ovf_cap) {
    // no room: the tile's output is left empty (so the level's waveforms stay
    // in time order) and the run reports the error
    if (lane == 0) atomicOr(A.err, ERR_OVERFLOW);
    Lane<uint32_t*> ln;
    ln.st = A.ovf;
    ln.sp = ln.top = ln.cnt = ln.nown = 0;
    finish(ln);
    return;
  }
  uint32_t bad = 0;
  finish(lane_walk<K, false>(X, D.ti, s_lut, s_dt, A.ovf + base, (uint32_t)need, bad, A.lo_rel,
                             A.hi_rel, A.first != 0));
  if (bad) atomicOr(A.err, ERR_ORDER);
}

// The tile's output inside the tile: wavefront scan of the lane aggregates
// (dedup across lane boundaries is an associative monoid).
struct TileSum {
  Agg carry;         // everything before this lane's first entry, inside the tile
  Agg tile;          // the whole tile
  uint32_t e_first;  // the lane's first stack entry
  uint32_t ft;       // time of the tile's first entry
};

template <typename SP>
__device__ __forceinline__ TileSum combine_tile(const Lane<SP>& ln) {
  const uint32_t sp = ln.sp;
  const uint32_t e_first = sp ? ln.st[0] : 0u;
  Agg mine;
  mine.cnt = ln.cnt;
  mine.has = sp > 0 ? 1u : 0u;
  mine.fv = e_first & 3u;
  mine.lv = ln.top & 3u;
  mine.lt = ln.top >> 2;
  const Agg inc = agg_wave_scan(mine);
  TileSum r;
  r.carry = agg_dpp<0x138>(inc);  // lane - 1 (empty for lane 0)
  r.tile = agg_readlane(inc, 63);
  r.e_first = e_first;
  const uint64_t has = __ballot(sp > 0);
  r.ft = has ? (uint32_t)__builtin_amdgcn_readlane(e_first, __ffsll((long long)has) - 1) >> 2 : 0u;
  return r;
}

// Store the tile's entries (deduplicated inside the tile) at dst, sum the
// SAIF durations between them (32 bits per lane: times < 2^30; the lanes'
// intervals are disjoint pieces of the tile's span, so the sums fit too) and
// write the tile summary.  out_cap: words of the tile's output region (never
// reached for time-ordered inputs; stores beyond it are dropped and the
// count is clamped, with ERR_ORDER).
template <typename SP>
__device__ __forceinline__ void emit_tile(const LevelArgs& A, uint64_t tk, const Lane<SP>& ln,
                                          const TileSum& T, uint32_t* dst, uint32_t out_cap,
                                          unsigned long long w_start, unsigned long long w_end,
                                          unsigned long long& wave_ev) {
  const int lane = threadIdx.x;
  const uint32_t sp = ln.sp;
  const SP st = ln.st;
  const Agg& carry = T.carry;
  const uint32_t s32 = (uint32_t)min(w_start, 0x7FFFFFFFull);
  const uint32_t e32 = (uint32_t)min(w_end, 0x7FFFFFFFull);
  uint32_t pvv = carry.has ? carry.lv : 4u;
  uint32_t pct = max(min(carry.has ? carry.lt : (T.e_first >> 2), e32), s32);
  const uint32_t ct0 = pct;
  uint32_t a0 = 0, a1 = 0;
  uint32_t w = carry.has ? carry.cnt : 0u;  // lane's first slot in the tile output
  // four stack reads in flight per step; indices past the last entry read
  // the last entry again, which adds nothing (same value: no store; same
  // time: no duration), so every step takes all four
  for (uint32_t i = 0; i < sp; i += 4) {
    uint32_t e4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) e4[q] = st[min(i + q, sp - 1)];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t e = e4[q];
      const uint32_t v = e & 3u;
      const uint32_t ct = max(min(e >> 2, e32), s32);
      const uint32_t len = ct - pct;
      a0 += pvv == 0u ? len : 0u;
      a1 += pvv == 3u ? len : 0u;
      if (v != pvv) {
        // past out_cap (inputs out of time order: the run fails with
        // ERR_ORDER) stores land on the region's last word
        dst[min(w, out_cap - 1u)] = e;
        ++w;
      }
      pvv = v;
      pct = ct;
    }
  }
  const uint32_t t0 = wave_sum(a0), t1 = wave_sum(a1), tx = wave_sum(pct - ct0 - a0 - a1);
  const uint32_t ev = wave_sum(ln.nown);
  if (lane == 0) {
    TileOut o;
    o.cnt = min(T.tile.cnt, out_cap);
    if (T.tile.cnt > out_cap) atomicOr(A.err, ERR_ORDER);
    o.ft = T.ft;
    o.lt = T.tile.lt;
    o.flags = agg_flags(T.tile);
    o.t0 = t0;
    o.t1 = t1;
    o.tx = tx;
    o.ev = ev;
    A.tout[tk] = o;
    wave_ev += ev;
  }
}

// A tile descriptor is read with one coalesced load (lane i < 32 holds
// dword i) and broadcast to every lane with readlane.  The fields then live
// in registers: no per-field loads (whose waits, s_waitcnt vmcnt, would also
// wait for the staging DMA in flight) on the walk's path.
__device__ __forceinline__ uint32_t desc_word(const TileDesc* p) {
  const uint32_t lane = threadIdx.x;
  return lane < 32 ? ((const uint32_t*)p)[lane] : 0u;
}

__device__ __forceinline__ TileDesc desc_of(uint32_t w) {
  uint32_t f[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) f[i] = __builtin_amdgcn_readlane(w, i);
  TileDesc D;
  __builtin_memcpy(&D, f, sizeof D);
  return D;
}

// Scratch word of a tile's output region: its gate's region, then the
// tile's first own event in the gate's merged input order plus one slack
// word per earlier tile (a tile keeps at most its own events + 1 entries).
__device__ __forceinline__ uint64_t tile_scratch(const uint64_t* scr_base, const TileDesc& D) {
  uint64_t p = scr_base[D.g] + D.ti;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j)
    if (j < D.k) p += D.L[j];
  return p;
}

// Persistent, software-pipelined walk kernel.  Tiles are independent (no
// tile waits for another): workgroup b starts with tile b and then draws
// tickets for the next one; the next tile's inputs are staged into LDS while
// the current tile is stored.  A static round-robin schedule left the
// workgroups that drew many tiles too long for LDS as a long tail (14%
// slower on the 200k-cell workload).
__global__ __launch_bounds__(kLanes) void k_level(LevelArgs A) {
  // dynamic LDS: inputs (cap + 128 words), stacks (cap + 64, unless they are
  // in HBM), truth table (64 words), delay table (512 words)
  extern __shared__ uint32_t s_dyn[];
  uint32_t* s_in = s_dyn;
  uint32_t* s_st = A.stk ? nullptr : s_dyn + A.cap + 2 * kLanes;
  uint32_t* s_lut = A.stk ? s_dyn + A.cap + 2 * kLanes : s_st + A.cap + kLanes;
  uint32_t* s_dt = s_lut + 64;
  const uint8_t* lut8 = (const uint8_t*)s_lut;
  const uint16_t* dt16 = (const uint16_t*)s_dt;
  const int lane = threadIdx.x;
  unsigned long long wave_ev = 0;  // gate events of this wavefront (lane 0)
  unsigned long long wave_tiles = 0, wave_hbm = 0;
  const uint64_t ntiles = *A.ntiles;  // uniform: one scalar load
  // Tiles are handed out in batches of B consecutive tiles (the first batch
  // of workgroup b is b's, later ones come from a ticket counter): one atomic
  // per batch instead of per tile (a single counter serves only so many
  // atomics per second), and consecutive tiles of one gate share its tables
  // in L2.  Small levels use B = 1 so every workgroup gets a tile.
  const uint64_t B = max(1ull, min(8ull, (unsigned long long)(ntiles / (16ull * gridDim.x))));
  uint64_t tk = (uint64_t)blockIdx.x * B;
  uint64_t bend = min(tk + B, ntiles);  // end of the current batch
  bool fits = false;
  uint32_t dw = 0;  // descriptor of tile tk, one dword per lane
  if (tk < ntiles) {
    dw = desc_word(A.tiles + tk);
    const TileDesc D0 = desc_of(dw);
    fits = tile_words(D0) <= A.cap;
    stage_tile(A, D0, fits, s_in, s_lut, s_dt);
  }
  while (tk < ntiles) {
    const TileDesc D = desc_of(dw);
    const bool wide = D.k > kTileK;  // uniform: the tile is k_wide's
    wait_vm();
    if (fits && !wide && lane == 0) {  // INF sentinel after every input range
      // one lane, constant indices: the descriptor stays in registers (an
      // index chosen by lane would put it in per-lane scratch memory)
      uint32_t P[4];
      range_layout(D, P);
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j)
        if (j < D.k) s_in[P[j] + D.E[j] - range_start(D.L[j])] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // the emit's scalars, loaded now so the walk covers their latency
    uint32_t* dst = A.scratch + tile_scratch(A.scr_base, D);
    const unsigned long long w_start = A.cwin[D.job].lo, w_end = A.cwin[D.job].hi;
    // the next tile: the next of this batch, or the first of a batch drawn
    // from the ticket counter (dynamic schedule: workgroups that drew slow
    // tiles, e.g. ones too long for LDS, take fewer, so the level's tail stays
    // short); its descriptor load completes during the walk
    uint64_t tn = tk + 1, bend_n = bend;
    if (tn >= bend) {
      uint64_t t0 = 0;
      if (lane == 0) t0 = (uint64_t)gridDim.x * B + atomicAdd(A.ticket, (unsigned long long)B);
      tn = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(t0 >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)t0);
      bend_n = min(tn + B, ntiles);
    }
    const uint32_t dw_n = tn < ntiles ? desc_word(A.tiles + tn) : 0u;
    bool fits_n = false;
    auto finish = [&](const auto& ln) {
      __syncthreads();  // inputs, truth and delay tables no longer read
      const TileSum T = combine_tile(ln);
      if (A.prefetch && tn < ntiles) {  // stage the next tile while this one is stored
        const TileDesc Dn = desc_of(dw_n);
        fits_n = tile_words(Dn) <= A.cap;
        stage_tile(A, Dn, fits_n, s_in, s_lut, s_dt);
      }
      emit_tile(A, tk, ln, T, dst, D.out_cap, w_start, w_end, wave_ev);
    };
    switch (wide ? 0u : D.k) {
      case 0: break;
      case 1: walk_k<1>(A, D, fits, s_in, s_st, lut8, dt16, finish); break;
      case 2: walk_k<2>(A, D, fits, s_in, s_st, lut8, dt16, finish); break;
      case 3: walk_k<3>(A, D, fits, s_in, s_st, lut8, dt16, finish); break;
      default: walk_k<4>(A, D, fits, s_in, s_st, lut8, dt16, finish); break;
    }
    if ((wide || !A.prefetch) && tn < ntiles) {
      __syncthreads();
      const TileDesc Dn = desc_of(dw_n);
      fits_n = tile_words(Dn) <= A.cap;
      stage_tile(A, Dn, fits_n, s_in, s_lut, s_dt);
    }
    __syncthreads();
    ++wave_tiles;
    wave_hbm += fits ? 0u : 1u;
    tk = tn;
    bend = bend_n;
    fits = fits_n;
    dw = dw_n;
  }
  if (lane == 0 && wave_ev) atomicAdd(A.ev_count, wave_ev);
  if (lane == 0 && wave_tiles) {
    atomicAdd(A.tiles_walked, wave_tiles);
    if (wave_hbm) atomicAdd(A.tiles_walked + 1, wave_hbm);
  }
}

// Per gate, in tile order: join the tiles' outputs (a tile whose first
// entry repeats the last value before it drops that entry), place each
// tile's output (offset inside the gate, bit 31: first entry dropped), and
// add up the SAIF durations: inside tiles, between tiles, and from the last
// entry to the window end.  One wavefront per gate, 64 tiles per step: each
// lane takes the last value and time of the nearest non-empty tile before
// it (a ballot of the non-empty lanes, then a shuffle; the step's first
// non-empty tile takes them from the previous step), the offsets are an
// exclusive scan of the joined counts, and the sums are reduced at the end.
// In a batch (ev_job set) the gate's events are added to its instance's
// count for the chunk.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) x += __shfl_xor(x, d);
  return x;
}

__global__ void k_gate_carry(const GateRec* __restrict__ gates, uint32_t ng,
                             const uint64_t* __restrict__ tile_base,
                             const uint64_t* __restrict__ ntile, const TileOut* __restrict__ tout,
                             uint32_t* __restrict__ tile_off, uint64_t* __restrict__ gate_cnt,
                             uint32_t* __restrict__ wlen, unsigned long long* __restrict__ saif,
                             const ChunkWin* __restrict__ cwin, const unsigned int* __restrict__ err,
                             TailArgs T, unsigned long long* __restrict__ ev_job) {
  const int lane = threadIdx.x & 63;
  const uint32_t g = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (g >= ng) return;  // uniform over the wavefront
  const GateRec& G = gates[g];
  if (*err & ERR_CAPACITY) {  // the level was skipped: its tiles were never walked
    if (lane == 0) {
      gate_cnt[g] = 0;
      wlen[G.out_net] = 0;
    }
    return;
  }
  const ChunkWin W = cwin[G.job];
  const unsigned long long start = W.lo, end = W.hi;
  const uint64_t t0 = tile_base[g], n = ntile[g];
  // carried from the previous step: count so far, last value and time.  The
  // gate's waveform starts with its tail (state entry and look-back
  // entries); after chunk 0 the first new entry is compared with the tail's
  // last value, which holds from before the chunk (time 0 is clipped to
  // the chunk's SAIF start).
  const uint32_t tl = T.tail_len[G.out_net];
  uint32_t cnt = tl, lv = 0, lt = 0, has = 0;
  if (!T.first && tl) {
    lv = T.tails[T.tail_off[G.out_net] + tl - 1] & 3u;
    has = 1;
  }
  unsigned long long acc[3] = {0, 0, 0}, tc = 0, gev = 0;
  for (uint64_t c = 0; c < n; c += 64) {
    const bool in = c + (uint64_t)lane < n;
    const uint64_t t = t0 + c + (uint64_t)lane;
    TileOut S = {};
    if (in) S = tout[t];
    const bool ne = in && ((S.flags >> 4) & 1u);
    const uint64_t mask = __ballot(ne);
    const uint32_t my_lv = (S.flags >> 2) & 3u, my_lt = S.lt, my_fv = S.flags & 3u;
    // nearest non-empty tile below this lane in the step, if any
    const uint64_t below = mask & ((1ull << lane) - 1ull);
    const int pl = below ? 63 - __clzll((long long)below) : 0;
    const uint32_t s_lv = __shfl(my_lv, pl), s_lt = __shfl(my_lt, pl);
    const uint32_t p_has = below ? 1u : has;
    const uint32_t p_lv = below ? s_lv : lv, p_lt = below ? s_lt : lt;
    uint32_t mine = 0, skip = 0;
    if (ne) {
      skip = (p_has && p_lv == my_fv) ? 1u : 0u;
      mine = S.cnt - skip;
      if (p_has) saif_acc(acc, p_lv, p_lt, S.ft, start, end);
      acc[0] += S.t0;
      acc[1] += S.t1;
      acc[2] += S.tx;
      // the tile's first entry against the previous tile's last value
      tc += (p_has && is_toggle(p_lv, my_fv) && S.ft > W.tc_lo && S.ft <= W.tc_hi) ? 1u : 0u;
      gev += S.ev;
    }
    // inclusive scan of the joined counts over the step
    uint32_t incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    if (in) tile_off[t] = (cnt + incl - mine) | (skip << 31);
    cnt += __shfl(incl, 63);
    if (mask) {  // the step's last non-empty tile carries on
      const int ll = 63 - __clzll((long long)mask);
      lv = __shfl(my_lv, ll);
      lt = __shfl(my_lt, ll);
      has = 1;
    }
  }
  if (has && lane == 0) saif_acc(acc, lv, lt, end > lt ? end : (unsigned long long)lt, start, end);
  const unsigned long long s0 = wave_sum_u64(acc[0]), s1 = wave_sum_u64(acc[1]);
  const unsigned long long s2 = wave_sum_u64(acc[2]), s3 = wave_sum_u64(tc);
  const unsigned long long s4 = ev_job ? wave_sum_u64(gev) : 0ull;
  if (lane == 0) {
    gate_cnt[g] = cnt;
    wlen[G.out_net] = cnt;
    unsigned long long* sa = saif + (uint64_t)kSaifWords * G.out_net;
    sa[0] = s0;
    sa[1] = s1;
    sa[2] = s2;
    sa[3] = s3;
    // a level whose tiles ran out of overflow stacks runs again: count it then
    if (ev_job && s4 && !(*err & ERR_OVERFLOW)) atomicAdd(ev_job + G.job, s4);
  }
}

// Copy every tile's joined output from the scratch into the gate's dense
// waveform in the pool (one wavefront per tile, coalesced); the gate's first
// tile also copies the gate's tail (state and look-back entries, moved to
// this chunk's time base) to the front.  When the level does not fit the
// pool (error flag set by k_pool_guard) nothing is copied and the level's
// waveforms are left empty.
__global__ void k_compact(const TileDesc* __restrict__ tiles, const uint64_t* __restrict__ ntiles,
                          const TileOut* __restrict__ tout, const uint32_t* __restrict__ tile_off,
                          const uint64_t* __restrict__ scr_base, const uint32_t* __restrict__ scratch,
                          const uint64_t* __restrict__ gate_base, uint32_t* __restrict__ pool,
                          uint64_t* __restrict__ woff, uint32_t* __restrict__ wlen,
                          const unsigned int* __restrict__ err, unsigned long long* __restrict__ saif,
                          const ChunkWin* __restrict__ cwin, uint32_t count_tc, TailArgs T) {
  const int lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  const bool failed = (*err & ERR_CAPACITY) != 0;
  const uint64_t nt = *ntiles;
  for (uint64_t t = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; t < nt; t += waves) {
    const TileDesc& D = tiles[t];
    if (failed) {
      if (D.ti == 0 && lane == 0) {
        woff[D.out_net] = 0;
        wlen[D.out_net] = 0;
      }
      continue;
    }
    if (D.ti == 0) {
      if (lane == 0) woff[D.out_net] = gate_base[D.g];
      const uint32_t tl = T.tail_len[D.out_net];
      const uint32_t* ts = T.tails + T.tail_off[D.out_net];
      uint32_t* td = pool + gate_base[D.g];
      for (uint32_t q = lane; q < tl; q += 64) {
        const uint32_t w = ts[q], tt = w >> 2;
        td[q] = ((tt > T.shift ? tt - T.shift : 0u) << 2) | (w & 3u);
      }
    }
    const TileOut S = tout[t];
    if (!((S.flags >> 4) & 1u)) continue;
    const uint32_t off = tile_off[t] & 0x7FFFFFFFu, skip = tile_off[t] >> 31;
    const uint32_t n = S.cnt - skip;
    const uint32_t* src = scratch + tile_scratch(scr_base, D) + skip;
    uint32_t* dst = pool + gate_base[D.g] + off;
    // four loads in flight per lane before their stores: the copy of a tile
    // waits for memory about n / 256 times instead of n / 64
    uint32_t q = lane;
    for (; q + 192 < n; q += 256) {
      const uint32_t a0 = src[q], a1 = src[q + 64], a2 = src[q + 128], a3 = src[q + 192];
      dst[q] = a0;
      dst[q + 64] = a1;
      dst[q + 128] = a2;
      dst[q + 192] = a3;
    }
    for (; q < n; q += 64) dst[q] = src[q];
    if (count_tc) {
      // SAIF TC inside the tile: 0<->1 changes between its consecutive
      // entries (the first entry against the previous tile is k_gate_carry's)
      const uint32_t* e = src - skip;
      const ChunkWin W = cwin[D.job];
      uint32_t c = 0;
      for (uint32_t q = lane + 1; q < S.cnt; q += 64) {
        const uint32_t a = e[q - 1], b = e[q];
        const unsigned long long tt = b >> 2;
        c += (is_toggle(a & 3u, b & 3u) && tt > W.tc_lo && tt <= W.tc_hi) ? 1u : 0u;
      }
      c = wave_sum(c);
      if (lane == 0 && c) atomicAdd(saif + (uint64_t)kSaifWords * D.out_net + 3, (unsigned long long)c);
    }
  }
}

// Wide gates (5 or 6 distinct inputs): one thread walks a gate's whole
// chunk, with the walk's rules in sequence (the same merge, truth table,
// edge class, AddDelay per input, preemption in closed form, chunk
// bounds), and leaves its tile output and summary exactly as a one-tile
// lane walk would (k_gate_carry and k_compact then treat it like any other
// tile).  The candidates are built in the tile's output region itself
// (it holds every input word + 1) and compacted in place.
__global__ void k_wide(WideArgs A) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.nwide) return;
  // a level over the captured capacities was skipped (no descriptors)
  if (*A.err & ERR_CAPACITY) return;
  const uint32_t g = A.wide[i] - A.g0;
  const GateRec& G = A.gates[g];
  const uint64_t tk = A.tile_base[g];
  const uint32_t cap = A.tiles[tk].out_cap;
  uint32_t* st = A.scratch + A.scr_base[g];
  const uint8_t* lut = A.luts + (uint32_t)G.lut * 256u;
  const uint16_t* tab = A.dtab + G.dt_off;
  const uint32_t k = G.k;
  const uint32_t kX = 1u, INF = 0xFFFFFFFFu;
  const uint32_t* in[kWideK];
  uint32_t n[kWideK], r[kWideK], h[kWideK];
  uint32_t idx = 0;
#pragma unroll
  for (uint32_t j = 0; j < kWideK; ++j) {
    if (j < k) {
      in[j] = A.pool + A.woff[G.in_net[j]];
      n[j] = A.wlen[G.in_net[j]];
      idx |= (in[j][0] & 3u) << (2 * j);  // the state entry
      r[j] = 1;
      h[j] = n[j] > 1 ? in[j][1] : INF;
    } else {
      in[j] = nullptr;
      n[j] = r[j] = 0;
      h[j] = INF;
    }
  }
  const bool init = A.first != 0;
  uint32_t fprev = init ? kX : lut[idx];
  uint32_t nc = 0, nown = 0, bad = 0;
  if (init) st[nc++] = (1u << 2) | kX;  // (0, X) at rel 1; cap >= 1
  for (;;) {
    uint32_t w = h[0];
#pragma unroll
    for (uint32_t j = 1; j < kWideK; ++j) w = min(w, h[j]);
    const uint32_t t = w >> 2;
    if (t >= 0x3FFFFFFFu) break;  // every input at its end
    const uint32_t lim = w | 3u;
    uint32_t mask = 0, nidx = idx;
#pragma unroll
    for (uint32_t j = 0; j < kWideK; ++j) {
      if (h[j] <= lim) {
        mask |= 1u << j;
        nidx = (nidx & ~(3u << (2 * j))) | ((h[j] & 3u) << (2 * j));
        ++r[j];
        h[j] = r[j] < n[j] ? in[j][r[j]] : INF;
      }
    }
    const uint32_t f = lut[nidx];
    const uint32_t sc = edge_cls(fprev, f);
    if (sc != 0u) {
      uint32_t d;
      if ((mask & (mask - 1u)) == 0u) {
        const uint32_t j = __builtin_ctz(mask);
        const uint32_t p = (idx >> (2 * j)) & 3u, c = (nidx >> (2 * j)) & 3u;
        const uint32_t fell = ((c == 0u && p != 0u) || (p == 3u && (c == 1u || c == 2u))) ? 1u : 0u;
        d = tab[wide_dt_index(j, sc, fell, 1)];
      } else {
        uint32_t m = 0xFFFFu;
#pragma unroll
        for (uint32_t j = 0; j < kWideK; ++j) {
          if (!((mask >> j) & 1u)) continue;
          const uint32_t p = (idx >> (2 * j)) & 3u, c = (nidx >> (2 * j)) & 3u;
          const uint32_t fell = ((c == 0u && p != 0u) || (p == 3u && (c == 1u || c == 2u))) ? 1u : 0u;
          m = min(m, (uint32_t)tab[wide_dt_index(j, sc, fell, 0)]);
        }
        d = m == 0xFFFFu ? 1u : m;
      }
      if (nc < cap) st[nc++] = ((t + d) << 2) | f;
      else bad = 1;
    }
    nown += t >= A.lo_rel ? 1u : 0u;
    idx = nidx;
    fprev = f;
  }
  // keep the candidates strictly earlier than every later one (and inside
  // [lo_rel, hi_rel)), compacted in time order to the end of the list
  uint32_t runmin = A.hi_rel, jj = nc, cnt = 0, top = 0, lastv = 4u, has = 0;
  for (uint32_t q = nc; q-- > 0;) {
    const uint32_t e = st[q];
    const uint32_t o = e >> 2;
    const bool keep = o >= A.lo_rel && o < runmin;
    runmin = min(runmin, o);
    if (keep) {
      st[--jj] = e;
      cnt += (e & 3u) != lastv ? 1u : 0u;
      lastv = e & 3u;
      if (!has) top = e;
      has = 1;
    }
  }
  // entries (deduplicated) to the front of the region, SAIF durations
  // between them, summary
  const ChunkWin W = A.cwin[G.job];
  const uint32_t s32 = (uint32_t)min(W.lo, 0x7FFFFFFFull);
  const uint32_t e32 = (uint32_t)min(W.hi, 0x7FFFFFFFull);
  const uint32_t sp = nc - jj;
  const uint32_t e_first = sp ? st[jj] : 0u;
  uint32_t pvv = 4u, pct = max(min(e_first >> 2, e32), s32);
  const uint32_t ct0 = pct;
  uint32_t a0 = 0, a1 = 0, wpos = 0;
  for (uint32_t q = 0; q < sp; ++q) {
    const uint32_t e = st[jj + q];
    const uint32_t v = e & 3u;
    const uint32_t ct = max(min(e >> 2, e32), s32);
    const uint32_t len = ct - pct;
    a0 += pvv == 0u ? len : 0u;
    a1 += pvv == 3u ? len : 0u;
    if (v != pvv) st[wpos++] = e;  // wpos <= q <= jj + q: read before written
    pvv = v;
    pct = ct;
  }
  TileOut o;
  o.cnt = min(cnt, cap);
  o.ft = sp ? e_first >> 2 : 0u;
  o.lt = has ? top >> 2 : 0u;
  o.flags = has ? ((e_first & 3u) | ((top & 3u) << 2) | (1u << 4)) : 0u;
  o.t0 = a0;
  o.t1 = a1;
  o.tx = pct - ct0 - a0 - a1;
  o.ev = nown;
  A.tout[tk] = o;
  if (cnt > cap || bad) atomicOr(A.err, ERR_ORDER);
  if (nown) atomicAdd(A.ev_count, (unsigned long long)nown);
}

hipError_t launch_wide(const WideArgs& a, hipStream_t st) {
  if (a.nwide == 0) return hipSuccess;
  hipLaunchKernelGGL(k_wide, dim3((a.nwide + 63) / 64), dim3(64), 0, st, a);
  return hipGetLastError();
}

// Start-of-run scalars: bump pointer past the stimulus words, all else 0.
__global__ void k_set_scal(uint64_t* scal, uint64_t bump) {
  if (blockIdx.x == 0 && threadIdx.x < 16) scal[threadIdx.x] = threadIdx.x == 0 ? bump : 0;
}

// per gate: leader, stride, segments, tiles, output bound
__global__ void k_plan2(const GateRec* __restrict__ gates, uint32_t ng,
                        const uint32_t* __restrict__ wlen, uint32_t seg_target,
                        GatePlan* __restrict__ plan, uint64_t* __restrict__ ntile,
                        uint64_t* __restrict__ bound, unsigned int* __restrict__ err,
                        const uint32_t* __restrict__ tail_len) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const GateRec& G = gates[g];
  uint64_t N = 0;
  uint32_t nl = 0, lead = 0;
  for (int j = 0; j < G.k; ++j) {
    uint32_t n = wlen[G.in_net[j]];
    N += n;
    if (j == 0 || n > nl) {
      nl = n;
      lead = j;
    }
  }
  // a tile's offset inside its gate's waveform travels in 31 bits
  if (N + 1 >= (1ull << 30)) atomicOr(err, ERR_RANGE);
  if (G.k > kTileK) {  // wide gate: one tile, walked by k_wide
    GatePlan P;
    P.nseg = 1;
    P.stride = nl > 0 ? nl : 1u;
    P.lead = lead;
    P.pad = 0;
    plan[g] = P;
    ntile[g] = 1;
    bound[g] = N + 2 + tail_len[G.out_net];
    return;
  }
  // leader events per segment so that a segment holds ~seg_target input
  // events; then the gate's tiles are filled evenly (the stride that spreads
  // the leader's events over all 64 segments of each of those tiles, never
  // larger), so a small gate's single tile uses every lane
  uint64_t s = N > 0 ? (uint64_t)seg_target * nl / N : 1;
  if (s < 1) s = 1;
  if (nl > 0) {
    const uint64_t nt0 = ((nl + s - 1) / s + kLanes - 1) / kLanes;
    s = (nl + nt0 * kLanes - 1) / (nt0 * kLanes);
  }
  uint32_t ns = nl > 0 ? (uint32_t)((nl + s - 1) / s) : 1u;
  GatePlan P;
  P.nseg = ns;
  P.stride = (uint32_t)s;
  P.lead = lead;
  P.pad = 0;
  plan[g] = P;
  const uint32_t nt = (ns + kLanes - 1) / kLanes;
  ntile[g] = nt;
  // scratch words of the gate's tile outputs: every tile keeps at most its
  // own input events + 1 entries; the gate's tail is added so that the
  // level's waveforms (tails + joined entries) never exceed the scratch
  bound[g] = N + nt + 1 + tail_len[G.out_net];
}

// per (tile, input): the tile's descriptor (gate, tile index within the
// gate, segmentation, tables) and input j's range
__global__ void k_tile_desc(const GateRec* __restrict__ gates, uint32_t ng,
                            const GatePlan* __restrict__ plan,
                            const uint64_t* __restrict__ tile_base,
                            const uint64_t* __restrict__ woff, const uint32_t* __restrict__ wlen,
                            const uint32_t* __restrict__ pool, const uint32_t* __restrict__ lutp,
                            TileDesc* __restrict__ tiles, const uint64_t* __restrict__ ntiles) {
  // a tile's first own word of each input is the previous tile's end: the
  // block's tiles pass them on through LDS (each boundary searched once)
  __shared__ uint32_t s_first[256];
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t t = x >> 2;
  const int j = (int)(x & 3);
  const uint64_t nt = *ntiles;
  const bool valid = t < nt;
  // gate of tile t: binary search over tile_base by the tile's first
  // thread, passed to the other three
  uint32_t g = 0;
  if (valid && j == 0) {
    uint32_t lo = 0, hi = ng;
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (tile_base[mid] <= t) lo = mid;
      else hi = mid;
    }
    g = lo;
  }
  g = __shfl(g, (int)((threadIdx.x & 63u) & ~3u));
  GateRec G;
  GatePlan P;
  uint32_t ti = 0, L = 0, c0 = 0, clast = 0, n = 0;
  const uint32_t* lw = nullptr;
  const uint32_t* iw = nullptr;
  bool narrow_in = false;  // input j of a gate of fan-in <= 4
  if (valid) {
    G = gates[g];
    P = plan[g];
    ti = (uint32_t)(t - tile_base[g]);
    narrow_in = G.k <= kTileK && (uint32_t)j < G.k;
    if (narrow_in) {
      lw = pool + woff[G.in_net[P.lead]];
      iw = pool + woff[G.in_net[j]];
      n = wlen[G.in_net[j]];
      c0 = ti * kLanes;
      clast = min(c0 + kLanes - 1, P.nseg - 1);
      if (c0 > 0) {
        const uint32_t tau_lo = lw[c0 * P.stride] >> 2;
        L = (uint32_t)j == P.lead ? c0 * P.stride : lb_words(iw, n, tau_lo << 2);
      }
    }
  }
  s_first[threadIdx.x] = L;
  __syncthreads();
  if (!valid) return;
  TileDesc& D = tiles[t];
  D.lutw[j] = lutp[(uint32_t)G.lut * 16u + j];
  if (j == 0) {
    D.g = g;
    D.ti = ti;
    D.nseg = P.nseg;
    D.stride = P.stride;
    D.lead = P.lead;
    D.k = G.k;
    D.dt_off = G.dt_off;
    D.lut = G.lut;
    D.dmax = G.dmax;
    D.out_net = G.out_net;
    D.job = G.job;
  }
  // own input words of the tile (up to the next tile's first words), summed
  // over the tile's four threads: its output region in the level scratch
  // holds that many entries plus one
  uint32_t span = 0;
  if (G.k > kTileK) {
    // wide gate (one tile): no staged ranges; the output region holds all
    // input words + 1 (threads 0 and 1 also count inputs 4 and 5)
    D.L[j] = 0;
    D.E[j] = 0;
    D.in_off[j] = 0;
    span = wlen[G.in_net[j]];
    if ((uint32_t)j + 4 < G.k) span += wlen[G.in_net[j + 4]];
  } else if (!narrow_in) {
    D.L[j] = 0;
    D.E[j] = 0;
    D.in_off[j] = 0;
  } else {
    uint32_t E = n, Ln = n;
    if (clast + 1 < P.nseg) {
      // the next tile is this gate's: its first words end this one
      const uint32_t tau_nx = lw[(clast + 1) * P.stride] >> 2;
      if (threadIdx.x + 4 < blockDim.x && t + 1 < nt)
        Ln = s_first[threadIdx.x + 4];
      else
        Ln = (uint32_t)j == P.lead ? (clast + 1) * P.stride : lb_words(iw, n, tau_nx << 2);
      // halo end: the first word at or after tau_nx + dmax lies at or after
      // Ln, usually at it; step forward with doubling strides, then search
      // only inside the last stride (every word before a is below the key)
      const uint32_t key = (tau_nx + G.dmax) << 2;
      E = Ln;
      if (E < n && iw[E] < key) {
        uint32_t a = E, d = 1, b = a + 1;
        while (b < n && iw[b] < key) {
          a = b;
          d <<= 1;
          b = a + d;
        }
        if (b > n) b = n;
        E = a + 1 + lb_words(iw + a + 1, b - a - 1, key);
      }
    }
    D.L[j] = L;
    D.E[j] = E;
    D.in_off[j] = woff[G.in_net[j]] + (L > 0 ? L - 1 : 0u);
    span = Ln - L;
  }
  span += __shfl_xor(span, 1);
  span += __shfl_xor(span, 2);
  if (j == 0) D.out_cap = span + 1;
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_plan2(const GateRec* gates, uint32_t ng, const uint32_t* wlen,
                        uint32_t seg_target, GatePlan* plan, uint64_t* ntile, uint64_t* bound,
                        unsigned int* err, const uint32_t* tail_len, hipStream_t st) {
  k_plan2<<<(ng + 255) / 256, 256, 0, st>>>(gates, ng, wlen, seg_target, plan, ntile, bound, err,
                                            tail_len);
  return hipGetLastError();
}

hipError_t launch_tile_desc(const GateRec* gates, uint32_t ng, const GatePlan* plan,
                            const uint64_t* tile_base, const uint64_t* woff, const uint32_t* wlen,
                            const uint32_t* pool, const uint32_t* lutp, TileDesc* tiles,
                            const uint64_t* ntiles_dev, uint64_t ntiles, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  const uint64_t threads = ntiles * 4;
  k_tile_desc<<<(unsigned)((threads + 255) / 256), 256, 0, st>>>(gates, ng, plan, tile_base, woff,
                                                                  wlen, pool, lutp, tiles, ntiles_dev);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v,
                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_fill_u64<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

// Device-driven levels: when the tile count or the level's scratch does not
// fit the arrays sized by the first run, the level is skipped (tile count 0)
// and the error flag says why.
__global__ void k_guard(uint64_t* scal, uint64_t tiles_cap, uint64_t scratch_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // scal: 1 tiles of the level, 2 scratch words of the level, 7 error flags
  if (scal[1] > tiles_cap || scal[2] > scratch_cap) {
    scal[1] = 0;
    scal[2] = 0;
    atomicOr((unsigned int*)(scal + 7), ERR_CAPACITY);
  }
}

// After the level's output size is known: when it does not fit the pool,
// flag the error (the compaction then leaves the level empty) and reserve
// nothing.  scal: 0 bump, 3 output words of the level.
__global__ void k_pool_guard(uint64_t* scal, uint64_t pool_words) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (scal[0] + scal[3] > pool_words) {
    scal[3] = 0;
    atomicOr((unsigned int*)(scal + 7), ERR_CAPACITY | ERR_POOL);
  }
}

hipError_t launch_pool_guard(uint64_t* scal, uint64_t pool_words, hipStream_t st) {
  k_pool_guard<<<1, 64, 0, st>>>(scal, pool_words);
  return hipGetLastError();
}

hipError_t launch_gate_carry(const GateRec* gates, uint32_t ng, const uint64_t* tile_base,
                             const uint64_t* ntile, const TileOut* tout, uint32_t* tile_off,
                             uint64_t* gate_cnt, uint32_t* wlen, unsigned long long* saif,
                             const ChunkWin* cwin, const unsigned int* err, const TailArgs& T,
                             unsigned long long* ev_job, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  k_gate_carry<<<(ng + 3) / 4, 256, 0, st>>>(gates, ng, tile_base, ntile, tout, tile_off, gate_cnt,
                                             wlen, saif, cwin, err, T, ev_job);  // one wavefront per gate
  return hipGetLastError();
}

hipError_t launch_compact(const TileDesc* tiles, const uint64_t* ntiles_dev, uint64_t ntiles_cap,
                          const TileOut* tout, const uint32_t* tile_off, const uint64_t* scr_base,
                          const uint32_t* scratch, const uint64_t* gate_base, uint32_t* pool,
                          uint64_t* woff, uint32_t* wlen, const unsigned int* err,
                          unsigned long long* saif, const ChunkWin* cwin, uint32_t count_tc,
                          const TailArgs& T, hipStream_t st) {
  if (ntiles_cap == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((ntiles_cap + 3) / 4, 8192);  // 4 wavefronts each
  k_compact<<<(unsigned)blocks, 256, 0, st>>>(tiles, ntiles_dev, tout, tile_off, scr_base, scratch,
                                              gate_base, pool, woff, wlen, err, saif, cwin, count_tc,
                                              T);
  return hipGetLastError();
}

hipError_t launch_guard(uint64_t* scal, uint64_t tiles_cap, uint64_t scratch_cap, hipStream_t st) {
  k_guard<<<1, 64, 0, st>>>(scal, tiles_cap, scratch_cap);
  return hipGetLastError();
}


hipError_t launch_set_scal(uint64_t* scal, uint64_t bump, hipStream_t st) {
  k_set_scal<<<1, 64, 0, st>>>(scal, bump);
  return hipGetLastError();
}

hipError_t launch_level(const LevelArgs& a, unsigned nblocks, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_level, dim3(nblocks), dim3(kLanes), level_lds_bytes(a.cap, a.stk == nullptr),
                     st, a);
  return hipGetLastError();
}

int level_max_blocks(uint32_t cap, bool stacks_in_lds) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 2048;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_level, kLanes,
                                                   level_lds_bytes(cap, stacks_in_lds)) != hipSuccess)
    per_cu = 4;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  return per_cu * cus;
}

}  // namespace gls
