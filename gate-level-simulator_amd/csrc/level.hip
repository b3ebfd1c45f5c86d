// Fused per-level gate evaluation kernel (gfx950).
//
// Work unit: a tile = 64 consecutive time segments of one gate, one segment
// per lane of a wavefront.  Segment c of a gate covers the events
// [c*stride, (c+1)*stride) of the gate's longest input (the "leader"), so
// segment boundaries are exact event times.  Per tile:
//   1. the input waveform ranges the tile needs (found beforehand by
//      k_tile_desc: one event before the tile for the starting values and a
//      D_max look-ahead halo after it) are copied into LDS with batched,
//      coalesced loads;
//   2. every lane merges its segment's input events in time order (branch
//      free: each LDS range ends with an INF sentinel), evaluates the 4-state
//      truth table (packed in registers for fan-in <= 3), classifies the
//      output edge, picks the SDF delay with one lookup in the gate's delay
//      table (AddDelay folded on the host, copied to LDS per tile) and applies
//      the preemption rule on a stack in LDS, counting value changes as it
//      goes; events in the halo only pre-empt;
//   3. a wavefront scan combines the lanes' kept lists (dedup of equal
//      adjacent values across lanes), a decoupled look-back over the gate's
//      earlier tiles supplies the carry (last value, output count), and the
//      tile's dense output is written with coalesced stores from LDS;
//   4. SAIF T0/T1/TX durations are reduced per wavefront and added with one
//      atomic per counter.
// Tiles are handed out by an atomic ticket in increasing order, so a tile
// only ever waits for tiles already taken by running wavefronts.  The wait
// is bounded: past the spin limit the kernel raises an error flag and
// finishes instead of hanging.
//
// Event rules: src/simulate.cpp:919-1201 (Calculate, AddDelay, CheckDelay),
// src/simulate.cpp:1365-1506 (Process, GetValueAt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device.h"

namespace gls {

constexpr int kLanes = 64;
constexpr uint32_t kSpinLimit = 1u << 24;

enum : unsigned { ERR_LOOKBACK = 1u, ERR_OVERFLOW = 2u, ERR_RANGE = 4u, ERR_CAPACITY = 8u };

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

__device__ __forceinline__ Agg agg_shfl_up(const Agg& a, int d) {
  Agg r;
  r.cnt = __shfl_up(a.cnt, d, 64);
  r.lt = __shfl_up(a.lt, d, 64);
  uint32_t f = a.fv | (a.lv << 2) | (a.has << 4);
  f = __shfl_up(f, d, 64);
  r.fv = f & 3u;
  r.lv = (f >> 2) & 3u;
  r.has = (f >> 4) & 1u;
  return r;
}

__device__ __forceinline__ Agg agg_shfl(const Agg& a, int src) {
  Agg r;
  r.cnt = __shfl(a.cnt, src, 64);
  r.lt = __shfl(a.lt, src, 64);
  uint32_t f = a.fv | (a.lv << 2) | (a.has << 4);
  f = __shfl(f, src, 64);
  r.fv = f & 3u;
  r.lv = (f >> 2) & 3u;
  r.has = (f >> 4) & 1u;
  return r;
}

// value of lane+d, or the empty aggregate past the last lane
__device__ __forceinline__ Agg agg_shfl_down_ordered(const Agg& a, int d, int lane) {
  Agg r;
  r.cnt = __shfl_down(a.cnt, d, 64);
  r.lt = __shfl_down(a.lt, d, 64);
  uint32_t f = a.fv | (a.lv << 2) | (a.has << 4);
  f = __shfl_down(f, d, 64);
  r.fv = f & 3u;
  r.lv = (f >> 2) & 3u;
  r.has = (f >> 4) & 1u;
  if (lane + d >= 64) {
    r.has = 0;
    r.cnt = 0;
  }
  return r;
}

// aggregate word: valid<<63 | lt<<33 | has<<32 | lv<<30 | fv<<28 | cnt(28).
// A tile's count is bounded by its input events: the LDS capacity, or the
// overflow buffer size (< 2^28 words) for tiles that do not fit LDS.
__device__ __forceinline__ unsigned long long pack_agg(const Agg& a) {
  return (1ull << 63) | ((unsigned long long)a.lt << 33) | ((unsigned long long)a.has << 32) |
         ((unsigned long long)a.lv << 30) | ((unsigned long long)a.fv << 28) |
         (a.cnt & 0x0FFFFFFFu);
}
__device__ __forceinline__ Agg unpack_agg(unsigned long long w) {
  Agg a;
  a.cnt = (uint32_t)(w & 0x0FFFFFFFu);
  a.fv = (uint32_t)(w >> 28) & 3u;
  a.lv = (uint32_t)(w >> 30) & 3u;
  a.has = (uint32_t)(w >> 32) & 1u;
  a.lt = (uint32_t)(w >> 33) & 0x3FFFFFFFu;
  return a;
}
// inclusive-prefix word: valid<<63 | lt<<33 | has<<32 | lv<<30 | cnt(30)
__device__ __forceinline__ unsigned long long pack_inc(const Agg& a) {
  return (1ull << 63) | ((unsigned long long)a.lt << 33) | ((unsigned long long)a.has << 32) |
         ((unsigned long long)a.lv << 30) | (a.cnt & 0x3FFFFFFFu);
}
__device__ __forceinline__ Agg unpack_inc(unsigned long long w) {
  Agg a;
  a.cnt = (uint32_t)(w & 0x3FFFFFFFu);
  a.lv = (uint32_t)(w >> 30) & 3u;
  a.has = (uint32_t)(w >> 32) & 1u;
  a.lt = (uint32_t)(w >> 33) & 0x3FFFFFFFu;
  a.fv = 0;  // not needed to the right of an inclusive prefix
  return a;
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

__device__ __forceinline__ uint32_t lb_words(const uint32_t* a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Uniform (per tile) data of one gate.
template <int K>
struct TileCtx {
  const uint32_t* src[K];  // input ranges (LDS or HBM), index relative to B
  uint32_t B[K], E[K], L[K];
  uint32_t nseg, stride, lead, dmax;
  uint32_t lut[4];          // packed truth table, fan-in <= 3
};

template <int K>
__device__ __forceinline__ uint32_t lut_eval(const TileCtx<K>& X, const uint8_t* s_lut,
                                             uint32_t idx) {
  if (K <= 2) return (X.lut[0] >> (2 * idx)) & 3u;
  if (K == 3) {
    // 64 entries = 128 bits: pick the 64-bit half, then shift
    const uint64_t lo = (uint64_t)X.lut[0] | ((uint64_t)X.lut[1] << 32);
    const uint64_t hi = (uint64_t)X.lut[2] | ((uint64_t)X.lut[3] << 32);
    const uint64_t w = (idx & 32u) ? hi : lo;
    return (uint32_t)(w >> (2 * (idx & 31u))) & 3u;
  }
  return s_lut[idx];
}

// Inputs that fell (CheckDelay == 2: to 0 from X/Z/1, or from 1 to X/Z)
// between the packed input values p and c (2 bits per input), as a bit mask.
__device__ __forceinline__ uint32_t fell_bits(uint32_t p, uint32_t c) {
  const uint32_t M = 0x55u;
  const uint32_t c0 = ~(c | (c >> 1)) & M;   // now 0
  const uint32_t pn0 = (p | (p >> 1)) & M;   // was not 0
  const uint32_t p1 = p & (p >> 1) & M;      // was 1
  const uint32_t cxz = (c ^ (c >> 1)) & M;   // now X or Z
  uint32_t x = (c0 & pn0) | (p1 & cxz);      // bit 2j per input j
  x = (x | (x >> 1)) & 0x33u;
  return (x | (x >> 2)) & 0x0Fu;
}

// Remove the top entry of a lane's stack, keeping the count of value changes.
__device__ __forceinline__ void stack_pop(const uint32_t* st, uint32_t& sp, uint32_t& top,
                                          uint32_t& cnt) {
  --sp;
  const uint32_t nt = sp ? st[sp - 1] : 0u;
  cnt -= (sp == 0 || ((nt ^ top) & 3u)) ? 1u : 0u;
  top = nt;
}

// What one lane's walk over its segment leaves behind: its stack of kept
// candidates (st[0..sp)), the top entry, the number of value changes in the
// stack, and the number of gate events it owned.
struct Lane {
  uint32_t* st;
  uint32_t sp, top, cnt, nown;
};

typedef __attribute__((address_space(3))) void lds_void;

// One 4-byte direct-to-LDS load per lane: lane i's word lands at l + i.
__device__ __forceinline__ void dma_word(const uint32_t* g, uint32_t* l) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)l, 4, 0, 0);
}

// Wait until this wavefront has no vector-memory access in flight (direct-to-
// LDS loads included): s_waitcnt vmcnt(0).
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ uint32_t range_start(uint32_t L) { return L > 0 ? L - 1 : 0u; }

__device__ __forceinline__ uint32_t tile_words(const TileDesc& D) {
  uint32_t total = 0;
  for (uint32_t j = 0; j < D.k; ++j) total += D.E[j] - range_start(D.L[j]);
  return total;
}

// Issue the loads of tile D into LDS: its input ranges (range j at
// sum_{i<j}(n_i + 1), leaving one word for the sentinel) when they fit, its
// delay table and (fan-in 4) its truth table.  Completion: wait_vm().
__device__ __forceinline__ void stage_tile(const LevelArgs& A, const TileDesc& D, bool fits,
                                           uint32_t* s_in, uint32_t* s_lut, uint32_t* s_dt) {
  const uint32_t lane = threadIdx.x;
  if (fits) {
    uint32_t pos = 0;
    for (uint32_t j = 0; j < D.k; ++j) {
      const uint32_t n = D.E[j] - range_start(D.L[j]);
      const uint32_t* src = A.pool + D.in_off[j];
      for (uint32_t i = 0; i < n; i += kLanes)
        if (i + lane < n) dma_word(src + i + lane, s_in + pos + i);
      pos += n + 1;
    }
  }
  const uint32_t nw = dt_size(D.k) / 2;
  const uint32_t* dsrc = (const uint32_t*)(A.dtab + D.dt_off);
  for (uint32_t i = 0; i < nw; i += kLanes)
    if (i + lane < nw) dma_word(dsrc + i + lane, s_dt + i);
  if (D.k == 4) dma_word((const uint32_t*)(A.luts + D.lut * 256u) + lane, s_lut);
}

// One lane's walk over its time segment: merge the inputs in time order,
// evaluate, pick the delay, keep the preemption stack.
template <int K, bool LDS>
__device__ __forceinline__ Lane lane_walk(const TileCtx<K>& X, uint32_t ti, const uint8_t* s_lut,
                                          const uint16_t* s_dt, uint32_t* stk_base) {
  const int lane = threadIdx.x;
  const uint32_t c = ti * kLanes + lane;
  const uint32_t kX = 1u;
  const uint32_t INF = 0xFFFFFFFFu;
  uint32_t r[K], h[K], bound[K];
  const uint32_t* srcL = X.src[0];
  uint32_t BL = X.B[0];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    bound[j] = X.E[j] - X.B[j];
    if ((uint32_t)j == X.lead) {
      srcL = X.src[j];
      BL = X.B[j];
    }
  }
  Lane ln;
  ln.st = stk_base;
  // sp: stack depth, top: top entry, cnt: value changes in the stack
  uint32_t sp = 0, top = 0, cnt = 0, nown = 0, hmin = INF;
  if (c < X.nseg) {
    // segment bounds from the leader
    uint32_t tau_c = 0, tau_n = 0x3FFFFFFFu, halo_end = 0x3FFFFFFFu;
    const uint32_t pl = c * X.stride;
    if (c > 0) tau_c = srcL[pl - BL] >> 2;
    if (c + 1 < X.nseg) {
      tau_n = srcL[pl + X.stride - BL] >> 2;
      halo_end = tau_n + X.dmax;
    }
    uint32_t soff = lane, idx = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      uint32_t pj;
      if (c == 0) pj = 0;
      else if ((uint32_t)j == X.lead) pj = pl;
      else pj = X.B[j] + lb_words(X.src[j], bound[j], tau_c << 2);
      soff += pj - X.L[j];
      r[j] = pj - X.B[j];
      idx |= (pj > 0 ? (X.src[j][r[j] - 1] & 3u) : kX) << (2 * j);
      // LDS ranges end with an INF sentinel
      h[j] = (LDS || r[j] < bound[j]) ? X.src[j][r[j]] : INF;
    }
    uint32_t* st = stk_base + soff;
    ln.st = st;
    uint32_t fprev = c == 0 ? kX : lut_eval<K>(X, s_lut, idx);
    if (c == 0) {  // every output starts with (0, X) (src/simulate.cpp:882-887)
      top = kX;
      st[0] = top;
      sp = 1;
      cnt = 1;
    }
    for (;;) {
      uint32_t w = h[0];
#pragma unroll
      for (int j = 1; j < K; ++j) w = min(w, h[j]);
      const uint32_t t = w >> 2;
      if (t >= halo_end) break;  // also the end of every input (INF)
      // every input with an event at time t
      const uint32_t lim = w | 3u;
      uint32_t mask = 0, nidx = idx;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const bool take = h[j] <= lim;
        mask |= take ? (1u << j) : 0u;
        nidx = take ? ((nidx & ~(3u << (2 * j))) | ((h[j] & 3u) << (2 * j))) : nidx;
        r[j] += take ? 1u : 0u;
      }
#pragma unroll
      for (int j = 0; j < K; ++j) h[j] = (LDS || r[j] < bound[j]) ? X.src[j][r[j]] : INF;
      const uint32_t f = lut_eval<K>(X, s_lut, nidx);
      const uint32_t sc = edge_cls(fprev, f);
      const uint32_t fl = fell_bits(idx, nidx);
      idx = nidx;
      fprev = f;
      const bool own = t < tau_n;
      nown += own ? 1u : 0u;
      if (sc != 0) {
        const uint32_t o = t + s_dt[dt_index(K, sc, mask, fl)];
        if (own) {
          // preemption: drop queued entries at or after o (src/simulate.cpp:991-1007)
          while (sp > 0 && (top >> 2) >= o) stack_pop(st, sp, top, cnt);
          const uint32_t e = (o << 2) | f;
          cnt += (sp == 0 || ((e ^ top) & 3u)) ? 1u : 0u;
          st[sp++] = e;
          top = e;
        } else {
          hmin = min(hmin, o);
        }
      }
    }
    // the next segment's candidates pre-empt this segment's tail
    while (sp > 0 && (top >> 2) >= hmin) stack_pop(st, sp, top, cnt);
  }
  ln.sp = sp;
  ln.top = top;
  ln.cnt = cnt;
  ln.nown = nown;
  return ln;
}

// Tile context and walk for fan-in K.  Tiles that do not fit LDS read their
// inputs from HBM and keep their stacks in the overflow buffer.
template <int K>
__device__ __forceinline__ Lane walk_k(const LevelArgs& A, const TileDesc& D, uint64_t tk,
                                       bool fits, const uint32_t* s_in, uint32_t* s_st,
                                       const uint8_t* s_lut, const uint16_t* s_dt) {
  const int lane = threadIdx.x;
  TileCtx<K> X;
  X.nseg = D.nseg;
  X.stride = D.stride;
  X.lead = D.lead;
  X.dmax = D.dmax;
#pragma unroll
  for (int q = 0; q < 4; ++q) X.lut[q] = D.lutw[q];
  uint32_t pos = 0, total = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    X.L[j] = D.L[j];
    X.E[j] = D.E[j];
    X.B[j] = range_start(D.L[j]);
    const uint32_t n = X.E[j] - X.B[j];
    X.src[j] = fits ? s_in + pos : A.pool + D.in_off[j];
    pos += n + 1;
    total += n;
  }
  if (fits) return lane_walk<K, true>(X, D.ti, s_lut, s_dt, s_st);
  unsigned long long base = 0;
  const unsigned long long need = (unsigned long long)total + kLanes;
  if (lane == 0) base = atomicAdd(A.ovf_used, need);
  base = __shfl(base, 0, 64);
  if (base + need > A.ovf_cap) {
    if (lane == 0) atomicOr(A.err, ERR_OVERFLOW);
    base = 0;  // results are invalid (the error is reported); keep every wavefront moving
    if (need > A.ovf_cap) {
      Lane ln;
      ln.st = A.ovf;
      ln.sp = ln.top = ln.cnt = ln.nown = 0;
      return ln;
    }
  }
  return lane_walk<K, false>(X, D.ti, s_lut, s_dt, A.ovf + base);
}

// Combine the lanes' stacks into the tile's output: wavefront scan of the
// lane aggregates, decoupled look-back over the gate's earlier tiles, dedup
// against the carry, stores, SAIF durations (summed per lane in 32 bits:
// times < 2^30) and event count.
// Wavefront scan of the lane aggregates and the gate prefix before the tile.
// chain: this wavefront finished the gate's previous tile just before, and
// prev_incl is its inclusive prefix (no look-back needed).
struct TileSum {
  Agg carry;          // everything before this lane's first entry (gate prefix + earlier lanes)
  Agg incl;           // inclusive prefix of the gate through this tile
  uint32_t tile_out0; // output index of the tile's first entry
  uint32_t e_first;   // the lane's first stack entry
};

__device__ __forceinline__ TileSum combine_tile(const LevelArgs& A, const TileDesc& D,
                                                uint64_t tk, const Lane& ln, bool chain,
                                                const Agg& prev_incl) {
  const int lane = threadIdx.x;
  const uint32_t ti = D.ti;
  const uint32_t sp = ln.sp;
  const uint32_t* st = ln.st;
  const uint32_t e_first = sp ? st[0] : 0u;
  Agg mine;
  mine.cnt = ln.cnt;
  mine.has = sp > 0 ? 1u : 0u;
  mine.fv = e_first & 3u;
  mine.lv = ln.top & 3u;
  mine.lt = ln.top >> 2;
  // wavefront inclusive scan of lane aggregates
  Agg inc = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    Agg o = agg_shfl_up(inc, d);
    if (lane >= d) inc = agg_combine(o, inc);
  }
  Agg lane_excl = agg_shfl_up(inc, 1);
  if (lane == 0) lane_excl.has = 0, lane_excl.cnt = 0;
  const Agg tile_agg = agg_shfl(inc, 63);

  // decoupled look-back over the gate's earlier tiles, 64 predecessors per
  // round: lane i reads tile tk-1-i (tiles before the gate's first one act
  // as an empty inclusive prefix).  A look-back word carries its whole
  // payload, so relaxed device-scope atomics suffice; release/acquire would
  // write back and invalidate the XCD's L2 on every tile.
  unsigned long long* ts = A.tstate;
  Agg excl;
  excl.has = 0;
  excl.cnt = 0;
  excl.fv = excl.lv = excl.lt = 0;
  if (ti == 0) {
    if (lane == 0)
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(tile_agg), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  } else if (chain) {
    excl = prev_incl;
    if (lane == 0)
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(agg_combine(excl, tile_agg)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  } else {
    if (lane == 0)
      __hip_atomic_store(&ts[2 * tk], pack_agg(tile_agg), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t first = tk - ti;  // the gate's first tile
    Agg acc;                         // combination of the tiles after the window
    acc.has = 0;
    acc.cnt = 0;
    acc.fv = acc.lv = acc.lt = 0;
    uint64_t top_tile = tk - 1;  // tile read by lane 0 in this round
    uint32_t spins = 0;
    for (;;) {
      const bool in_gate = top_tile >= first + lane;  // tile top_tile-lane belongs to the gate
      const uint64_t p = top_tile - lane;
      unsigned long long wi = 0, wa = 0;
      if (in_gate) {
        wi = __hip_atomic_load(&ts[2 * p + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(wi >> 63))
          wa = __hip_atomic_load(&ts[2 * p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // status: 2 = inclusive prefix (or before the gate), 1 = aggregate, 0 = not ready
      const uint32_t status = !in_gate ? 2u : ((wi >> 63) ? 2u : ((wa >> 63) ? 1u : 0u));
      const uint64_t have_inc = __ballot(status == 2u);
      const uint64_t not_ready = __ballot(status == 0u);
      const int ip = have_inc ? __ffsll((long long)have_inc) - 1 : 64;  // nearest inclusive
      const uint64_t need = ip >= 64 ? ~0ull : ((2ull << ip) - 1);      // lanes 0..ip
      if (not_ready & need) {
        if (++spins > kSpinLimit) {
          if (lane == 0) atomicOr(A.err, ERR_LOOKBACK);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      // ordered reduction: lane i holds tile top_tile-i; earlier tiles are
      // at higher lanes, so x_i = x_{i+d} (earlier) + x_i
      Agg x;
      if (lane < ip) x = unpack_agg(wa);
      else if (lane == ip) x = in_gate ? unpack_inc(wi) : Agg{0, 0, 0, 0, 0};
      else x = Agg{0, 0, 0, 0, 0};
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        Agg y = agg_shfl_down_ordered(x, d, lane);
        x = agg_combine(y, x);
      }
      x = agg_shfl(x, 0);
      acc = agg_combine(x, acc);
      if (ip < 64) break;
      top_tile -= 64;  // no inclusive prefix in this window: look further back
    }
    excl = acc;
    if (lane == 0)
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(agg_combine(excl, tile_agg)), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  TileSum r;
  r.incl = agg_combine(excl, tile_agg);
  r.carry = agg_combine(excl, lane_excl);
  r.tile_out0 = excl.cnt;
  r.e_first = e_first;
  return r;
}

// Emit: dedup against the carry; each lane stores its run of the tile's
// output straight into the gate's output region.  SAIF durations are added
// to sacc (lane 0), which the caller flushes per gate.
__device__ __forceinline__ void emit_tile(const LevelArgs& A, const TileDesc& D, const Lane& ln,
                                          const TileSum& T, unsigned long long (&sacc)[3],
                                          unsigned long long& wave_ev) {
  const int lane = threadIdx.x;
  const uint32_t ti = D.ti;
  const uint32_t sp = ln.sp;
  const uint32_t* st = ln.st;
  const Agg& carry = T.carry;
  const Agg& incl = T.incl;
  const uint32_t e_first = T.e_first;
  const uint64_t gbase = A.out_base[D.g];
  const uint32_t tile_out0 = T.tile_out0;
  const unsigned long long w_start = A.win[2 * D.job], w_end = A.win[2 * D.job + 1];
  const uint32_t s32 = (uint32_t)min(w_start, 0x7FFFFFFFull);
  const uint32_t e32 = (uint32_t)min(w_end, 0x7FFFFFFFull);
  uint32_t pvv = carry.has ? carry.lv : 4u;
  uint32_t pct = max(min(carry.has ? carry.lt : (e_first >> 2), e32), s32);
  const uint32_t ct0 = pct;
  uint32_t a0 = 0, a1 = 0;
  {
    uint32_t w = carry.has ? carry.cnt - tile_out0 : 0u;  // lane's first slot in the tile output
    for (uint32_t i = 0; i < sp; ++i) {
      const uint32_t e = st[i];
      const uint32_t v = e & 3u;
      const uint32_t ct = max(min(e >> 2, e32), s32);
      const uint32_t len = ct - pct;
      a0 += pvv == 0u ? len : 0u;
      a1 += pvv == 3u ? len : 0u;
      if (v != pvv) {
        A.pool[gbase + tile_out0 + w] = e;
        ++w;
      }
      pvv = v;
      pct = ct;
    }
  }
  // SAIF durations and event count of the tile, reduced over the wavefront
  unsigned long long acc[3] = {a0, a1, (unsigned long long)(pct - ct0 - a0 - a1)};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
  unsigned long long ev = ln.nown;
  for (int o = 32; o > 0; o >>= 1) ev += __shfl_down(ev, o, 64);
  if (lane == 0) {
    const uint32_t ntile_g = (D.nseg + kLanes - 1) / kLanes;
    if (ti + 1 == ntile_g) {  // last tile: waveform length and the tail duration
      A.wlen[D.out_net] = incl.cnt;
      if (incl.has)
        saif_acc(acc, incl.lv, incl.lt, w_end > incl.lt ? w_end : incl.lt, w_start, w_end);
    }
    if (ti == 0) A.woff[D.out_net] = gbase;
    for (int k = 0; k < 3; ++k) sacc[k] += acc[k];
    wave_ev += ev;
  }
}

// Add a wavefront's SAIF durations of one net to the counters (lane 0).
__device__ __forceinline__ void flush_saif(const LevelArgs& A, uint32_t net,
                                           unsigned long long (&sacc)[3]) {
  if (threadIdx.x == 0) {
    unsigned long long* sa = A.saif + 3ull * net;
    for (int k = 0; k < 3; ++k)
      if (sacc[k]) atomicAdd(sa + k, sacc[k]);
  }
  sacc[0] = sacc[1] = sacc[2] = 0;
}

// First tile of the next block of A.blk consecutive tiles.  Blocks are
// handed out in increasing order, so a tile only ever waits for tiles that
// running wavefronts hold; one device-scope atomic per block instead of per
// tile keeps the counter from serialising the grid.
__device__ __forceinline__ uint64_t take_block(const LevelArgs& A) {
  unsigned int t = 0;
  if (threadIdx.x == 0) t = atomicAdd(A.ticket, 1u);
  return (uint64_t)__builtin_amdgcn_readfirstlane(__shfl(t, 0, 64)) * A.blk;
}

// Persistent, software-pipelined level kernel: while tile n is combined and
// stored, the inputs of tile n+1 are already on their way into LDS.  Every
// wavefront exits when the ticket passes the last tile.
__global__ __launch_bounds__(kLanes) void k_level(LevelArgs A) {
  // dynamic LDS: inputs (cap + 128 words), stacks (cap + 64), truth table
  // (64 words), delay table (512 words)
  extern __shared__ uint32_t s_dyn[];
  uint32_t* s_in = s_dyn;
  uint32_t* s_st = s_dyn + A.cap + 2 * kLanes;
  uint32_t* s_lut = s_st + A.cap + kLanes;
  uint32_t* s_dt = s_lut + 64;
  const uint8_t* lut8 = (const uint8_t*)s_lut;
  const uint16_t* dt16 = (const uint16_t*)s_dt;
  const int lane = threadIdx.x;
  unsigned long long wave_ev = 0;  // gate events of this wavefront (lane 0)
  unsigned long long sacc[3] = {0, 0, 0};
  const uint64_t ntiles = *A.ntiles;  // uniform: one scalar load
  uint64_t tk = take_block(A), blk_end = tk + A.blk;
  bool fits = false;
  if (tk < ntiles) {
    fits = tile_words(A.tiles[tk]) <= A.cap;
    stage_tile(A, A.tiles[tk], fits, s_in, s_lut, s_dt);
  }
  bool have_prev = false;
  uint64_t prev_tk = 0;
  Agg prev_incl = Agg{0, 0, 0, 0, 0};
  while (tk < ntiles) {
    const TileDesc& D = A.tiles[tk];
    wait_vm();
    if (fits) {  // INF sentinel after every input range
      uint32_t pos = 0;
      for (uint32_t j = 0; j < D.k; ++j) {
        pos += D.E[j] - range_start(D.L[j]);
        if ((uint32_t)lane == j) s_in[pos] = 0xFFFFFFFFu;
        ++pos;
      }
    }
    __syncthreads();
    Lane ln;
    switch (D.k) {
      case 1: ln = walk_k<1>(A, D, tk, fits, s_in, s_st, lut8, dt16); break;
      case 2: ln = walk_k<2>(A, D, tk, fits, s_in, s_st, lut8, dt16); break;
      case 3: ln = walk_k<3>(A, D, tk, fits, s_in, s_st, lut8, dt16); break;
      default: ln = walk_k<4>(A, D, tk, fits, s_in, s_st, lut8, dt16); break;
    }
    __syncthreads();  // inputs, truth and delay tables no longer read
    const bool chain = have_prev && prev_tk + 1 == tk && D.ti > 0;
    // the look-back may wait for other wavefronts: it runs before this
    // wavefront takes its next tile, so no tile is ever held by a waiting one
    const TileSum T = combine_tile(A, D, tk, ln, chain, prev_incl);
    uint64_t tn = tk + 1;
    bool fits_n = false;
    if (A.prefetch) {  // stage the next tile while this one is stored
      if (tn >= blk_end) {
        tn = take_block(A);
        blk_end = tn + A.blk;
      }
      if (tn < ntiles) {
        fits_n = tile_words(A.tiles[tn]) <= A.cap;
        stage_tile(A, A.tiles[tn], fits_n, s_in, s_lut, s_dt);
      }
    }
    emit_tile(A, D, ln, T, sacc, wave_ev);
    prev_incl = T.incl;
    if (!A.prefetch) {
      if (tn >= blk_end) {
        tn = take_block(A);
        blk_end = tn + A.blk;
      }
      if (tn < ntiles) {
        fits_n = tile_words(A.tiles[tn]) <= A.cap;
        stage_tile(A, A.tiles[tn], fits_n, s_in, s_lut, s_dt);
      }
    }
    if (tn >= ntiles || A.tiles[tn].out_net != D.out_net) flush_saif(A, D.out_net, sacc);
    have_prev = true;
    prev_tk = tk;
    __syncthreads();
    tk = tn;
    fits = fits_n;
  }
  if (lane == 0 && wave_ev) atomicAdd(A.ev_count, wave_ev);
}

// per gate: leader, stride, segments, tiles, output bound
__global__ void k_plan2(const GateRec* __restrict__ gates, uint32_t ng,
                        const uint32_t* __restrict__ wlen, uint32_t seg_target,
                        GatePlan* __restrict__ plan, uint64_t* __restrict__ ntile,
                        uint64_t* __restrict__ bound, unsigned int* __restrict__ err) {
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
  // output counts travel in 30-bit fields of the look-back words
  if (N + 1 >= (1ull << 30)) atomicOr(err, ERR_RANGE);
  // leader events per segment so that a segment holds ~seg_target input events
  uint64_t s = N > 0 ? (uint64_t)seg_target * nl / N : 1;
  if (s < 1) s = 1;
  uint32_t ns = nl > 0 ? (uint32_t)((nl + s - 1) / s) : 1u;
  GatePlan P;
  P.nseg = ns;
  P.stride = (uint32_t)s;
  P.lead = lead;
  P.pad = 0;
  plan[g] = P;
  ntile[g] = (ns + kLanes - 1) / kLanes;
  bound[g] = N + 1;
}

// per (tile, input): the tile's descriptor (gate, tile index within the
// gate, segmentation, tables) and input j's range
__global__ void k_tile_desc(const GateRec* __restrict__ gates, uint32_t ng,
                            const GatePlan* __restrict__ plan,
                            const uint64_t* __restrict__ tile_base,
                            const uint64_t* __restrict__ woff, const uint32_t* __restrict__ wlen,
                            const uint32_t* __restrict__ pool, const uint32_t* __restrict__ lutp,
                            TileDesc* __restrict__ tiles, const uint64_t* __restrict__ ntiles) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t t = x >> 2;
  const int j = (int)(x & 3);
  if (t >= *ntiles) return;
  // gate of tile t: binary search over tile_base
  uint32_t lo = 0, hi = ng;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (tile_base[mid] <= t) lo = mid;
    else hi = mid;
  }
  const uint32_t g = lo;
  const GateRec& G = gates[g];
  const GatePlan P = plan[g];
  const uint32_t ti = (uint32_t)(t - tile_base[g]);
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
    D.pad = 0;
  }
  if (j >= G.k) {
    D.L[j] = 0;
    D.E[j] = 0;
    D.in_off[j] = 0;
    return;
  }
  const uint32_t* lw = pool + woff[G.in_net[P.lead]];
  const uint32_t* iw = pool + woff[G.in_net[j]];
  const uint32_t n = wlen[G.in_net[j]];
  const uint32_t c0 = ti * kLanes;
  const uint32_t clast = min(c0 + kLanes - 1, P.nseg - 1);
  uint32_t L = 0, E = n;
  if (c0 > 0) {
    const uint32_t tau_lo = lw[c0 * P.stride] >> 2;
    L = (uint32_t)j == P.lead ? c0 * P.stride : lb_words(iw, n, tau_lo << 2);
  }
  if (clast + 1 < P.nseg) {
    const uint32_t tau_hi = (lw[(clast + 1) * P.stride] >> 2) + G.dmax;
    E = lb_words(iw, n, tau_hi << 2);
  }
  D.L[j] = L;
  D.E[j] = E;
  D.in_off[j] = woff[G.in_net[j]] + (L > 0 ? L - 1 : 0u);
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_plan2(const GateRec* gates, uint32_t ng, const uint32_t* wlen,
                        uint32_t seg_target, GatePlan* plan, uint64_t* ntile, uint64_t* bound,
                        unsigned int* err, hipStream_t st) {
  k_plan2<<<(ng + 255) / 256, 256, 0, st>>>(gates, ng, wlen, seg_target, plan, ntile, bound, err);
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

// Device-driven levels: when the tile count does not fit the tile arrays, or
// the level's output reservations do not fit the pool, the level is skipped
// (tile count 0) and the error flag says why.
__global__ void k_guard(uint64_t* scal, uint64_t tiles_cap, uint64_t pool_words) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // scal: 0 bump, 1 tiles of the level, 2 output words reserved, 7 error flags
  if (scal[1] > tiles_cap || scal[0] + scal[2] > pool_words) {
    scal[1] = 0;
    scal[2] = 0;
    atomicOr((unsigned int*)(scal + 7), ERR_CAPACITY);
  }
}

hipError_t launch_guard(uint64_t* scal, uint64_t tiles_cap, uint64_t pool_words, hipStream_t st) {
  k_guard<<<1, 64, 0, st>>>(scal, tiles_cap, pool_words);
  return hipGetLastError();
}

// p[i] = v for i < mult * (*n)
__global__ void k_fill_u64_dev(unsigned long long* p, const uint64_t* n, uint64_t mult,
                               uint64_t cap, unsigned long long v) {
  const uint64_t lim = min(cap, mult * *n);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lim;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t launch_fill_u64_dev(unsigned long long* p, const uint64_t* n, uint64_t mult, uint64_t cap,
                               unsigned long long v, hipStream_t st) {
  if (cap == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((cap + 255) / 256, 4096);
  k_fill_u64_dev<<<(unsigned)blocks, 256, 0, st>>>(p, n, mult, cap, v);
  return hipGetLastError();
}

// Start-of-run scalars: bump pointer past the stimulus words, all else 0.
__global__ void k_set_scal(uint64_t* scal, uint64_t bump) {
  if (blockIdx.x == 0 && threadIdx.x < 16) scal[threadIdx.x] = threadIdx.x == 0 ? bump : 0;
}

hipError_t launch_set_scal(uint64_t* scal, uint64_t bump, hipStream_t st) {
  k_set_scal<<<1, 64, 0, st>>>(scal, bump);
  return hipGetLastError();
}

hipError_t launch_level(const LevelArgs& a, unsigned nblocks, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_level, dim3(nblocks), dim3(kLanes), level_lds_bytes(a.cap), st, a);
  return hipGetLastError();
}

int level_max_blocks(uint32_t cap) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 2048;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_level, kLanes,
                                                   level_lds_bytes(cap)) != hipSuccess)
    per_cu = 4;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  return per_cu * cus;
}

}  // namespace gls
