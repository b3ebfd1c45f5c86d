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
//   2. every lane merges its segment's input events in time order, evaluates
//      the 4-state truth table (packed in registers for fan-in <= 3),
//      classifies the output edge, picks the SDF delay (AddDelay folded into
//      per-gate tables, the common rise/fall cases in registers) and applies
//      the preemption rule as a strict suffix-min filter (stack in LDS);
//      events in the halo only pre-empt;
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

#include <cstdint>

#include "device.h"

namespace gls {

constexpr int kLanes = 64;
constexpr uint32_t kSpinLimit = 1u << 24;

enum : unsigned { ERR_LOOKBACK = 1u, ERR_OVERFLOW = 2u, ERR_RANGE = 4u };

__device__ __forceinline__ uint32_t edge_cls(uint32_t pre, uint32_t cur) {
  // CheckDelay (src/simulate.cpp:1178-1201); codes 0=0 1=X 2=Z 3=1
  const uint64_t T = 0x0222ull << 48 | 0x1042ull << 32 | 0x1302ull << 16 | 0x1110ull;
  return (uint32_t)(T >> ((pre * 4 + cur) * 4)) & 0xF;
}

// input j fell at this time (CheckDelay(last, cur) == 2)
__device__ __forceinline__ uint32_t fell(uint32_t pre, uint32_t cur) {
  return edge_cls(pre, cur) == 2 ? 1u : 0u;
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
  uint64_t tmin2[K], tone2[K];  // delay tables for rise/fall output edges
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

// AddDelay (src/simulate.cpp:1012-1176) on the folded tables.
template <int K>
__device__ __forceinline__ uint32_t delay_of(const TileCtx<K>& X, const uint16_t* s_tmin,
                                             const uint16_t* s_tone, uint32_t mask, uint32_t sc,
                                             const uint32_t (&pv)[K], const uint32_t (&cur)[K]) {
  const bool fast = sc <= 2;  // rise / fall: tables in registers
  if (__popc(mask) == 1) {
    uint32_t d = 1;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      if (mask == (1u << j)) {
        const uint32_t f = fell(pv[j], cur[j]);
        const uint32_t q = (sc - 1) * 2 + f;
        d = fast ? (uint32_t)(X.tone2[j] >> (16 * q)) & 0xFFFFu : s_tone[j * 8 + q];
      }
    }
    return d;
  }
  uint32_t m = 0xFFFFu;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if ((mask >> j) & 1u) {
      const uint32_t f = fell(pv[j], cur[j]);
      const uint32_t q = (sc - 1) * 2 + f;
      const uint32_t v = fast ? (uint32_t)(X.tmin2[j] >> (16 * q)) & 0xFFFFu : s_tmin[j * 8 + q];
      m = min(m, v);
    }
  }
  return m == 0xFFFFu ? 1u : m;
}

template <int K, bool LDS>
__device__ __forceinline__ void run_tile(const LevelArgs& A, const GateRec& G, uint32_t g,
                                         uint32_t ti, uint64_t tk, const TileCtx<K>& X,
                                         uint32_t* s_in, const uint8_t* s_lut,
                                         const uint16_t* s_tmin, const uint16_t* s_tone,
                                         uint32_t* stk_base, unsigned long long& wave_ev) {
  const int lane = threadIdx.x;
  const uint32_t c = ti * kLanes + lane;
  const bool active = c < X.nseg;
  const uint32_t kX = 1u;
  const uint32_t INF = 0xFFFFFFFFu;

  uint32_t r[K], cur[K], h[K], bound[K];
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
  uint32_t sp = 0, top = 0, nown = 0, hmin = INF;
  uint32_t* st = stk_base;
  if (active) {
    // segment bounds from the leader
    uint32_t tau_c = 0, tau_n = 0x3FFFFFFFu, halo_end = 0x3FFFFFFFu;
    const uint32_t pl = c * X.stride;
    if (c > 0) tau_c = srcL[pl - BL] >> 2;
    if (c + 1 < X.nseg) {
      tau_n = srcL[pl + X.stride - BL] >> 2;
      halo_end = tau_n + X.dmax;
    }
    uint32_t soff = lane;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      uint32_t pj;
      if (c == 0) pj = 0;
      else if ((uint32_t)j == X.lead) pj = pl;
      else pj = X.B[j] + lb_words(X.src[j], bound[j], tau_c << 2);
      soff += pj - X.L[j];
      r[j] = pj - X.B[j];
      cur[j] = pj > 0 ? (X.src[j][r[j] - 1] & 3u) : kX;
      h[j] = r[j] < bound[j] ? X.src[j][r[j]] : INF;
    }
    st = stk_base + soff;
    uint32_t li = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) li |= cur[j] << (2 * j);
    uint32_t fprev = c == 0 ? kX : lut_eval<K>(X, s_lut, li);
    if (c == 0) {  // every output starts with (0, X) (src/simulate.cpp:882-887)
      top = kX;
      st[sp++] = top;
    }
    for (;;) {
      uint32_t w = h[0];
#pragma unroll
      for (int j = 1; j < K; ++j) w = min(w, h[j]);
      const uint32_t t = w >> 2;
      if (w == INF || t >= halo_end) break;
      const uint32_t lim = w | 3u;
      uint32_t pv[K], mask = 0, idx = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        pv[j] = cur[j];
        if (h[j] <= lim) {
          cur[j] = h[j] & 3u;
          mask |= 1u << j;
          ++r[j];
          h[j] = r[j] < bound[j] ? X.src[j][r[j]] : INF;
        }
        idx |= cur[j] << (2 * j);
      }
      const uint32_t f = lut_eval<K>(X, s_lut, idx);
      const uint32_t sc = edge_cls(fprev, f);
      fprev = f;
      const bool own = t < tau_n;
      nown += own ? 1u : 0u;
      if (sc == 0) continue;
      const uint32_t o = t + delay_of<K>(X, s_tmin, s_tone, mask, sc, pv, cur);
      if (own) {
        while (sp > 0 && (top >> 2) >= o) {
          --sp;
          top = sp ? st[sp - 1] : 0u;
        }
        top = (o << 2) | f;
        st[sp++] = top;
      } else {
        hmin = min(hmin, o);
      }
    }
    while (sp > 0 && (top >> 2) >= hmin) {
      --sp;
      top = sp ? st[sp - 1] : 0u;
    }
  }

  // lane aggregate over its kept candidates
  Agg mine;
  mine.cnt = 0;
  mine.has = sp > 0 ? 1u : 0u;
  mine.fv = mine.lv = mine.lt = 0;
  {
    uint32_t pvv = 4;
    for (uint32_t i = 0; i < sp; ++i) {
      uint32_t e = st[i];
      if ((e & 3u) != pvv) mine.cnt++;
      pvv = e & 3u;
    }
    if (sp) {
      mine.fv = st[0] & 3u;
      mine.lv = top & 3u;
      mine.lt = top >> 2;
    }
  }
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
  // as an empty inclusive prefix)
  unsigned long long* ts = A.tstate;
  Agg excl;
  excl.has = 0;
  excl.cnt = 0;
  excl.fv = excl.lv = excl.lt = 0;
  if (ti == 0) {
    if (lane == 0)
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(tile_agg), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
  } else {
    if (lane == 0)
      __hip_atomic_store(&ts[2 * tk], pack_agg(tile_agg), __ATOMIC_RELEASE,
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
        wi = __hip_atomic_load(&ts[2 * p + 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (!(wi >> 63))
          wa = __hip_atomic_load(&ts[2 * p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
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
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(agg_combine(excl, tile_agg)), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  const Agg incl = agg_combine(excl, tile_agg);

  // emit: dedup against the carry, stage in LDS, coalesced store
  const Agg carry = agg_combine(excl, lane_excl);
  const uint64_t gbase = A.out_base[g];
  const uint32_t tile_out0 = excl.cnt;
  uint32_t pvv = carry.has ? carry.lv : 4u;
  uint32_t pt = carry.lt;
  unsigned long long acc[3] = {0, 0, 0};
  __syncthreads();
  {
    uint32_t w = carry.has ? carry.cnt - tile_out0 : 0u;  // lane's first slot in the tile output
    for (uint32_t i = 0; i < sp; ++i) {
      const uint32_t e = st[i];
      const uint32_t v = e & 3u, t = e >> 2;
      if (pvv != 4u) saif_acc(acc, pvv, pt, t, A.start, A.end);
      if (v != pvv) {
        if (LDS) s_in[w] = e;
        else A.pool[gbase + tile_out0 + w] = e;
        ++w;
      }
      pvv = v;
      pt = t;
    }
  }
  __syncthreads();
  const uint32_t tile_cnt = incl.cnt - tile_out0;
  if (LDS) {
    uint32_t* dst = A.pool + gbase + tile_out0;
    for (uint32_t q = lane; q < tile_cnt; q += kLanes) dst[q] = s_in[q];
  }
  // SAIF and event count: one atomic per counter per tile
#pragma unroll
  for (int k = 0; k < 3; ++k)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
  unsigned long long ev = nown;
  for (int o = 32; o > 0; o >>= 1) ev += __shfl_down(ev, o, 64);
  if (lane == 0) {
    unsigned long long* sa = A.saif + 3ull * G.out_net;
    const uint32_t ntile_g = (X.nseg + kLanes - 1) / kLanes;
    if (ti + 1 == ntile_g) {  // last tile: waveform length and the tail duration
      A.wlen[G.out_net] = incl.cnt;
      if (incl.has)
        saif_acc(acc, incl.lv, incl.lt, A.end > incl.lt ? A.end : incl.lt, A.start, A.end);
    }
    if (ti == 0) A.woff[G.out_net] = gbase;
    for (int k = 0; k < 3; ++k)
      if (acc[k]) atomicAdd(sa + k, acc[k]);
    wave_ev += ev;
  }
}

template <int K>
__device__ __forceinline__ void tile_k(const LevelArgs& A, const GateRec& G, const TileDesc& D,
                                       uint64_t tk, uint32_t* s_in, uint32_t* s_st,
                                       const uint8_t* s_lut, const uint16_t* s_tmin,
                                       const uint16_t* s_tone, unsigned long long& wave_ev) {
  const int lane = threadIdx.x;
  const uint32_t g = D.g, ti = D.ti;
  const GatePlan P = A.plan[g];
  TileCtx<K> X;
  X.nseg = P.nseg;
  X.stride = P.stride;
  X.lead = P.lead;
  X.dmax = G.dmax;
  const uint32_t* lp = A.lutp + (uint32_t)G.lut * 16u;
#pragma unroll
  for (int q = 0; q < 4; ++q) X.lut[q] = lp[q];
  const uint32_t* gin[K];
  uint32_t off[K];
  uint32_t total = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    X.tmin2[j] = *(const uint64_t*)&G.tmin[j][0][0];
    X.tone2[j] = *(const uint64_t*)&G.tone[j][0][0];
    X.L[j] = D.L[j];
    X.E[j] = D.E[j];
    X.B[j] = D.L[j] > 0 ? D.L[j] - 1 : 0u;
    gin[j] = A.pool + A.woff[G.in_net[j]] + X.B[j];
    off[j] = total;
    total += X.E[j] - X.B[j];
  }
  if (total <= A.cap) {
    // batched copy of the concatenated ranges: 8 loads in flight per lane
    for (uint32_t x0 = 0; x0 < total; x0 += 8 * kLanes) {
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t x = x0 + u * kLanes + lane;
        v[u] = 0;
        if (x < total) {
          const uint32_t* s = gin[0] + x;
#pragma unroll
          for (int j = 1; j < K; ++j)
            if (x >= off[j]) s = gin[j] + (x - off[j]);
          v[u] = *s;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t x = x0 + u * kLanes + lane;
        if (x < total) s_in[x] = v[u];
      }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) X.src[j] = s_in + off[j];
    __syncthreads();
    run_tile<K, true>(A, G, g, ti, tk, X, s_in, s_lut, s_tmin, s_tone, s_st, wave_ev);
  } else {
    // does not fit LDS: read inputs from HBM, stacks in the overflow buffer
    unsigned long long base = 0;
    const unsigned long long need = (unsigned long long)total + kLanes;
    if (lane == 0) base = atomicAdd(A.ovf_used, need);
    base = __shfl(base, 0, 64);
#pragma unroll
    for (int j = 0; j < K; ++j) X.src[j] = gin[j];
    if (base + need > A.ovf_cap) {
      if (lane == 0) atomicOr(A.err, ERR_OVERFLOW);
      base = 0;  // results are invalid; keep every wavefront moving
      if (need > A.ovf_cap) {
        // cannot even alias the buffer: publish an empty tile and stop
        if (lane == 0) {
          A.tstate[2 * tk] = 1ull << 63;
          __hip_atomic_store(&A.tstate[2 * tk + 1], 1ull << 63, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
      }
    }
    run_tile<K, false>(A, G, g, ti, tk, X, s_in, s_lut, s_tmin, s_tone, A.ovf + base, wave_ev);
  }
}

__global__ __launch_bounds__(kLanes) void k_level(LevelArgs A) {
  // dynamic LDS: inputs (cap + 64 words), stacks (cap + 64), truth table, delay tables
  extern __shared__ uint32_t s_dyn[];
  uint32_t* s_in = s_dyn;
  uint32_t* s_st = s_dyn + A.cap + kLanes;
  uint8_t* s_lut = (uint8_t*)(s_st + A.cap + kLanes);
  uint16_t* s_tab = (uint16_t*)(s_lut + 256);
  const int lane = threadIdx.x;
  unsigned long long wave_ev = 0;  // gate events of this wavefront (lane 0)
  for (;;) {
    unsigned int t = 0;
    if (lane == 0) t = atomicAdd(A.ticket, 1u);
    const uint64_t tk = (uint64_t)__builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (tk >= A.ntiles) break;
    const TileDesc& D = A.tiles[tk];
    const uint32_t g = __builtin_amdgcn_readfirstlane(D.g);
    const GateRec& G = A.gates[g];
    const uint32_t k = G.k;
    // full truth table (fan-in 4) and delay tables (X/Z edges) into LDS
    if (k == 4) {
      const uint32_t* lsrc = (const uint32_t*)(A.luts + (uint32_t)G.lut * 256u);
      ((uint32_t*)s_lut)[lane] = lsrc[lane];
    }
    if (lane < 32) {
      const uint32_t* tb = (const uint32_t*)&G.tmin[0][0][0];
      ((uint32_t*)s_tab)[lane] = tb[lane];
    }
    __syncthreads();
    switch (k) {
      case 1: tile_k<1>(A, G, D, tk, s_in, s_st, s_lut, s_tab, s_tab + 32, wave_ev); break;
      case 2: tile_k<2>(A, G, D, tk, s_in, s_st, s_lut, s_tab, s_tab + 32, wave_ev); break;
      case 3: tile_k<3>(A, G, D, tk, s_in, s_st, s_lut, s_tab, s_tab + 32, wave_ev); break;
      default: tile_k<4>(A, G, D, tk, s_in, s_st, s_lut, s_tab, s_tab + 32, wave_ev); break;
    }
    __syncthreads();
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

// per (tile, input): gate, tile index within the gate, input range
__global__ void k_tile_desc(const GateRec* __restrict__ gates, uint32_t ng,
                            const GatePlan* __restrict__ plan,
                            const uint64_t* __restrict__ tile_base,
                            const uint64_t* __restrict__ ntile, const uint64_t* __restrict__ woff,
                            const uint32_t* __restrict__ wlen, const uint32_t* __restrict__ pool,
                            TileDesc* __restrict__ tiles, uint64_t ntiles) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t t = x >> 2;
  const int j = (int)(x & 3);
  if (t >= ntiles) return;
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
  if (j == 0) {
    D.g = g;
    D.ti = ti;
    D.pad[0] = D.pad[1] = 0;
  }
  if (j >= G.k) {
    D.L[j] = 0;
    D.E[j] = 0;
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
  (void)ntile;
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
                            const uint64_t* tile_base, const uint64_t* ntile, const uint64_t* woff,
                            const uint32_t* wlen, const uint32_t* pool, TileDesc* tiles,
                            uint64_t ntiles, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  const uint64_t threads = ntiles * 4;
  k_tile_desc<<<(unsigned)((threads + 255) / 256), 256, 0, st>>>(gates, ng, plan, tile_base, ntile,
                                                                  woff, wlen, pool, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v,
                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_fill_u64<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_level(const LevelArgs& a, unsigned nblocks, hipStream_t st) {
  if (a.ntiles == 0) return hipSuccess;
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
