// Fused per-level gate evaluation kernel (gfx950).
//
// Work unit: a tile = 64 consecutive time segments of one gate, one segment
// per lane of a wavefront.  Segments are cut at evenly spaced events of the
// gate's longest input (the "leader"), so their boundaries are exact event
// times.  Per tile:
//   1. the input waveform ranges the tile needs (plus one event before the
//      tile for the starting values and a D_max look-ahead halo after it)
//      are copied into LDS with coalesced loads;
//   2. every lane merges its segment's input events in time order, evaluates
//      the 4-state truth table, classifies the output edge, picks the SDF
//      delay (AddDelay rules folded into per-gate tables) and applies the
//      preemption rule as a strict suffix-min filter (stack in LDS); events
//      in the halo only pre-empt;
//   3. a wavefront scan combines the lanes' kept lists (dedup of equal
//      adjacent values across lanes), a decoupled look-back over the
//      gate's earlier tiles supplies the carry (last value, output count),
//      and the tile's dense output is written with coalesced stores from
//      LDS;
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
constexpr int kCapIn = 2560;            // staged input words per tile
constexpr int kCapSt = kCapIn + kLanes;  // stack words per tile
constexpr uint32_t kSpinLimit = 1u << 24;

enum : unsigned { ERR_LOOKBACK = 1u, ERR_OVERFLOW = 2u };

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

// aggregate word: valid<<63 | lt<<21 | has<<20 | lv<<18 | fv<<16 | cnt(16)
__device__ __forceinline__ unsigned long long pack_agg(const Agg& a) {
  return (1ull << 63) | ((unsigned long long)a.lt << 21) | ((unsigned long long)a.has << 20) |
         ((unsigned long long)a.lv << 18) | ((unsigned long long)a.fv << 16) | (a.cnt & 0xFFFFu);
}
__device__ __forceinline__ Agg unpack_agg(unsigned long long w) {
  Agg a;
  a.cnt = (uint32_t)(w & 0xFFFFu);
  a.fv = (uint32_t)(w >> 16) & 3u;
  a.lv = (uint32_t)(w >> 18) & 3u;
  a.has = (uint32_t)(w >> 20) & 1u;
  a.lt = (uint32_t)(w >> 21) & 0x3FFFFFFFu;
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

template <int K, bool LDS>
__device__ __forceinline__ void run_tile(const LevelArgs& A, const GateRec& G, uint32_t g,
                                         uint64_t ti, uint32_t* s_in, uint32_t* s_st,
                                         const uint8_t* s_lut, const uint16_t* s_tmin,
                                         const uint16_t* s_tone, const uint32_t* const (&src_base)[K],
                                         const uint32_t (&Bv)[K], const uint32_t (&Ev)[K],
                                         const uint32_t (&Lv)[K], uint32_t* stk_base, int lead,
                                         uint32_t nl, uint32_t nseg, uint64_t tk) {
  const int lane = threadIdx.x;
  const uint32_t c0 = (uint32_t)(ti * kLanes);
  const uint32_t c = c0 + lane;
  const bool active = c < nseg;
  const uint32_t kX = 1u;
  const uint32_t INF = 0xFFFFFFFFu;

  const uint32_t* src[K];
  uint32_t bound[K], r[K], cur[K], h[K];
  // the leader's array and base as selects (no dynamic register indexing)
  const uint32_t* srcL = src_base[0];
  uint32_t BL = Bv[0];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    src[j] = src_base[j];
    bound[j] = Ev[j] - Bv[j];
    if (j == lead) {
      srcL = src_base[j];
      BL = Bv[j];
    }
  }
  uint32_t sp = 0, top = 0, nown = 0, hmin = INF;
  uint32_t* st = stk_base;
  if (active) {
    // segment bounds from the leader
    uint32_t tau_c = 0, tau_n = 0x3FFFFFFFu, halo_end = 0x3FFFFFFFu;
    uint32_t pl = (uint32_t)((uint64_t)c * nl / nseg);
    if (c > 0) tau_c = srcL[pl - BL] >> 2;
    if (c + 1 < nseg) {
      uint32_t pn = (uint32_t)((uint64_t)(c + 1) * nl / nseg);
      tau_n = srcL[pn - BL] >> 2;
      halo_end = tau_n + G.dmax;
    }
    uint32_t soff = lane;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      uint32_t pj;
      if (c == 0) pj = 0;
      else if (j == lead) pj = pl;
      else pj = Bv[j] + lb_words(src[j], bound[j], tau_c << 2);
      soff += pj - Lv[j];
      r[j] = pj - Bv[j];
      cur[j] = pj > 0 ? (src[j][r[j] - 1] & 3u) : kX;
      h[j] = r[j] < bound[j] ? src[j][r[j]] : INF;
    }
    st = stk_base + soff;
    uint32_t li = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) li |= cur[j] << (2 * j);
    uint32_t fprev = c == 0 ? kX : s_lut[li];
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
          h[j] = r[j] < bound[j] ? src[j][r[j]] : INF;
        }
        idx |= cur[j] << (2 * j);
      }
      const uint32_t f = s_lut[idx];
      const uint32_t sc = edge_cls(fprev, f);
      fprev = f;
      const bool own = t < tau_n;
      nown += own ? 1u : 0u;
      if (sc == 0) continue;
      uint32_t d;
      if (__popc(mask) == 1) {
        d = 1;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          if (mask == (1u << j)) {
            const uint32_t fall = edge_cls(pv[j], cur[j]) == 2 ? 1u : 0u;
            d = s_tone[(j * 4 + sc - 1) * 2 + fall];
          }
        }
      } else {
        uint32_t m = 0xFFFFu;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          if ((mask >> j) & 1u) {
            const uint32_t fall = edge_cls(pv[j], cur[j]) == 2 ? 1u : 0u;
            m = min(m, (uint32_t)s_tmin[(j * 4 + sc - 1) * 2 + fall]);
          }
        }
        d = m == 0xFFFFu ? 1u : m;
      }
      const uint32_t o = t + d;
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
      mine.lv = st[sp - 1] & 3u;
      mine.lt = st[sp - 1] >> 2;
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

  // decoupled look-back over the gate's earlier tiles
  unsigned long long* ts = A.tstate;
  Agg excl;
  excl.has = 0;
  excl.cnt = 0;
  excl.fv = excl.lv = excl.lt = 0;
  if (lane == 0) {
    if (ti == 0) {
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(tile_agg), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(&ts[2 * tk], pack_agg(tile_agg), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      Agg acc;
      acc.has = 0;
      acc.cnt = 0;
      acc.fv = acc.lv = acc.lt = 0;
      uint64_t p = tk - 1;
      uint32_t spins = 0;
      for (;;) {
        unsigned long long wi =
            __hip_atomic_load(&ts[2 * p + 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (wi >> 63) {
          acc = agg_combine(unpack_inc(wi), acc);
          break;
        }
        unsigned long long wa =
            __hip_atomic_load(&ts[2 * p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (wa >> 63) {
          acc = agg_combine(unpack_agg(wa), acc);
          --p;  // the gate's first tile always publishes an inclusive prefix
          continue;
        }
        if (++spins > kSpinLimit) {
          atomicOr(A.err, ERR_LOOKBACK);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      excl = acc;
      __hip_atomic_store(&ts[2 * tk + 1], pack_inc(agg_combine(excl, tile_agg)), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  excl = agg_shfl(excl, 0);
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
    uint32_t w = carry.cnt - tile_out0;  // lane's first slot in the tile output
    if (!carry.has) w = 0;
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
    const uint64_t ntile_g = (nseg + kLanes - 1) / kLanes;
    if (ti + 1 == ntile_g) {  // last tile: waveform length and the tail duration
      A.wlen[G.out_net] = incl.cnt;
      if (incl.has) saif_acc(acc, incl.lv, incl.lt, A.end > incl.lt ? A.end : incl.lt, A.start, A.end);
    }
    if (ti == 0) A.woff[G.out_net] = gbase;
    for (int k = 0; k < 3; ++k)
      if (acc[k]) atomicAdd(sa + k, acc[k]);
    if (ev) atomicAdd(A.ev_count, ev);
  }
}

template <int K>
__device__ __forceinline__ void tile_k(const LevelArgs& A, const GateRec& G, uint32_t g, uint64_t ti, uint64_t tk,
                       uint32_t* s_in, uint32_t* s_st, const uint8_t* s_lut,
                       const uint16_t* s_tmin, const uint16_t* s_tone) {
  const int lane = threadIdx.x;
  const uint32_t nseg = A.gnseg[g];
  const uint32_t* gin[K];
  uint32_t n[K];
  int lead = 0;
  uint32_t nl = 0;
  const uint32_t* gL = nullptr;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    gin[j] = A.pool + A.woff[G.in_net[j]];
    n[j] = A.wlen[G.in_net[j]];
    if (j == 0 || n[j] > nl) {
      lead = j;
      nl = n[j];
      gL = gin[j];
    }
  }
  const uint32_t c0 = (uint32_t)(ti * kLanes);
  const uint32_t clast = min(c0 + kLanes - 1, nseg - 1);
  const bool has_hi = clast + 1 < nseg;
  const uint32_t tau_lo = c0 == 0 ? 0u : (gL[(uint64_t)c0 * nl / nseg] >> 2);
  const uint32_t tau_hi = has_hi ? (gL[(uint64_t)(clast + 1) * nl / nseg] >> 2) + G.dmax : 0u;
  // lanes 0..K-1 find the tile start in input j, lanes K..2K-1 the end
  const int jl = lane < K ? lane : lane - K;
  const uint32_t* gj = gin[0];
  uint32_t nj = n[0];
#pragma unroll
  for (int j = 1; j < K; ++j)
    if (jl == j) {
      gj = gin[j];
      nj = n[j];
    }
  uint32_t myv = 0;
  if (lane < K) myv = c0 == 0 ? 0u : lb_words(gj, nj, tau_lo << 2);
  else if (lane < 2 * K) myv = has_hi ? lb_words(gj, nj, tau_hi << 2) : nj;
  uint32_t Bv[K], Ev[K], Lv[K];
  uint32_t total = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    Lv[j] = __shfl(myv, j, 64);
    Ev[j] = __shfl(myv, K + j, 64);
    Bv[j] = Lv[j] > 0 ? Lv[j] - 1 : 0;
    total += Ev[j] - Bv[j];
  }
  if (total <= (uint32_t)kCapIn) {
    const uint32_t* src[K];
    uint32_t off = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const uint32_t len = Ev[j] - Bv[j];
      const uint32_t* s = gin[j] + Bv[j];
      for (uint32_t q = lane; q < len; q += kLanes) s_in[off + q] = s[q];
      src[j] = s_in + off;
      off += len;
    }
    __syncthreads();
    run_tile<K, true>(A, G, g, ti, s_in, s_st, s_lut, s_tmin, s_tone, src, Bv, Ev, Lv, s_st, lead,
                      nl, nseg, tk);
  } else {
    // does not fit LDS: read inputs from HBM, stacks in the overflow buffer
    unsigned long long base = 0;
    const unsigned long long need = (unsigned long long)total + kLanes;
    if (lane == 0) base = atomicAdd(A.ovf_used, need);
    base = __shfl(base, 0, 64);
    const uint32_t* src[K];
#pragma unroll
    for (int j = 0; j < K; ++j) src[j] = gin[j] + Bv[j];
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
    run_tile<K, false>(A, G, g, ti, s_in, s_st, s_lut, s_tmin, s_tone, src, Bv, Ev, Lv,
                       A.ovf + base, lead, nl, nseg, tk);
  }
}

__global__ __launch_bounds__(kLanes) void k_level(LevelArgs A) {
  __shared__ uint32_t s_in[kCapIn + kLanes];
  __shared__ uint32_t s_st[kCapSt];
  __shared__ uint8_t s_lut[256];
  __shared__ uint16_t s_tab[64];
  const int lane = threadIdx.x;
  for (;;) {
    unsigned int t = 0;
    if (lane == 0) t = atomicAdd(A.ticket, 1u);
    const uint64_t tk = (uint64_t)__builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
    if (tk >= A.ntiles) break;
    const uint32_t g = __builtin_amdgcn_readfirstlane(A.tilemap[tk]);
    const GateRec& G = A.gates[g];
    const uint64_t ti = tk - A.tile_base[g];
    // truth table and delay tables of this gate into LDS
    const uint32_t* lsrc = (const uint32_t*)(A.luts + (uint32_t)G.lut * 256u);
    ((uint32_t*)s_lut)[lane] = lsrc[lane];
    if (lane < 32) {
      const uint32_t* tb = (const uint32_t*)&G.tmin[0][0][0];
      ((uint32_t*)s_tab)[lane] = tb[lane];
    }
    __syncthreads();
    switch (G.k) {
      case 1: tile_k<1>(A, G, g, ti, tk, s_in, s_st, s_lut, s_tab, s_tab + 32); break;
      case 2: tile_k<2>(A, G, g, ti, tk, s_in, s_st, s_lut, s_tab, s_tab + 32); break;
      case 3: tile_k<3>(A, G, g, ti, tk, s_in, s_st, s_lut, s_tab, s_tab + 32); break;
      default: tile_k<4>(A, G, g, ti, tk, s_in, s_st, s_lut, s_tab, s_tab + 32); break;
    }
    __syncthreads();
  }
}

// per gate: segments, tiles, output bound
__global__ void k_plan2(const GateRec* __restrict__ gates, uint32_t ng,
                        const uint32_t* __restrict__ wlen, uint32_t seg_target,
                        uint64_t* __restrict__ ntile, uint64_t* __restrict__ bound,
                        uint32_t* __restrict__ gnseg) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const GateRec& G = gates[g];
  uint64_t N = 0;
  uint32_t nl = 0;
  for (int j = 0; j < G.k; ++j) {
    uint32_t n = wlen[G.in_net[j]];
    N += n;
    nl = max(nl, n);
  }
  uint64_t ns = N / seg_target + 1;
  if (ns > nl) ns = nl > 0 ? nl : 1;
  gnseg[g] = (uint32_t)ns;
  ntile[g] = (ns + kLanes - 1) / kLanes;
  bound[g] = N + 1;
}

__global__ void k_tilemap(uint32_t ng, const uint64_t* __restrict__ tile_base,
                          const uint64_t* __restrict__ ntile, uint32_t* __restrict__ tilemap) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  uint64_t b = tile_base[g];
  for (uint64_t i = 0; i < ntile[g]; ++i) tilemap[b + i] = g;
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_plan2(const GateRec* gates, uint32_t ng, const uint32_t* wlen,
                        uint32_t seg_target, uint64_t* ntile, uint64_t* bound, uint32_t* gnseg,
                        hipStream_t st) {
  k_plan2<<<(ng + 255) / 256, 256, 0, st>>>(gates, ng, wlen, seg_target, ntile, bound, gnseg);
  return hipGetLastError();
}

hipError_t launch_tilemap(uint32_t ng, const uint64_t* tile_base, const uint64_t* ntile,
                          uint32_t* tilemap, hipStream_t st) {
  k_tilemap<<<(ng + 255) / 256, 256, 0, st>>>(ng, tile_base, ntile, tilemap);
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
  hipLaunchKernelGGL(k_level, dim3(nblocks), dim3(kLanes), 0, st, a);
  return hipGetLastError();
}

int level_max_blocks() {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 2048;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_level, kLanes, 0) != hipSuccess)
    per_cu = 4;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  return per_cu * cus;
}

}  // namespace gls
