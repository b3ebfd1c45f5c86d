// Level kernels of the plain per-level simulator (gfx950).
//
// For each topological level:
//   k_plan         per gate: input event count N, leader input, segment count
//   scan (x2)      segment bases and scratch bases over the level's gates
//   k_segmap       segment -> gate map
//   k_eval         one lane per time segment of one gate: k-way merge of the
//                  input waveforms, 4-state truth table, edge class, SDF delay
//                  selection, o = t + d, and the preemption rule as a strict
//                  suffix-min filter (stack), with a D_max look-ahead halo into
//                  the next segment
//   k_gate_scan    per gate: cross-segment dedup carry, output length, tail of
//                  the SAIF durations
//   scan           output offsets in the event pool
//   k_emit         one lane per segment: dedup + write the dense waveform and
//                  accumulate the SAIF T0/T1/TX durations
//
// The event semantics restate src/simulate.cpp:919-1201 (Calculate,
// AddDelay, CheckDelay): see the per-function comments.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace gls {

#define GLS_DEV __device__ __forceinline__

GLS_DEV uint32_t lower_bound_w(const uint32_t* __restrict__ a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// CheckDelay (src/simulate.cpp:1178-1201): 1 rise, 2 fall, 3 X->Z, 4 Z->X,
// 0 unchanged.  Codes: 0=0, 1=X, 2=Z, 3=1.  Packed 4-bit table indexed by
// pre*4+cur.
GLS_DEV uint32_t edge_class(uint32_t pre, uint32_t cur) {
  // pre\cur:   0  X  Z  1
  //   0     :  0  1  1  1
  //   X     :  2  0  3  1
  //   Z     :  2  4  0  1
  //   1     :  2  2  2  0
  const uint64_t T = 0x0222ull << 48 | 0x1042ull << 32 | 0x1302ull << 16 | 0x1110ull;
  uint32_t idx = pre * 4 + cur;
  return (uint32_t)(T >> (idx * 4)) & 0xF;
}

// Delay select by output edge class (src/simulate.cpp:1029-1046).
GLS_DEV uint32_t sel_delay(uint32_t r, uint32_t f, uint32_t s) {
  return s == 1 ? r : (s == 2 ? f : (s == 3 ? max(r, f) : min(r, f)));
}

// AddDelay (src/simulate.cpp:1012-1176) in integer time units.
//   mask: inputs with an entry at this time (num_in = popcount)
//   prev/cur: input values before / at this time (last/cur of GetValueAt)
GLS_DEV uint32_t pick_delay(const GateDev& G, uint32_t mask, const uint32_t* prev,
                            const uint32_t* cur, uint32_t s) {
  const uint16_t NO = 0xFFFF;
  if (__popc(mask) == 1) {
    int j = __ffs(mask) - 1;
    uint32_t nr = G.nrec[j];
    if (nr == 0) return 1;
    if (nr == 1) {
      for (int kd = 0; kd < 3; ++kd)
        if (G.dly[j][kd][0] != NO) return sel_delay(G.dly[j][kd][0], G.dly[j][kd][1], s);
      return 1;
    }
    uint32_t ie = edge_class(prev[j], cur[j]);
    int kd = (ie == 2) ? 2 : 1;  // fall -> negedge, else posedge
    if (G.dly[j][kd][0] != NO) return sel_delay(G.dly[j][kd][0], G.dly[j][kd][1], s);
    return 1;
  }
  uint32_t m = 0xFFFFFFFFu;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (!((mask >> j) & 1)) continue;
    if (G.dly[j][0][0] != NO) m = min(m, sel_delay(G.dly[j][0][0], G.dly[j][0][1], s));
    uint32_t ie = edge_class(prev[j], cur[j]);
    int kd = (ie == 2) ? 2 : 1;
    if (G.dly[j][kd][0] != NO) m = min(m, sel_delay(G.dly[j][kd][0], G.dly[j][kd][1], s));
  }
  return m == 0xFFFFFFFFu ? 1 : m;
}

// ------------------------------------------------------------------ plan ----
__global__ void k_plan(const GateDev* __restrict__ gates, uint32_t ng,
                       const uint32_t* __restrict__ wlen, uint32_t seg_target,
                       uint64_t* __restrict__ nseg_out, uint64_t* __restrict__ scr_out,
                       uint32_t* __restrict__ gnseg) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const GateDev& G = gates[g];
  uint64_t N = 0;
  uint32_t nl = 0;
  for (int j = 0; j < G.k; ++j) {
    uint32_t n = wlen[G.in_net[j]];
    N += n;
    nl = max(nl, n);
  }
  uint64_t ns = N / seg_target + 1;
  if (ns > nl) ns = nl > 0 ? nl : 1;
  nseg_out[g] = ns;
  scr_out[g] = N + ns;
  gnseg[g] = (uint32_t)ns;
}

__global__ void k_segmap(uint32_t ng, const uint64_t* __restrict__ seg_base,
                         const uint32_t* __restrict__ gnseg, uint32_t* __restrict__ segmap) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  uint64_t b = seg_base[g];
  for (uint32_t c = 0; c < gnseg[g]; ++c) segmap[b + c] = g;
}

// ------------------------------------------------------------------ eval ----
__global__ __launch_bounds__(256) void k_eval(
    const GateDev* __restrict__ gates, const uint64_t* __restrict__ seg_base,
    const uint64_t* __restrict__ scr_base, const uint32_t* __restrict__ gnseg,
    const uint32_t* __restrict__ segmap, uint64_t nsegs, const uint64_t* __restrict__ woff,
    const uint32_t* __restrict__ wlen, const uint32_t* __restrict__ pool,
    const uint8_t* __restrict__ luts, uint32_t* __restrict__ scratch,
    SegSum* __restrict__ sums, unsigned long long* __restrict__ ev_count) {
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t own_events = 0;
  if (s < nsegs) {
    uint32_t g = segmap[s];
    const GateDev G = gates[g];
    const uint32_t c = (uint32_t)(s - seg_base[g]);
    const uint32_t nseg = gnseg[g];
    const int k = G.k;
    const uint8_t* lut = luts + (uint32_t)G.lut * 256u;

    const uint32_t* in[4];
    uint32_t n[4];
    int lead = 0;
    for (int j = 0; j < 4; ++j) {
      if (j < k) {
        in[j] = pool + woff[G.in_net[j]];
        n[j] = wlen[G.in_net[j]];
        if (n[j] > n[lead]) lead = j;
      } else {
        in[j] = pool;
        n[j] = 0;
      }
    }
    // Segment bounds from evenly spaced leader events.
    uint32_t tau0 = 0, tau1 = kTimeInf, halo_end = kTimeInf;
    if (c > 0) tau0 = in[lead][(uint64_t)c * n[lead] / nseg] >> 2;
    if (c + 1 < nseg) {
      tau1 = in[lead][(uint64_t)(c + 1) * n[lead] / nseg] >> 2;
      halo_end = tau1 + G.dmax;
    }
    uint32_t pos[4], head[4], cur[4];
    uint64_t before = 0;
    for (int j = 0; j < 4; ++j) {
      pos[j] = (j < k && c > 0) ? lower_bound_w(in[j], n[j], tau0 << 2) : 0;
      before += pos[j];
      cur[j] = (j < k && pos[j] > 0) ? (in[j][pos[j] - 1] & 3u) : 1u;  // X before the first entry
      head[j] = (j < k && pos[j] < n[j]) ? in[j][pos[j]] : kWordInf;
    }
    uint32_t li = cur[0] | (cur[1] << 2) | (cur[2] << 4) | (cur[3] << 6);
    uint32_t fprev = (c > 0) ? lut[li & ((1u << (2 * k)) - 1)] : 1u;  // f_{-1} = X

    const uint64_t off = scr_base[g] + before + c;
    uint32_t* st = scratch + off;
    uint32_t sp = 0;
    uint32_t top = 0;  // word of st[sp-1]
    if (c == 0) {      // every output starts with (0, X) (src/simulate.cpp:882-887)
      top = (0u << 2) | 1u;
      st[sp++] = top;
    }
    uint32_t hmin = 0xFFFFFFFFu;
    const uint32_t kmask = (1u << (2 * k)) - 1;
    for (;;) {
      uint32_t w = min(min(head[0], head[1]), min(head[2], head[3]));
      uint32_t t = w >> 2;
      if (w == kWordInf || t >= halo_end) break;
      uint32_t prev[4], mask = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        prev[j] = cur[j];
        if ((head[j] >> 2) == t && head[j] != kWordInf) {
          cur[j] = head[j] & 3u;
          mask |= 1u << j;
          ++pos[j];
          head[j] = pos[j] < n[j] ? in[j][pos[j]] : kWordInf;
        }
      }
      uint32_t idx = (cur[0] | (cur[1] << 2) | (cur[2] << 4) | (cur[3] << 6)) & kmask;
      uint32_t f = lut[idx];
      uint32_t sc = edge_class(fprev, f);
      fprev = f;
      bool own = t < tau1;
      own_events += own ? 1u : 0u;
      if (sc == 0) continue;
      uint32_t d = pick_delay(G, mask, prev, cur, sc);
      uint32_t o = t + d;
      if (own) {
        // pop every pending candidate at or after o, then push (o, f)
        while (sp > 0 && (top >> 2) >= o) {
          --sp;
          top = sp > 0 ? st[sp - 1] : 0u;
        }
        top = (o << 2) | f;
        st[sp++] = top;
      } else {
        hmin = min(hmin, o);
      }
    }
    while (sp > 0 && (top >> 2) >= hmin) {
      --sp;
      top = sp > 0 ? st[sp - 1] : 0u;
    }
    // in-segment dedup count
    uint32_t ndd = 0, firstv = 0, lastv = 0, lasto = 0, pv = 4;
    for (uint32_t i = 0; i < sp; ++i) {
      uint32_t e = st[i];
      if ((e & 3u) != pv) ++ndd;
      pv = e & 3u;
    }
    if (sp > 0) {
      firstv = st[0] & 3u;
      lastv = st[sp - 1] & 3u;
      lasto = st[sp - 1] >> 2;
    }
    SegSum S;
    S.scr_off = off;
    S.n_raw = sp;
    S.n_dd = ndd;
    S.last_o = lasto;
    S.first_v = (uint8_t)firstv;
    S.last_v = (uint8_t)lastv;
    S.pad[0] = S.pad[1] = 0;
    sums[s] = S;
  }
  // one atomic per wavefront for the event count
  unsigned long long v = own_events;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(ev_count, v);
}

// ------------------------------------------------------------- gate scan ----
GLS_DEV void saif_add(unsigned long long* acc, uint32_t v, uint64_t a, uint64_t b,
                      uint64_t start, uint64_t end) {
  // duration of [a, b) inside [start, end) in state v (src/simulate.cpp:102-163)
  uint64_t lo = a > start ? a : start;
  uint64_t hi = b < end ? b : end;
  if (hi <= lo) return;
  uint64_t len = hi - lo;
  if (v == 0) acc[0] += len;
  else if (v == 3) acc[1] += len;
  else acc[2] += len;
}

__global__ void k_gate_scan(const GateDev* __restrict__ gates, uint32_t ng,
                            const uint64_t* __restrict__ seg_base,
                            const uint32_t* __restrict__ gnseg, const SegSum* __restrict__ sums,
                            SegCarry* __restrict__ carry, uint64_t* __restrict__ out_len,
                            unsigned long long* __restrict__ saif, uint64_t start, uint64_t end) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  uint64_t b = seg_base[g];
  uint32_t total = 0, cv = 4, ct = 0;
  for (uint32_t c = 0; c < gnseg[g]; ++c) {
    SegSum S = sums[b + c];
    SegCarry C;
    C.pos = total;
    C.carry_t = ct;
    C.carry_v = (uint8_t)cv;
    C.pad[0] = C.pad[1] = C.pad[2] = 0;
    carry[b + c] = C;
    if (S.n_raw) {
      total += S.n_dd - (S.first_v == cv ? 1u : 0u);
      cv = S.last_v;
      ct = S.last_o;
    }
  }
  out_len[g] = total;
  // tail of the waveform: last value until `end`; written before k_emit's
  // atomics, so plain stores initialise the counters
  unsigned long long acc[3] = {0, 0, 0};
  if (cv != 4) saif_add(acc, cv, ct, end > ct ? end : ct, start, end);
  unsigned long long* dst = saif + (uint64_t)kSaifWords * gates[g].out_net;
  dst[0] = acc[0];
  dst[1] = acc[1];
  dst[2] = acc[2];
  dst[3] = 0;
}

// ------------------------------------------------------------------ emit ----
__global__ __launch_bounds__(256) void k_emit(
    const GateDev* __restrict__ gates, const uint64_t* __restrict__ seg_base,
    const uint32_t* __restrict__ segmap, uint64_t nsegs, const SegSum* __restrict__ sums,
    const SegCarry* __restrict__ carry, const uint64_t* __restrict__ out_off,
    const uint64_t* __restrict__ out_len, const uint32_t* __restrict__ scratch,
    uint32_t* __restrict__ pool, uint64_t* __restrict__ woff, uint32_t* __restrict__ wlen,
    unsigned long long* __restrict__ saif, uint64_t start, uint64_t end) {
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsegs) return;
  uint32_t g = segmap[s];
  uint32_t out_net = gates[g].out_net;
  uint64_t base = out_off[g];
  if (s == seg_base[g]) {
    woff[out_net] = base;
    wlen[out_net] = (uint32_t)out_len[g];
  }
  SegSum S = sums[s];
  if (S.n_raw == 0) return;
  SegCarry C = carry[s];
  const uint32_t* src = scratch + S.scr_off;
  uint32_t* dst = pool + base + C.pos;
  uint32_t pv = C.carry_v, pt = C.carry_t;
  unsigned long long acc[3] = {0, 0, 0}, tc = 0;
  uint32_t w = 0;
  for (uint32_t i = 0; i < S.n_raw; ++i) {
    uint32_t e = src[i];
    uint32_t v = e & 3u, t = e >> 2;
    if (pv != 4) saif_add(acc, pv, pt, t, start, end);
    else if (t > 0) saif_add(acc, 1u, 0, t, start, end);
    if (pv != 4 && is_toggle(pv, v) && t > start && t <= end) ++tc;
    if (v != pv) dst[w++] = e;
    pv = v;
    pt = t;
  }
  unsigned long long* a = saif + (uint64_t)kSaifWords * out_net;
  if (acc[0]) atomicAdd(a + 0, acc[0]);
  if (acc[1]) atomicAdd(a + 1, acc[1]);
  if (acc[2]) atomicAdd(a + 2, acc[2]);
  if (tc) atomicAdd(a + 3, tc);
}

// Stimulus / constant waveforms: SAIF durations, one thread per net.
// SAIF durations of nets that no gate drives (stimulus signals, constants):
// one wavefront per net, lane i takes entries i, i+64, ... (coalesced); entry
// q holds its value until entry q+1 (the last one until the window end), and
// the net is X before its first entry.  Window of net i: win[2*job[i]].
__global__ void k_saif_nets(const uint32_t* __restrict__ nets, const uint32_t* __restrict__ job,
                            uint32_t count, const uint64_t* __restrict__ woff,
                            const uint32_t* __restrict__ wlen, const uint32_t* __restrict__ pool,
                            unsigned long long* __restrict__ saif,
                            const unsigned long long* __restrict__ win) {
  const uint32_t i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (i >= count) return;
  const uint32_t net = nets[i];
  const uint64_t start = win[2 * job[i]], end = win[2 * job[i] + 1];
  const uint32_t* w = pool + woff[net];
  const uint32_t n = wlen[net];
  unsigned long long acc[4] = {0, 0, 0, 0};
  if (lane == 0) saif_add(acc, 1u, 0, n ? (w[0] >> 2) : (end > 0 ? end : 0), start, end);
  for (uint32_t q = lane; q < n; q += 64) {
    const uint32_t e = w[q];
    const uint64_t t = e >> 2;
    const uint64_t tn = q + 1 < n ? (uint64_t)(w[q + 1] >> 2) : (end > t ? end : t);
    saif_add(acc, e & 3u, t, tn, start, end);
    // stimulus waveforms may repeat a value (scalars are not de-duplicated)
    if (q > 0 && is_toggle(w[q - 1] & 3u, e & 3u) && t > start && t <= end) ++acc[3];
  }
  for (int k = 0; k < 4; ++k)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
  if (lane == 0) {
    unsigned long long* d = saif + (uint64_t)kSaifWords * net;
    d[0] = acc[0];
    d[1] = acc[1];
    d[2] = acc[2];
    d[3] = acc[3];
  }
}

// ------------------------------------------------------------------ scan ----
// Exclusive scan of u64 in three passes: block sums, scan of block sums,
// block-local scan plus block offset.  `base` (device scalar, may be null)
// is added to every output; `total` (device scalar) receives the sum.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

GLS_DEV uint64_t block_excl_scan(uint64_t v, uint64_t* sh, uint64_t& block_total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  uint64_t wpre = 0, tot = 0;
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if (i < wid) wpre += sh[i];
    tot += sh[i];
  }
  __syncthreads();
  block_total = tot;
  return wpre + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint64_t* __restrict__ in,
                                                               uint64_t n,
                                                               uint64_t* __restrict__ part) {
  __shared__ uint64_t sh[kScanThreads / 64];
  uint64_t b0 = (uint64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = b0 + (uint64_t)threadIdx.x * kScanItems + i;
    if (idx < n) s += in[idx];
  }
  uint64_t tot;
  block_excl_scan(s, sh, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_part(uint64_t* __restrict__ part,
                                                             uint64_t nb,
                                                             uint64_t* __restrict__ total) {
  __shared__ uint64_t sh[kScanThreads / 64];
  uint64_t carry = 0;
  for (uint64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    uint64_t idx = b0 + threadIdx.x;
    uint64_t v = idx < nb ? part[idx] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_scan(v, sh, tot);
    if (idx < nb) part[idx] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(
    const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n,
    const uint64_t* __restrict__ part, const uint64_t* __restrict__ base) {
  __shared__ uint64_t sh[kScanThreads / 64];
  uint64_t b0 = (uint64_t)blockIdx.x * kScanTile;
  uint64_t vals[kScanItems];
  uint64_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = b0 + (uint64_t)threadIdx.x * kScanItems + i;
    vals[i] = idx < n ? in[idx] : 0;
    s += vals[i];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan(s, sh, tot);
  uint64_t run = part[blockIdx.x] + ex + (base ? *base : 0);
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t idx = b0 + (uint64_t)threadIdx.x * kScanItems + i;
    if (idx < n) out[idx] = run;
    run += vals[i];
  }
}

__global__ void k_bump(uint64_t* __restrict__ bump, const uint64_t* __restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *bump += *total;
}

// ------------------------------------------------------- launch helpers ----
#define GLS_CHECK_LAUNCH() \
  do {                      \
    hipError_t e_ = hipGetLastError(); \
    if (e_ != hipSuccess) return e_;   \
  } while (0)

static inline unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Exclusive scan of n u64 values; part must hold ceil(n / kScanTile) + 1.
hipError_t launch_scan(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* part,
                       uint64_t* total, const uint64_t* base, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t nb = (n + kScanTile - 1) / kScanTile;
  k_scan_reduce<<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, part);
  GLS_CHECK_LAUNCH();
  k_scan_part<<<1, kScanThreads, 0, st>>>(part, nb, total);
  GLS_CHECK_LAUNCH();
  k_scan_apply<<<(unsigned)nb, kScanThreads, 0, st>>>(in, out, n, part, base);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

uint64_t scan_part_size(uint64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t launch_plan(const GateDev* gates, uint32_t ng, const uint32_t* wlen, uint32_t seg_target,
                       uint64_t* nseg, uint64_t* scr, uint32_t* gnseg, hipStream_t st) {
  k_plan<<<blocks_for(ng, 256), 256, 0, st>>>(gates, ng, wlen, seg_target, nseg, scr, gnseg);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_segmap(uint32_t ng, const uint64_t* seg_base, const uint32_t* gnseg,
                         uint32_t* segmap, hipStream_t st) {
  k_segmap<<<blocks_for(ng, 256), 256, 0, st>>>(ng, seg_base, gnseg, segmap);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_eval(const GateDev* gates, const uint64_t* seg_base, const uint64_t* scr_base,
                       const uint32_t* gnseg, const uint32_t* segmap, uint64_t nsegs,
                       const uint64_t* woff, const uint32_t* wlen, const uint32_t* pool,
                       const uint8_t* luts, uint32_t* scratch, SegSum* sums,
                       unsigned long long* ev_count, hipStream_t st) {
  if (nsegs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_eval, dim3(blocks_for(nsegs, 256)), dim3(256), 0, st, gates, seg_base,
                     scr_base, gnseg, segmap, nsegs, woff, wlen, pool, luts, scratch, sums,
                     ev_count);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_gate_scan(const GateDev* gates, uint32_t ng, const uint64_t* seg_base,
                            const uint32_t* gnseg, const SegSum* sums, SegCarry* carry,
                            uint64_t* out_len, unsigned long long* saif, uint64_t start,
                            uint64_t end, hipStream_t st) {
  k_gate_scan<<<blocks_for(ng, 256), 256, 0, st>>>(gates, ng, seg_base, gnseg, sums, carry,
                                                   out_len, saif, start, end);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_emit(const GateDev* gates, const uint64_t* seg_base, const uint32_t* segmap,
                       uint64_t nsegs, const SegSum* sums, const SegCarry* carry,
                       const uint64_t* out_off, const uint64_t* out_len, const uint32_t* scratch,
                       uint32_t* pool, uint64_t* woff, uint32_t* wlen, unsigned long long* saif,
                       uint64_t start, uint64_t end, hipStream_t st) {
  if (nsegs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit, dim3(blocks_for(nsegs, 256)), dim3(256), 0, st, gates, seg_base,
                     segmap, nsegs, sums, carry, out_off, out_len, scratch, pool, woff, wlen,
                     saif, start, end);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_bump(uint64_t* bump, const uint64_t* total, hipStream_t st) {
  k_bump<<<1, 64, 0, st>>>(bump, total);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

hipError_t launch_saif_nets(const uint32_t* nets, const uint32_t* job, uint32_t count,
                            const uint64_t* woff, const uint32_t* wlen, const uint32_t* pool,
                            unsigned long long* saif, const unsigned long long* win,
                            hipStream_t st) {
  if (count == 0) return hipSuccess;
  k_saif_nets<<<blocks_for(count, 4), 256, 0, st>>>(nets, job, count, woff, wlen, pool, saif, win);
  GLS_CHECK_LAUNCH();
  return hipSuccess;
}

}  // namespace gls
