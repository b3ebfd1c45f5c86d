// Time-chunk kernels of the fused path (gfx950): the per-chunk stimulus
// waveforms, the gate-output tails carried from one chunk to the next, the
// per-chunk SAIF accumulation, and the SAIF of the resident stimulus.
//
// A simulation runs as time chunks [F_k, F_k+1) (device.h, ChunkWin).  The
// stimulus stays resident for the whole run as 64-bit words
// (absolute time << 2 | value); each chunk cuts every signal's range
// [lb_k, F_k+1) out of it, in the chunk's relative time, behind a state entry
// (the value before lb_k, X when there is none: src/simulate.cpp:1459-1506
// reads X before a net's first entry).  Gate outputs of chunk k are final
// below F_k+1 (a candidate made at time t has output time >= t, so events at
// or after F_k+1 cannot pre-empt them); their tails (the state entry before
// lb_k+1 and the entries at or after it) seed chunk k+1.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace gls {

// first index of a sorted 64-bit word range whose time (word >> 2) is at or
// after `t`
__device__ __forceinline__ uint32_t lb_time64(const uint64_t* a, uint32_t n, uint64_t t) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((a[mid] >> 2) < t) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Per stimulus signal: the range [lb, ub) of its resident waveform and the
// chunk words it needs (one state entry + the range).
__global__ void k_stim_count(const uint64_t* __restrict__ store, const uint64_t* __restrict__ soff,
                             const uint32_t* __restrict__ slen, uint32_t ns, uint64_t lb, uint64_t ub,
                             uint64_t* __restrict__ cnt, uint32_t* __restrict__ first) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const uint64_t* w = store + soff[s];
  const uint32_t n = slen[s];
  const uint32_t i = lb ? lb_time64(w, n, lb) : 0u;
  const uint32_t j = lb_time64(w, n, ub);
  first[s] = i;
  cnt[s] = 1ull + (j - i);
}

// One wavefront per stimulus signal: state entry at relative time 0, then
// the range in relative time (abs - lb + 1 >= 1); sets the net's waveform.
// A chunk whose stimulus does not fit the pool (a replayed sequence that
// outgrew its first run) flags ERR_CAPACITY | ERR_POOL and copies nothing.
__global__ void k_stim_copy(const uint64_t* __restrict__ store, const uint64_t* __restrict__ soff,
                            const uint32_t* __restrict__ snet, uint32_t ns, uint64_t lb,
                            const uint64_t* __restrict__ cnt, const uint32_t* __restrict__ first,
                            const uint64_t* __restrict__ base, uint32_t* __restrict__ pool,
                            uint64_t pool_words, uint64_t* __restrict__ woff,
                            uint32_t* __restrict__ wlen, unsigned int* __restrict__ err) {
  const uint32_t s = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (s >= ns) return;
  if (base[s] + cnt[s] > pool_words) {
    if (lane == 0) {
      atomicOr(err, 8u | 16u);  // ERR_CAPACITY | ERR_POOL
      woff[snet[s]] = 0;
      wlen[snet[s]] = 0;
    }
    return;
  }
  const uint64_t* w = store + soff[s] + first[s];
  const uint32_t n = (uint32_t)(cnt[s] - 1);
  uint32_t* d = pool + base[s];
  if (lane == 0) {
    d[0] = first[s] > 0 ? (uint32_t)(w[-1] & 3u) : 1u;  // X before the first entry
    woff[snet[s]] = base[s];
    wlen[snet[s]] = n + 1;
  }
  for (uint32_t q = lane; q < n; q += 64) {
    const uint64_t e = w[q];
    d[1 + q] = (uint32_t)((((e >> 2) - lb + 1) << 2) | (e & 3u));
  }
}

// Chunk 0: every gate output's tail is one X state entry (word 1 at
// relative time 0), tail of gate g at tails[g].
__global__ void k_tails_init(const GateRec* __restrict__ gates, uint32_t ng,
                             uint32_t* __restrict__ tail_len, uint64_t* __restrict__ tail_off,
                             uint32_t* __restrict__ tails) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const uint32_t net = gates[g].out_net;
  tail_len[net] = 1;
  tail_off[net] = g;
  tails[g] = 1u;
}

// End of chunk k, per gate output: SAIF durations and toggles of the chunk
// join the run's totals, and the net's new entries (after its tail) are
// added to its waveform length.
__global__ void k_saif_accum(const GateRec* __restrict__ gates, uint32_t ng,
                             const unsigned long long* __restrict__ chunk,
                             unsigned long long* __restrict__ saif, const uint32_t* __restrict__ wlen,
                             const uint32_t* __restrict__ tail_len, unsigned long long* __restrict__ wtot) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const uint32_t net = gates[g].out_net;
  const unsigned long long* c = chunk + (uint64_t)kSaifWords * net;
  unsigned long long* d = saif + (uint64_t)kSaifWords * net;
  d[0] += c[0];
  d[1] += c[1];
  d[2] += c[2];
  d[3] += c[3];
  wtot[net] += wlen[net] - tail_len[net];
}

// End of chunk k, per gate output: where its tail for chunk k+1 starts in
// its chunk waveform (relative time of chunk k): the last entry before
// `key` (lb_k+1 in chunk k's time; index 0, the state entry, is below it)
// and everything after it.  Tails are short, so the search gallops back
// from the end.
__global__ void k_tail_len(const GateRec* __restrict__ gates, uint32_t ng,
                           const uint64_t* __restrict__ woff, const uint32_t* __restrict__ wlen,
                           const uint32_t* __restrict__ pool, uint32_t key,
                           uint64_t* __restrict__ len, uint64_t* __restrict__ src) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const uint32_t net = gates[g].out_net;
  const uint32_t* w = pool + woff[net];
  const uint32_t n = wlen[net];
  const uint32_t kw = key << 2;
  // w[lo] < kw; every index at or after hi is >= kw (hi = n: none checked)
  uint32_t lo = 0, hi = n, d = 1;
  while (hi > lo + 1) {
    const uint32_t p = hi - min(d, hi - lo - 1);
    if (w[p] < kw) {
      lo = p;
      break;
    }
    hi = p;
    d <<= 1;
  }
  while (hi > lo + 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (w[mid] < kw) lo = mid;
    else hi = mid;
  }
  len[g] = n - lo;
  src[g] = woff[net] + lo;
}

// Copies each gate output's tail into the tail store (kept in chunk k's
// relative time; k_compact moves it to chunk k+1's).  A store too small for
// the chunk (a replayed sequence that outgrew its first run) flags
// ERR_CAPACITY and copies nothing past its end.
__global__ void k_tail_copy(const GateRec* __restrict__ gates, uint32_t ng,
                            const uint64_t* __restrict__ len, const uint64_t* __restrict__ src,
                            const uint64_t* __restrict__ off, const uint32_t* __restrict__ pool,
                            uint32_t* __restrict__ tails, uint64_t cap,
                            uint32_t* __restrict__ tail_len, uint64_t* __restrict__ tail_off,
                            unsigned int* __restrict__ err) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const uint32_t net = gates[g].out_net;
  const uint64_t n = len[g], o = off[g];
  if (o + n > cap) {
    atomicOr(err, 8u);  // ERR_CAPACITY
    tail_len[net] = 0;
    tail_off[net] = 0;
    return;
  }
  const uint32_t* s = pool + src[g];
  for (uint64_t q = 0; q < n; ++q) tails[o + q] = s[q];
  tail_len[net] = (uint32_t)n;
  tail_off[net] = o;
}

// SAIF of the resident stimulus signals over the whole window (once per
// run; they are never re-computed per chunk): one wavefront per signal, the
// net is X before its first entry, an entry holds until the next one (the
// last until the window end), src/simulate.cpp:102-163.
__device__ __forceinline__ void saif_add64(unsigned long long* acc, uint32_t v, uint64_t a,
                                           uint64_t b, uint64_t start, uint64_t end) {
  const uint64_t lo = a > start ? a : start, hi = b < end ? b : end;
  if (hi <= lo) return;
  const uint64_t len = hi - lo;
  if (v == 0) acc[0] += len;
  else if (v == 3) acc[1] += len;
  else acc[2] += len;
}

__global__ void k_saif_stim(const uint64_t* __restrict__ store, const uint64_t* __restrict__ soff,
                            const uint32_t* __restrict__ slen, const uint32_t* __restrict__ snet,
                            const uint32_t* __restrict__ sjob, uint32_t ns,
                            unsigned long long* __restrict__ saif,
                            const unsigned long long* __restrict__ win) {
  const uint32_t s = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (s >= ns) return;
  const uint64_t start = win[2 * sjob[s]], end = win[2 * sjob[s] + 1];
  const uint64_t* w = store + soff[s];
  const uint32_t n = slen[s];
  unsigned long long acc[4] = {0, 0, 0, 0};
  if (lane == 0) saif_add64(acc, 1u, 0, n ? (w[0] >> 2) : end, start, end);
  for (uint32_t q = lane; q < n; q += 64) {
    const uint64_t e = w[q], t = e >> 2;
    const uint64_t tn = q + 1 < n ? (w[q + 1] >> 2) : (end > t ? end : t);
    saif_add64(acc, (uint32_t)(e & 3u), t, tn, start, end);
    // stimulus waveforms may repeat a value (scalars are not de-duplicated)
    if (q > 0 && is_toggle((uint32_t)(w[q - 1] & 3u), (uint32_t)(e & 3u)) && t > start && t <= end)
      ++acc[3];
  }
  for (int k = 0; k < 4; ++k)
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_down(acc[k], o, 64);
  if (lane == 0) {
    unsigned long long* d = saif + (uint64_t)kSaifWords * snet[s];
    d[0] = acc[0];
    d[1] = acc[1];
    d[2] = acc[2];
    d[3] = acc[3];
  }
}

static inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_stim_count(const uint64_t* store, const uint64_t* soff, const uint32_t* slen,
                             uint32_t ns, uint64_t lb, uint64_t ub, uint64_t* cnt, uint32_t* first,
                             hipStream_t st) {
  if (ns == 0) return hipSuccess;
  k_stim_count<<<nblk(ns, 256), 256, 0, st>>>(store, soff, slen, ns, lb, ub, cnt, first);
  return hipGetLastError();
}

hipError_t launch_stim_copy(const uint64_t* store, const uint64_t* soff, const uint32_t* snet,
                            uint32_t ns, uint64_t lb, const uint64_t* cnt, const uint32_t* first,
                            const uint64_t* base, uint32_t* pool, uint64_t pool_words,
                            uint64_t* woff, uint32_t* wlen, unsigned int* err, hipStream_t st) {
  if (ns == 0) return hipSuccess;
  k_stim_copy<<<nblk(ns, 4), 256, 0, st>>>(store, soff, snet, ns, lb, cnt, first, base, pool,
                                           pool_words, woff, wlen, err);
  return hipGetLastError();
}

hipError_t launch_tails_init(const GateRec* gates, uint32_t ng, uint32_t* tail_len,
                             uint64_t* tail_off, uint32_t* tails, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  k_tails_init<<<nblk(ng, 256), 256, 0, st>>>(gates, ng, tail_len, tail_off, tails);
  return hipGetLastError();
}

hipError_t launch_saif_accum(const GateRec* gates, uint32_t ng, const unsigned long long* chunk,
                             unsigned long long* saif, const uint32_t* wlen,
                             const uint32_t* tail_len, unsigned long long* wtot, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  k_saif_accum<<<nblk(ng, 256), 256, 0, st>>>(gates, ng, chunk, saif, wlen, tail_len, wtot);
  return hipGetLastError();
}

hipError_t launch_tail_len(const GateRec* gates, uint32_t ng, const uint64_t* woff,
                           const uint32_t* wlen, const uint32_t* pool, uint32_t key, uint64_t* len,
                           uint64_t* src, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  k_tail_len<<<nblk(ng, 256), 256, 0, st>>>(gates, ng, woff, wlen, pool, key, len, src);
  return hipGetLastError();
}

hipError_t launch_tail_copy(const GateRec* gates, uint32_t ng, const uint64_t* len,
                            const uint64_t* src, const uint64_t* off, const uint32_t* pool,
                            uint32_t* tails, uint64_t cap, uint32_t* tail_len, uint64_t* tail_off,
                            unsigned int* err, hipStream_t st) {
  if (ng == 0) return hipSuccess;
  k_tail_copy<<<nblk(ng, 256), 256, 0, st>>>(gates, ng, len, src, off, pool, tails, cap, tail_len,
                                             tail_off, err);
  return hipGetLastError();
}

// dst[i] += src[i]: a chunk's per-instance gate events join the run's.
__global__ void k_add_u64(unsigned long long* dst, const unsigned long long* src, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

hipError_t launch_add_u64(unsigned long long* dst, const unsigned long long* src, uint64_t n,
                          hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_add_u64<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_saif_stim(const uint64_t* store, const uint64_t* soff, const uint32_t* slen,
                            const uint32_t* snet, const uint32_t* sjob, uint32_t ns,
                            unsigned long long* saif, const unsigned long long* win,
                            hipStream_t st) {
  if (ns == 0) return hipSuccess;
  k_saif_stim<<<nblk(ns, 4), 256, 0, st>>>(store, soff, slen, snet, sjob, ns, saif, win);
  return hipGetLastError();
}

}  // namespace gls
