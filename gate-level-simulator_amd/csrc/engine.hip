#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>

#include "../../include/gls_status.h"

namespace gls {

// launch helpers (kernels.hip: two-pass path and scans)
hipError_t launch_scan(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* part,
                       uint64_t* total, const uint64_t* base, hipStream_t st);
uint64_t scan_part_size(uint64_t n);
hipError_t launch_plan(const GateDev* gates, uint32_t ng, const uint32_t* wlen, uint32_t seg_target,
                       uint64_t* nseg, uint64_t* scr, uint32_t* gnseg, hipStream_t st);
hipError_t launch_segmap(uint32_t ng, const uint64_t* seg_base, const uint32_t* gnseg,
                         uint32_t* segmap, hipStream_t st);
hipError_t launch_eval(const GateDev* gates, const uint64_t* seg_base, const uint64_t* scr_base,
                       const uint32_t* gnseg, const uint32_t* segmap, uint64_t nsegs,
                       const uint64_t* woff, const uint32_t* wlen, const uint32_t* pool,
                       const uint8_t* luts, uint32_t* scratch, SegSum* sums,
                       unsigned long long* ev_count, hipStream_t st);
hipError_t launch_gate_scan(const GateDev* gates, uint32_t ng, const uint64_t* seg_base,
                            const uint32_t* gnseg, const SegSum* sums, SegCarry* carry,
                            uint64_t* out_len, unsigned long long* saif, uint64_t start,
                            uint64_t end, hipStream_t st);
hipError_t launch_emit(const GateDev* gates, const uint64_t* seg_base, const uint32_t* segmap,
                       uint64_t nsegs, const SegSum* sums, const SegCarry* carry,
                       const uint64_t* out_off, const uint64_t* out_len, const uint32_t* scratch,
                       uint32_t* pool, uint64_t* woff, uint32_t* wlen, unsigned long long* saif,
                       uint64_t start, uint64_t end, hipStream_t st);
hipError_t launch_bump(uint64_t* bump, const uint64_t* total, hipStream_t st);
hipError_t launch_saif_nets(const uint32_t* nets, const uint32_t* job, uint32_t count,
                            const uint64_t* woff, const uint32_t* wlen, const uint32_t* pool,
                            unsigned long long* saif, const unsigned long long* win,
                            hipStream_t st);

// level.hip: fused path
hipError_t launch_plan2(const GateRec* gates, uint32_t ng, const uint32_t* wlen,
                        uint32_t seg_target, GatePlan* plan, uint64_t* ntile, uint64_t* bound,
                        unsigned int* err, const uint32_t* tail_len, hipStream_t st);
hipError_t launch_tile_desc(const GateRec* gates, uint32_t ng, const GatePlan* plan,
                            const uint64_t* tile_base, const uint64_t* woff, const uint32_t* wlen,
                            const uint32_t* pool, const uint32_t* lutp, TileDesc* tiles,
                            const uint64_t* ntiles_dev, uint64_t ntiles, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t st);
hipError_t launch_guard(uint64_t* scal, uint64_t tiles_cap, uint64_t scratch_cap, hipStream_t st);
hipError_t launch_pool_guard(uint64_t* scal, uint64_t pool_words, hipStream_t st);
hipError_t launch_gate_carry(const GateRec* gates, uint32_t ng, const uint64_t* tile_base,
                             const uint64_t* ntile, const TileOut* tout, uint32_t* tile_off,
                             uint64_t* gate_cnt, uint32_t* wlen, unsigned long long* saif,
                             const ChunkWin* cwin, const unsigned int* err, const TailArgs& T,
                             unsigned long long* ev_job, hipStream_t st);
hipError_t launch_add_u64(unsigned long long* dst, const unsigned long long* src, uint64_t n,
                          hipStream_t st);
hipError_t launch_compact(const TileDesc* tiles, const uint64_t* ntiles_dev, uint64_t ntiles_cap,
                          const TileOut* tout, const uint32_t* tile_off, const uint64_t* scr_base,
                          const uint32_t* scratch, const uint64_t* gate_base, uint32_t* pool,
                          uint64_t* woff, uint32_t* wlen, const unsigned int* err,
                          unsigned long long* saif, const ChunkWin* cwin, uint32_t count_tc,
                          const TailArgs& T, hipStream_t st);
hipError_t launch_set_scal(uint64_t* scal, uint64_t bump, hipStream_t st);
hipError_t launch_level(const LevelArgs& a, unsigned nblocks, hipStream_t st);
hipError_t launch_wide(const WideArgs& a, hipStream_t st);
// chunk.hip: time chunks
hipError_t launch_stim_count(const uint64_t* store, const uint64_t* soff, const uint32_t* slen,
                             uint32_t ns, uint64_t lb, uint64_t ub, uint64_t* cnt, uint32_t* first,
                             hipStream_t st);
hipError_t launch_stim_copy(const uint64_t* store, const uint64_t* soff, const uint32_t* snet,
                            uint32_t ns, uint64_t lb, const uint64_t* cnt, const uint32_t* first,
                            const uint64_t* base, uint32_t* pool, uint64_t pool_words,
                            uint64_t* woff, uint32_t* wlen, unsigned int* err, hipStream_t st);
hipError_t launch_tails_init(const GateRec* gates, uint32_t ng, uint32_t* tail_len,
                             uint64_t* tail_off, uint32_t* tails, hipStream_t st);
hipError_t launch_saif_accum(const GateRec* gates, uint32_t ng, const unsigned long long* chunk,
                             unsigned long long* saif, const uint32_t* wlen,
                             const uint32_t* tail_len, unsigned long long* wtot, hipStream_t st);
hipError_t launch_tail_len(const GateRec* gates, uint32_t ng, const uint64_t* woff,
                           const uint32_t* wlen, const uint32_t* pool, uint32_t key, uint64_t* len,
                           uint64_t* src, hipStream_t st);
hipError_t launch_tail_copy(const GateRec* gates, uint32_t ng, const uint64_t* len,
                            const uint64_t* src, const uint64_t* off, const uint32_t* pool,
                            uint32_t* tails, uint64_t cap, uint32_t* tail_len, uint64_t* tail_off,
                            unsigned int* err, hipStream_t st);
hipError_t launch_saif_stim(const uint64_t* store, const uint64_t* soff, const uint32_t* slen,
                            const uint32_t* snet, const uint32_t* sjob, uint32_t ns,
                            unsigned long long* saif, const unsigned long long* win,
                            hipStream_t st);
int level_max_blocks(uint32_t cap, bool stacks_in_lds);

// LDS input capacity of a fused tile: 64 segments of `seg` input events
// plus a little halo headroom.  Tiles above it (their density varies) take
// the HBM path; a tight capacity keeps more workgroups resident per CU,
// which pays more than the HBM path costs (200k-cell workload: 8 instead of
// 6 workgroups per CU, 9% faster than a 1.3x headroom).
static uint32_t tile_cap(uint32_t seg) {
  uint64_t c = (uint64_t)seg * 64 + 64;
  c = (c + 63) / 64 * 64;
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(c, 1024), 16384);
}

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  // the error travels in the exception: clear HIP's last-error state, which
  // the launch wrappers read (hipGetLastError) and would otherwise report
  // again for a later, good launch once the caller has handled this one
  (void)hipGetLastError();
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GLS_E_OOM : GLS_E_DEVICE;
  throw GlsError(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HC(x) hip_check((x), #x)

// Words allocated past the end of the event pool: the fused kernel stages
// input ranges with 16-byte copies that may read up to 3 words beyond a
// range's last word.
static constexpr uint64_t kPoolPad = 16;

// scalar slots in d_scal_
enum { S_BUMP = 0, S_TOT_A = 1, S_TOT_B = 2, S_TOT_OUT = 3, S_EVC = 4, S_TICKET = 5,
       S_OVF_USED = 6, S_ERR = 7, S_TILES = 8, S_TILES_HBM = 9, S_TAILS = 10 };

static uint32_t sel_delay(uint32_t r, uint32_t f, int s) {
  return s == 1 ? r : (s == 2 ? f : (s == 3 ? std::max(r, f) : std::min(r, f)));
}

// AddDelay (src/simulate.cpp:1012-1176, asserts compiled out) folded into
// one table per gate, indexed by dt_index(k, s, mask, fell): s = output edge
// class (1..4), mask = inputs with an event at this time, fell = the subset
// of those whose own edge class is "fall" (CheckDelay == 2).
//   * one input with an event: with a single IOPATH record for that input the
//     record is used whatever its edge; with several, the first record whose
//     edge matches the input's edge (fall -> negedge, else posedge), else 1;
//   * several inputs: the minimum over every edge-independent record and
//     every record whose edge matches its input's edge, else 1.
static void fill_delay_table(const GateDelays& D, int k, uint16_t* tab) {
  const uint16_t NO = kNoDelay;
  uint32_t tmin[4][4][2], tone[4][4][2];
  for (int j = 0; j < 4; ++j)
    for (int s = 1; s <= 4; ++s)
      for (int fall = 0; fall < 2; ++fall) {
        uint32_t m = 0xFFFF, one = 1;
        if (j < k) {
          const uint16_t* both = D.v[j][DK_BOTH];
          const uint16_t* kind = D.v[j][fall ? DK_NEG : DK_POS];
          if (both[0] != NO) m = std::min(m, sel_delay(both[0], both[1], s));
          if (kind[0] != NO) m = std::min(m, sel_delay(kind[0], kind[1], s));
          if (D.nrec[j] == 1) {
            for (int kd = 0; kd < 3; ++kd)
              if (D.v[j][kd][0] != NO) one = sel_delay(D.v[j][kd][0], D.v[j][kd][1], s);
          } else if (D.nrec[j] >= 2) {
            one = kind[0] != NO ? sel_delay(kind[0], kind[1], s) : 1;
          }
        }
        tmin[j][s - 1][fall] = m;
        tone[j][s - 1][fall] = one;
      }
  const uint32_t full = (1u << k) - 1;
  for (uint32_t s = 1; s <= 4; ++s)
    for (uint32_t mask = 0; mask <= full; ++mask)
      for (uint32_t fell = 0; fell <= full; ++fell) {
        uint32_t d = 1;
        if (mask != 0 && (mask & (mask - 1)) == 0) {
          const int j = __builtin_ctz(mask);
          d = tone[j][s - 1][(fell >> j) & 1u];
        } else if (mask != 0) {
          uint32_t m = 0xFFFF;
          for (int j = 0; j < k; ++j)
            if ((mask >> j) & 1u) m = std::min(m, tmin[j][s - 1][(fell >> j) & 1u]);
          d = m == 0xFFFF ? 1u : m;
        }
        tab[dt_index((uint32_t)k, s, mask, fell)] = (uint16_t)d;
      }
}

// AddDelay for gates of fan-in 5..6, the same rules as fill_delay_table
// kept per input (a full table would hold 4 << 12 entries): tmin and tone
// of every input, output edge class and input edge, at wide_dt_index().
static void fill_wide_delays(const GateDelays& D, int k, uint16_t* tab) {
  const uint16_t NO = kNoDelay;
  for (int j = 0; j < (int)kWideK; ++j)
    for (uint32_t s = 1; s <= 4; ++s)
      for (uint32_t fall = 0; fall < 2; ++fall) {
        uint32_t m = 0xFFFF, one = 1;
        if (j < k) {
          const uint16_t* both = D.v[j][DK_BOTH];
          const uint16_t* kind = D.v[j][fall ? DK_NEG : DK_POS];
          if (both[0] != NO) m = std::min(m, sel_delay(both[0], both[1], (int)s));
          if (kind[0] != NO) m = std::min(m, sel_delay(kind[0], kind[1], (int)s));
          if (D.nrec[j] == 1) {
            for (int kd = 0; kd < 3; ++kd)
              if (D.v[j][kd][0] != NO) one = sel_delay(D.v[j][kd][0], D.v[j][kd][1], (int)s);
          } else if (D.nrec[j] >= 2) {
            one = kind[0] != NO ? sel_delay(kind[0], kind[1], (int)s) : 1;
          }
        }
        tab[wide_dt_index((uint32_t)j, s, fall, 0)] = (uint16_t)m;
        tab[wide_dt_index((uint32_t)j, s, fall, 1)] = (uint16_t)one;
      }
}

// ------------------------------------------------------------- programs ---

DevProgram::~DevProgram() {
  if (d_block) (void)hipFree(d_block);
}

size_t DevProgram::device_bytes() const {
  return h_grec.size() * sizeof(GateRec) + h_gdev.size() * sizeof(GateDev) +
         h_dtab.size() * 2 + h_luts.size() + h_lutp.size() * 4 + h_wide.size() * 4;
}

bool DevProgram::is_resident() const { return resident.load(std::memory_order_acquire); }

// The six tables share one device allocation and are published only after
// it and every copy succeeded: a failed upload leaves the image
// non-resident (all pointers null) and the next upload starts over.
void DevProgram::upload(int dev, hipStream_t st) {
  std::lock_guard<std::mutex> lk(upload_mu);
  if (resident.load(std::memory_order_relaxed)) {
    if (dev != device) throw GlsError(GLS_E_ARG, "program image belongs to another device");
    return;
  }
  const size_t sizes[6] = {h_grec.size() * sizeof(GateRec), h_gdev.size() * sizeof(GateDev),
                           h_dtab.size() * 2, h_luts.size(), h_lutp.size() * 4, h_wide.size() * 4};
  const void* srcs[6] = {h_grec.data(), h_gdev.data(), h_dtab.data(), h_luts.data(), h_lutp.data(),
                         h_wide.data()};
  size_t off[6], total = 0;
  for (int i = 0; i < 6; ++i) {
    off[i] = total;
    total += (std::max<size_t>(sizes[i], 1) + 255) / 256 * 256;
  }
  char* blk = nullptr;
  HC(hipMalloc((void**)&blk, total));
  try {
    for (int i = 0; i < 6; ++i)
      if (sizes[i]) HC(hipMemcpyAsync(blk + off[i], srcs[i], sizes[i], hipMemcpyHostToDevice, st));
    HC(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(blk);
    throw;
  }
  device = dev;
  d_block = blk;
  d_grec = (GateRec*)(blk + off[0]);
  d_gdev = (GateDev*)(blk + off[1]);
  d_dtab = (uint16_t*)(blk + off[2]);
  d_luts = (uint8_t*)(blk + off[3]);
  d_lutp = (uint32_t*)(blk + off[4]);
  d_wide = (uint32_t*)(blk + off[5]);
  resident.store(true, std::memory_order_release);
}

// Wide gates of an image, level by level.
static void index_wide(DevProgram& P) {
  P.h_wide.clear();
  P.wide_start.assign(1, 0);
  P.max_k = 0;
  for (size_t l = 0; l + 1 < P.level_start.size(); ++l) {
    for (uint32_t g = P.level_start[l]; g < P.level_start[l + 1]; ++g) {
      P.max_k = std::max<uint32_t>(P.max_k, P.h_grec[g].k);
      if (P.h_grec[g].k > kTileK) P.h_wide.push_back(g);
    }
    P.wide_start.push_back((uint32_t)P.h_wide.size());
  }
}

std::shared_ptr<DevProgram> build_program_image(const Program& pg, uint32_t timescale_ps) {
  auto P = std::make_shared<DevProgram>();
  P->timescale_ps = timescale_ps;
  P->n_gates = (uint32_t)pg.gates.size();
  P->n_nets = (uint32_t)pg.net_names.size();
  P->n_pins = (uint32_t)pg.n_pins;
  P->level_start = pg.level_start;
  // truth tables: 256-byte blocks, a fan-in 5 / 6 table spans 4 / 16
  std::vector<uint32_t> blk_of(pg.luts.size());
  size_t nblk = 0;
  for (size_t i = 0; i < pg.luts.size(); ++i) {
    blk_of[i] = (uint32_t)nblk;
    nblk += std::max<size_t>(1, (pg.luts[i].size() + 255) / 256);
  }
  if (nblk > 65535) throw GlsError(GLS_E_RANGE, "too many truth tables");
  P->h_grec.resize(P->n_gates);
  P->h_gdev.resize(P->n_gates);
  for (uint32_t i = 0; i < P->n_gates; ++i) {
    const CompiledGate& g = pg.gates[i];
    if (g.k < 1) throw GlsError(GLS_E_UNSUPPORTED, "gate without inputs");
    if (g.k > (int)kWideK) throw GlsError(GLS_E_UNSUPPORTED, "gate with more than 6 distinct inputs");
    GateDelays D = gate_delays(g, timescale_ps);
    GateDev& r = P->h_gdev[i];
    std::memset(&r, 0, sizeof r);
    GateRec& q = P->h_grec[i];
    std::memset(&q, 0, sizeof q);
    for (int j = 0; j < (int)kWideK; ++j) q.in_net[j] = j < g.k ? (uint32_t)g.in_net[j] : 0u;
    for (int j = 0; j < 4; ++j) {
      r.in_net[j] = q.in_net[j];
      r.nrec[j] = D.nrec[j];
      for (int kd = 0; kd < 3; ++kd) {
        r.dly[j][kd][0] = D.v[j][kd][0];
        r.dly[j][kd][1] = D.v[j][kd][1];
      }
    }
    r.out_net = q.out_net = (uint32_t)g.out_net;
    r.lut = q.lut = (uint16_t)blk_of[g.lut];
    r.k = q.k = (uint8_t)g.k;
    r.dmax = q.dmax = D.dmax;
    const uint32_t dts = g.k > (int)kTileK ? kWideDtSize : dt_size(g.k);
    if (P->h_dtab.size() + dts >= (1ull << 32))
      throw GlsError(GLS_E_RANGE, "delay tables exceed 2^32 entries");
    q.dt_off = (uint32_t)P->h_dtab.size();
    P->h_dtab.resize(P->h_dtab.size() + dts);
    if (g.k > (int)kTileK) fill_wide_delays(D, g.k, P->h_dtab.data() + q.dt_off);
    else fill_delay_table(D, g.k, P->h_dtab.data() + q.dt_off);
  }
  nblk = std::max<size_t>(1, nblk);
  P->h_luts.assign(nblk * 256, 1);
  P->h_lutp.assign(nblk * 16, 0);
  for (size_t i = 0; i < pg.luts.size(); ++i) {
    std::memcpy(&P->h_luts[(size_t)blk_of[i] * 256], pg.luts[i].data(), pg.luts[i].size());
    // packed 2 bits per entry for register lookups (fan-in <= 3)
    for (size_t e = 0; e < pg.luts[i].size() && e < 256; ++e)
      P->h_lutp[(size_t)blk_of[i] * 16 + e / 16] |= (uint32_t)(pg.luts[i][e] & 3u) << (2 * (e % 16));
  }
  index_wide(*P);
  return P;
}

std::shared_ptr<DevProgram> ProgramCache::get(const std::shared_ptr<const Program>& pg,
                                             uint32_t timescale_ps) {
  std::lock_guard<std::mutex> lk(mu_);
  auto key = std::make_pair(pg.get(), timescale_ps);
  auto it = map_.find(key);
  if (it != map_.end()) return it->second.second;
  auto img = build_program_image(*pg, timescale_ps);
  map_[key] = {pg, img};
  return img;
}

bool ProgramCache::resident(const Program* pg, uint32_t timescale_ps) const {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = map_.find(std::make_pair(pg, timescale_ps));
  return it != map_.end() && it->second.second->is_resident();
}

size_t ProgramCache::device_bytes() const {
  std::lock_guard<std::mutex> lk(mu_);
  size_t b = 0;
  for (auto& kv : map_)
    if (kv.second.second->is_resident()) b += kv.second.second->device_bytes();
  return b;
}

size_t ProgramCache::size() const {
  std::lock_guard<std::mutex> lk(mu_);
  return map_.size();
}

void ProgramCache::clear() {
  std::lock_guard<std::mutex> lk(mu_);
  map_.clear();
}

// Union of several instances' images: nets of instance i are shifted by
// net_base[i], gates are merged level by level, each gate records its
// instance; truth and delay tables are shared per distinct image.
// Gates are ordered group by group (instance groups [groups[q], groups[q+1])),
// each group level by level and each level instance by instance; the
// level table is flattened: group q's level l is level q * nlev + l.
static std::shared_ptr<DevProgram> union_image(const std::vector<std::shared_ptr<DevProgram>>& parts,
                                               const std::vector<uint32_t>& net_base,
                                               const std::vector<uint32_t>& groups) {
  auto U = std::make_shared<DevProgram>();
  std::map<const DevProgram*, std::pair<uint32_t, uint32_t>> bases;  // dtab, lut
  for (auto& p : parts) {
    if (bases.count(p.get())) continue;
    bases[p.get()] = {(uint32_t)U->h_dtab.size(), (uint32_t)(U->h_luts.size() / 256)};
    if (U->h_dtab.size() + p->h_dtab.size() >= (1ull << 32))
      throw GlsError(GLS_E_RANGE, "delay tables exceed 2^32 entries");
    U->h_dtab.insert(U->h_dtab.end(), p->h_dtab.begin(), p->h_dtab.end());
    U->h_luts.insert(U->h_luts.end(), p->h_luts.begin(), p->h_luts.end());
    U->h_lutp.insert(U->h_lutp.end(), p->h_lutp.begin(), p->h_lutp.end());
  }
  if (U->h_luts.size() / 256 > 65535) throw GlsError(GLS_E_RANGE, "too many truth tables in one batch");
  size_t nlev = 0;
  for (auto& p : parts) nlev = std::max(nlev, p->level_start.size() - 1);
  U->level_start.assign(1, 0);
  for (size_t q = 0; q + 1 < groups.size(); ++q)
  for (size_t l = 0; l < nlev; ++l) {
    for (size_t i = groups[q]; i < groups[q + 1]; ++i) {
      const DevProgram& p = *parts[i];
      if (l + 1 >= p.level_start.size()) continue;
      const auto b = bases[&p];
      for (uint32_t g = p.level_start[l]; g < p.level_start[l + 1]; ++g) {
        GateRec q = p.h_grec[g];
        GateDev r = p.h_gdev[g];
        for (int j = 0; j < (int)kWideK; ++j)
          if (j < q.k) q.in_net[j] += net_base[i];
        for (int j = 0; j < 4; ++j) r.in_net[j] = q.in_net[j];
        q.out_net += net_base[i];
        r.out_net = q.out_net;
        q.dt_off += b.first;
        q.lut = r.lut = (uint16_t)(q.lut + b.second);
        q.job = (uint16_t)i;
        U->h_grec.push_back(q);
        U->h_gdev.push_back(r);
      }
    }
    U->level_start.push_back((uint32_t)U->h_grec.size());
  }
  U->n_gates = (uint32_t)U->h_grec.size();
  index_wide(*U);
  for (size_t i = 0; i < parts.size(); ++i) {
    U->n_nets = std::max(U->n_nets, net_base[i] + parts[i]->n_nets);
    U->n_pins += parts[i]->n_pins;
  }
  return U;
}

// GLS_TRACE=1: one line per time chunk on stderr (bounds, outcome, pool
// words used, host time, pool reallocations).
static bool trace_on() {
  static const bool on = [] {
    const char* e = std::getenv("GLS_TRACE");
    return e && *e && *e != '0';
  }();
  return on;
}

// ---------------------------------------------------------------- arena ---

DevArena::~DevArena() {
  std::lock_guard<std::mutex> lk(mu_);
  trim_locked();
}

size_t DevArena::size_class(size_t bytes) {
  // 4 KiB at least; up to 1 MiB powers of two, above in 1 MiB steps
  if (bytes <= (1u << 20)) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
  }
  return (bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
}

bool DevArena::get(size_t bytes, void** p) {
  const size_t c = size_class(bytes);
  std::lock_guard<std::mutex> lk(mu_);
  auto it = free_.lower_bound(c);
  if (it != free_.end() && it->first <= c + c / 2) {
    *p = it->second;
    live_[*p] = it->first;
    cached_ -= it->first;
    free_.erase(it);
    ++hits_;
    return true;
  }
  ++misses_;
  for (int attempt = 0; attempt < 2; ++attempt) {
    void* q = nullptr;
    if (hipMalloc(&q, c) == hipSuccess) {
      live_[q] = c;
      *p = q;
      keep_headroom_locked();
      return true;
    }
    (void)hipGetLastError();
    if (free_.empty()) break;
    trim_locked();  // the cached blocks make room
  }
  *p = nullptr;
  return false;
}

bool DevArena::get_large(size_t want, size_t min_bytes, void** p, size_t* got) {
  min_bytes = size_class(std::max<size_t>(min_bytes, 1));
  std::lock_guard<std::mutex> lk(mu_);
  if (!free_.empty()) {
    auto last = std::prev(free_.end());
    if (last->first >= min_bytes) {
      *p = last->second;
      *got = last->first;
      live_[*p] = last->first;
      cached_ -= last->first;
      free_.erase(last);
      ++hits_;
      return true;
    }
  }
  ++misses_;
  const size_t first = size_class(std::max(want, min_bytes));
  for (int attempt = 0; attempt < 2; ++attempt) {
    // 10% smaller after each refusal, down to min_bytes
    for (size_t m = first;;) {
      void* q = nullptr;
      if (hipMalloc(&q, m) == hipSuccess) {
        live_[q] = m;
        *p = q;
        *got = m;
        keep_headroom_locked();
        return true;
      }
      (void)hipGetLastError();
      const size_t next = size_class(m / 10 * 9);
      if (m <= min_bytes || next >= m) break;
      m = std::max(next, min_bytes);
    }
    if (free_.empty()) break;
    trim_locked();  // the cached blocks make room
  }
  *p = nullptr;
  *got = 0;
  return false;
}

// The HIP runtime needs free device memory of its own (code objects loaded
// at a kernel's first launch, launch arguments): when an allocation leaves
// less than 4 GiB (or 1.5% of the device) free, the cached blocks go.
void DevArena::keep_headroom_locked() {
  if (free_.empty()) return;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  if (f < std::max<size_t>(4ull << 30, t / 64)) trim_locked();
}

void DevArena::keep_headroom() {
  std::lock_guard<std::mutex> lk(mu_);
  keep_headroom_locked();
}

uint64_t DevArena::cached_bytes() {
  std::lock_guard<std::mutex> lk(mu_);
  return cached_;
}

void DevArena::put(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  const size_t c = it->second;
  live_.erase(it);
  free_.emplace(c, p);
  cached_ += c;
  // over the cap: the largest cached blocks go first
  while (cached_ > max_cached_ && !free_.empty()) {
    auto last = std::prev(free_.end());
    (void)hipFree(last->second);
    cached_ -= last->first;
    free_.erase(last);
  }
}

void DevArena::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  trim_locked();
}

void DevArena::trim_locked() {
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_ = 0;
}

// ------------------------------------------------------------------ job ---

Job::Job(int device, hipStream_t stream, DevArena* arena)
    : device_(device), stream_(stream), arena_(arena) {}

Job::~Job() {
  const auto t0 = std::chrono::steady_clock::now();
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  // blocks go back to the arena only once this job's work has drained
  if (arena_) (void)hipStreamSynchronize(stream_);
  for (void* p : owned_)
    if (p) arena_ ? arena_->put(p) : (void)hipFree(p);
  void* grown[] = {d_segmap_, d_sums_, d_carry_, d_scratch_, d_tiles_, d_tout_, d_toff_, d_stk_,
                   d_pool_,   d_ovf_,  d_tails_, d_cwin_};
  for (void* p : grown)
    if (p) arena_ ? arena_->put(p) : (void)hipFree(p);
  if (trace_on())
    std::fprintf(stderr, "gls: job buffers released in %.1f ms\n",
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}

bool Job::dev_alloc(void** p, size_t bytes) {
  const auto t0 = std::chrono::steady_clock::now();
  bool ok;
  if (arena_) {
    ok = arena_->get(bytes, p);
  } else {
    ok = hipMalloc(p, bytes) == hipSuccess;
    if (!ok) {
      (void)hipGetLastError();
      *p = nullptr;
    }
  }
  if (trace_on() && bytes >= (1ull << 30))
    std::fprintf(stderr, "gls: device allocation of %.1f GiB %s in %.1f ms\n", bytes / 1073741824.0,
                 ok ? "done" : "failed",
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  return ok;
}

void Job::dev_free(void* p) {
  if (!p) return;
  const auto t0 = std::chrono::steady_clock::now();
  if (arena_) {
    HC(hipStreamSynchronize(stream_));
    arena_->put(p);
  } else {
    HC(hipFree(p));
  }
  if (trace_on())
    std::fprintf(stderr, "gls: device free in %.1f ms\n",
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
}

void* Job::dmalloc(size_t bytes) {
  void* p = nullptr;
  bytes = std::max<size_t>(bytes, 256);
  if (static_) throw GlsError(GLS_E_STATE, "device allocation during an allocation-free execution");
  ++allocs_;
  if (!dev_alloc(&p, bytes)) HC(hipErrorOutOfMemory);
  owned_.push_back(p);
  dev_bytes_ += bytes;
  return p;
}

void Job::ensure(void** p, size_t& cap, size_t need) {
  if (try_ensure(p, cap, need)) return;
  if (static_)
    throw GlsError(GLS_E_RANGE, "the run needs more scratch than the allocation-free execution's "
                                "static array holds");
  HC(hipErrorOutOfMemory);
}

// As ensure(), but an allocation that fails leaves the buffer released
// (cap 0) and returns false instead of throwing.
bool Job::try_ensure(void** p, size_t& cap, size_t need) {
  if (need <= cap && *p) return true;
  // the allocation-free execution never frees or allocates: a buffer of the
  // static scratch that is too small fails the request
  if (static_) return false;
  ++allocs_;
  size_t n = std::max<size_t>(need, cap + cap / 2);
  n = std::max<size_t>(n, 4096);
  dev_free(*p);
  *p = nullptr;
  dev_bytes_ -= cap;
  cap = 0;
  if (!dev_alloc(p, n)) {
    if (n == need || !dev_alloc(p, need)) return false;
    n = need;
  }
  dev_bytes_ += n;
  cap = n;
  return true;
}

// Grows an automatically sized pool to at least `need` words, keeping its
// first `used` words (the waveforms so far; offsets into the pool stay
// valid).  Doubles when the device has room, else takes exactly `need`.
// Returns whether the pool now holds `need` words; a caller-sized pool never
// grows.  A captured graph holds the old pool address and is dropped.
bool Job::grow_pool(uint64_t need, uint64_t used) {
  if (need <= pool_words_) return true;
  if (!pool_auto_ || static_) return false;
  size_t free_b = 0, total_b = 0;
  HC(hipMemGetInfo(&free_b, &total_b));
  // the old pool stays during the copy; cached blocks make room when needed
  const uint64_t room = (uint64_t)((free_b + (arena_ ? arena_->cached_bytes() : 0)) * 0.9) / 4;
  uint64_t n = std::max<uint64_t>(need, 2 * pool_words_);
  if (n > room) n = need;
  if (n > room) return false;
  uint32_t* p = nullptr;
  if (!dev_alloc((void**)&p, (n + kPoolPad) * sizeof(uint32_t))) return false;
  ++pool_reallocs_;
  ++allocs_;
  if (used) HC(hipMemcpyAsync(p, d_pool_, used * 4, hipMemcpyDeviceToDevice, stream_));
  HC(hipStreamSynchronize(stream_));
  dev_free(d_pool_);
  d_pool_ = p;
  dev_bytes_ += (n - pool_words_) * sizeof(uint32_t);
  pool_words_ = n;
  if (gexec_) {
    HC(hipGraphExecDestroy(gexec_));
    gexec_ = nullptr;
    graph_sig_.clear();
  }
  return true;
}

// Largest pool an automatically sized pool may take between chunks (when
// nothing in it needs keeping): 80% of the free HBM and the arena's cached
// blocks, plus the current pool.
uint64_t Job::pool_max_words() const {
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  const double cached = arena_ ? (double)arena_->cached_bytes() : 0.0;
  return (uint64_t)((free_b + cached + pool_words_ * 4.0) * 0.8) / 4;
}

// Between chunks (the pool holds nothing to keep): replaces an automatically
// sized pool by one of `words` words when that is larger (or the largest
// that can be allocated above the current size).  Returns whether the pool
// grew.
bool Job::refresh_pool(uint64_t words) {
  // a new pool only for a real gain: each allocation of a large pool costs
  // host time
  if (!pool_auto_ || static_ || words <= pool_words_ + pool_words_ / 4) return false;
  ++pool_reallocs_;
  ++allocs_;
  HC(hipStreamSynchronize(stream_));
  dev_free(d_pool_);
  d_pool_ = nullptr;
  dev_bytes_ -= (pool_words_ + kPoolPad) * sizeof(uint32_t);
  uint64_t n = words;
  while (n > pool_words_) {
    if (dev_alloc((void**)&d_pool_, (n + kPoolPad) * sizeof(uint32_t))) break;
    d_pool_ = nullptr;
    n = n / 10 * 9;
  }
  if (!d_pool_) {
    n = pool_words_;
    if (!dev_alloc((void**)&d_pool_, (n + kPoolPad) * sizeof(uint32_t))) HC(hipErrorOutOfMemory);
  }
  dev_bytes_ += (n + kPoolPad) * sizeof(uint32_t);
  const bool grew = n > pool_words_;
  pool_words_ = n;
  if (gexec_) {
    HC(hipGraphExecDestroy(gexec_));
    gexec_ = nullptr;
    graph_sig_.clear();
  }
  return grew;
}

// The resident stimulus store (64-bit words (time << 2) | value, signal
// after signal) goes to the device through two pinned staging buffers:
// host threads convert block b + 1 while block b is copied.  The times of
// every signal are checked to increase on the way.
void Job::upload_store(const std::vector<SigSrc>& src, const std::vector<uint64_t>& soff,
                       uint64_t total) {
  if (total == 0) return;
  const uint64_t B = std::min<uint64_t>(total, kStageWords);
  uint64_t* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  std::atomic<bool> unordered{false};
  auto release = [&]() {
    for (int i = 0; i < 2; ++i) {
      if (done[i]) (void)hipEventDestroy(done[i]);
      if (buf[i]) (void)hipHostFree(buf[i]);
    }
  };
  try {
    for (int i = 0; i < 2; ++i) {
      HC(hipHostMalloc((void**)&buf[i], B * 8, hipHostMallocDefault));
      HC(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
    }
    const int nth = (int)std::max<uint64_t>(1, std::min<uint64_t>(host_threads(0), B / 65536));
    for (uint64_t b0 = 0, k = 0; b0 < total; b0 += B, ++k) {
      const uint64_t n = std::min(B, total - b0);
      uint64_t* dst = buf[k & 1];
      if (used[k & 1]) HC(hipEventSynchronize(done[k & 1]));
      // words [b0, b0 + n), cut into one stretch per thread
      auto fill = [&](uint64_t x0, uint64_t x1) {
        size_t sig = (size_t)(std::upper_bound(soff.begin(), soff.end(), x0) - soff.begin()) - 1;
        for (uint64_t x = x0; x < x1;) {
          while (sig + 1 < soff.size() && soff[sig + 1] <= x) ++sig;
          const SigSrc& S = src[sig];
          const uint64_t q0 = x - soff[sig];
          const uint64_t q1 = std::min<uint64_t>(S.n, q0 + (x1 - x));
          for (uint64_t q = q0; q < q1; ++q) {
            dst[x - b0 + (q - q0)] = ((uint64_t)S.t[q] << 2) | (S.v[q] & 3u);
            if (q > 0 && S.t[q] <= S.t[q - 1]) unordered.store(true, std::memory_order_relaxed);
          }
          x += q1 - q0;
        }
      };
      if (nth == 1) {
        fill(b0, b0 + n);
      } else {
        std::vector<std::thread> th;
        for (int i = 0; i < nth; ++i)
          th.emplace_back(fill, b0 + n * i / nth, b0 + n * (i + 1) / nth);
        for (auto& t : th) t.join();
      }
      HC(hipMemcpyAsync(d_store_ + b0, dst, n * 8, hipMemcpyHostToDevice, stream_));
      HC(hipEventRecord(done[k & 1], stream_));
      used[k & 1] = true;
    }
    HC(hipStreamSynchronize(stream_));
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    release();
    throw;
  }
  release();
  if (unordered.load()) throw GlsError(GLS_E_FORMAT, "stimulus times must increase");
}

void Job::prepare(const std::vector<SimInstance>& inst, uint64_t pool_words) {
  const auto t_prep = std::chrono::steady_clock::now();
  HC(hipSetDevice(device_));
  if (inst.empty()) throw GlsError(GLS_E_ARG, "no instance to simulate");
  if (inst.size() > 65535) throw GlsError(GLS_E_ARG, "at most 65535 instances per batch");
  if (mode == 1 && inst.size() > 1)
    throw GlsError(GLS_E_UNSUPPORTED, "the two-pass path simulates one instance at a time");
  // program images and net spaces
  std::vector<std::shared_ptr<DevProgram>> parts;
  net_base_.clear();
  win_.clear();
  uint64_t nets = 0;
  for (const SimInstance& in : inst) {
    if (in.end <= in.start) throw GlsError(GLS_E_ARG, "dumpoff time must be above dumpon time");
    auto img = in.image ? in.image : build_program_image(*in.pg, (uint32_t)in.st->timescale_ps);
    if (img->timescale_ps != (uint32_t)in.st->timescale_ps)
      throw GlsError(GLS_E_ARG, "program image built for another time scale");
    parts.push_back(img);
    net_base_.push_back((uint32_t)nets);
    uint64_t n = img->n_nets;
    for (int sn : in.stim_nets) n = std::max<uint64_t>(n, (uint64_t)sn + 1);
    nets += n;
    if (nets >= (1ull << 32) - 1) throw GlsError(GLS_E_RANGE, "batch exceeds 2^32 nets");
    win_.push_back(in.start);
    win_.push_back(in.end);
  }
  // Instance groups of a batch: consecutive instances whose estimated
  // waveforms (stimulus, and each gate about ten times its inputs' mean
  // events: on the generator's netlists about three times what they make)
  // fit the HBM that can hold the event pool.  Each group runs its own
  // sequence of time chunks (usually one), so per-gate work is not repeated
  // once per time chunk of the whole batch.
  std::vector<uint32_t> gsplit{0};
  std::vector<double> inst_est(inst.size());
  for (size_t i = 0; i < inst.size(); ++i) {
    double stim = 0;
    for (const Wave& w : inst[i].st->waves) stim += (double)w.size();
    const double per_input = stim / (double)std::max<size_t>(1, inst[i].st->waves.size()) + 1;
    inst_est[i] = stim + 10.0 * (double)parts[i]->n_gates * per_input;
  }
  if (mode == 0 && inst.size() > 1) {
    size_t free_b = 0, total_b = 0;
    HC(hipMemGetInfo(&free_b, &total_b));
    const double cached = arena_ ? (double)arena_->cached_bytes() : 0.0;  // reusable too
    const double target = pool_words ? (double)pool_words : 0.8 * ((double)free_b + cached) / 4.0;
    double acc = 0;
    for (size_t i = 0; i < inst.size(); ++i) {
      const double est = inst_est[i];
      if (i > gsplit.back() && acc + est > target) {
        gsplit.push_back((uint32_t)i);
        acc = 0;
      }
      acc += est;
    }
  }
  gsplit.push_back((uint32_t)inst.size());
  prog_ = parts.size() == 1 ? parts[0] : union_image(parts, net_base_, gsplit);
  if (mode == 1 && prog_->max_k > kTileK)
    throw GlsError(GLS_E_UNSUPPORTED, "the two-pass path takes gates of at most 4 distinct inputs");
  try {
    prog_->upload(device_, stream_);
  } catch (const GlsError& e) {
    // cached blocks of finished runs may hold the room the image needs
    if (e.code != GLS_E_OOM || !arena_) throw;
    arena_->trim();
    prog_->upload(device_, stream_);
  }
  if (arena_) arena_->keep_headroom();
  n_gates_ = prog_->n_gates;
  n_nets_ = (uint32_t)nets;
  level_start_ = prog_->level_start;
  wmax_ = 1;
  for (const GateRec& g : prog_->h_grec) wmax_ = std::max<uint32_t>(wmax_, g.dmax);

  // stimulus signals, then constants, per instance
  saif_net_list_.clear();
  saif_net_job_.clear();
  std::vector<uint32_t> words;    // two-pass: initial waveforms in the pool
  std::vector<SigSrc> src;        // fused: the signals of the resident store, in order
  uint64_t store_words = 0;
  std::vector<uint64_t> soff;
  h_slen_.clear();
  h_snet_.clear();
  std::vector<uint32_t> sjob;
  h_woff_init_.clear();
  h_wlen_init_.clear();
  if (mode == 1) {
    h_woff_init_.assign(n_nets_, 0);
    h_wlen_init_.assign(n_nets_, 0);
  }
  t_last_ = 0;
  auto add_signal = [&](uint32_t net, uint32_t job, const uint32_t* t, const uint8_t* v, size_t n) {
    if (mode == 1) {
      h_woff_init_[net] = words.size();
      h_wlen_init_[net] = (uint32_t)n;
      for (size_t q = 0; q < n; ++q) {
        if (t[q] >= (1u << 30) - (1u << 17))
          throw GlsError(GLS_E_RANGE, "the two-pass path keeps times below 2^30 - 2^17");
        words.push_back((t[q] << 2) | (v[q] & 3u));
      }
    } else {
      soff.push_back(store_words);
      h_slen_.push_back((uint32_t)n);
      h_snet_.push_back(net);
      sjob.push_back(job);
      src.push_back({t, v, n});
      store_words += n;
      if (n) t_last_ = std::max<uint64_t>(t_last_, t[n - 1]);
    }
    saif_net_list_.push_back(net);
    saif_net_job_.push_back(job);
  };
  // constants: one entry (0, value) each; the four possible entries live here
  static const uint32_t kT0[1] = {0};
  static const uint8_t kV[4] = {0, 1, 2, 3};
  std::vector<uint32_t> sig_start(inst.size() + 1, 0);
  std::vector<uint64_t> inst_t_last(inst.size(), 0);
  for (size_t i = 0; i < inst.size(); ++i) {
    const SimInstance& in = inst[i];
    const uint32_t nb = net_base_[i];
    sig_start[i] = (uint32_t)h_snet_.size();
    for (const Wave& w : in.st->waves)
      if (w.size()) inst_t_last[i] = std::max<uint64_t>(inst_t_last[i], w.t[w.size() - 1]);
    for (size_t s = 0; s < in.st->waves.size(); ++s) {
      const Wave& w = in.st->waves[s];
      if (mode == 1)
        for (size_t q = 1; q < w.size(); ++q)
          if (w.t[q] <= w.t[q - 1]) throw GlsError(GLS_E_FORMAT, "stimulus times must increase");
      add_signal(nb + (uint32_t)in.stim_nets[s], (uint32_t)i, w.t.data(), w.v.data(), w.size());
    }
    const Program& pg = *in.pg;
    for (uint32_t n = 0; n < (uint32_t)pg.net_names.size(); ++n) {
      if (pg.role[n] != NR_CONST) continue;
      add_signal(nb + n, (uint32_t)i, kT0, &kV[pg.const_val[n] & 3], 1);
    }
  }
  n_sig_ = (uint32_t)h_snet_.size();
  sig_start[inst.size()] = n_sig_;
  const size_t ngroups = gsplit.size() - 1;
  const size_t nlev = (level_start_.size() - 1) / ngroups;
  horizon_ = t_last_ + (uint64_t)(nlev + 1) * wmax_ + 2;
  groups_.clear();
  for (size_t q = 0; q < ngroups; ++q) {
    Group G;
    G.i0 = gsplit[q];
    G.i1 = gsplit[q + 1];
    G.L0 = (uint32_t)(q * nlev);
    G.L1 = (uint32_t)((q + 1) * nlev);
    G.g0 = level_start_[G.L0];
    G.g1 = level_start_[G.L1];
    G.s0 = mode == 0 ? sig_start[G.i0] : 0;
    G.s1 = mode == 0 ? sig_start[G.i1] : n_sig_;
    uint64_t tl = 0;
    for (uint32_t i = G.i0; i < G.i1; ++i) tl = std::max(tl, inst_t_last[i]);
    G.horizon = tl + (uint64_t)(nlev + 1) * wmax_ + 2;
    G.est_words = 0;
    for (uint32_t i = G.i0; i < G.i1; ++i) G.est_words += inst_est[i];
    groups_.push_back(G);
  }

  d_woff_ = (uint64_t*)dmalloc(n_nets_ * sizeof(uint64_t));
  d_wlen_ = (uint32_t*)dmalloc(n_nets_ * sizeof(uint32_t));
  d_saif_ = (unsigned long long*)dmalloc((size_t)n_nets_ * kSaifWords * sizeof(unsigned long long));
  d_win_ = (unsigned long long*)dmalloc(win_.size() * 8);
  d_scal_ = (uint64_t*)dmalloc(16 * sizeof(uint64_t));
  uint32_t maxw = 1;
  for (size_t l = 0; l + 1 < level_start_.size(); ++l)
    maxw = std::max(maxw, level_start_[l + 1] - level_start_[l]);
  for (uint64_t** p : {&d_outoff_, &d_ntile_, &d_bound_, &d_tilebase_, &d_scrbase_})
    *p = (uint64_t*)dmalloc(maxw * 8ull);
  if (mode == 1) {
    for (uint64_t** p : {&d_nseg_, &d_scr_, &d_segbase_, &d_outlen_})
      *p = (uint64_t*)dmalloc(maxw * 8ull);
    d_saif_nets_ = (uint32_t*)dmalloc(saif_net_list_.size() * 4);
    d_saif_job_ = (uint32_t*)dmalloc(saif_net_job_.size() * 4);
  }
  d_part_ = (uint64_t*)dmalloc(scan_part_size(maxw) * 8ull);
  d_gnseg_ = (uint32_t*)dmalloc(maxw * 4ull);
  d_plan_ = (GatePlan*)dmalloc(maxw * sizeof(GatePlan));
  // stacks of tiles that do not fit LDS; an automatic size grows when a
  // level runs out of it (the level is run again)
  ovf_words_ = ovf_words_opt ? ovf_words_opt : (1ull << 24);
  ensure((void**)&d_ovf_, cap_ovf_, ovf_words_ * 4);
  if (mode == 0) {
    d_store_ = (uint64_t*)dmalloc(std::max<uint64_t>(store_words, 1) * 8);
    d_soff_ = (uint64_t*)dmalloc(n_sig_ * 8ull);
    d_slen_ = (uint32_t*)dmalloc(n_sig_ * 4ull);
    d_snet_ = (uint32_t*)dmalloc(n_sig_ * 4ull);
    d_sjob_ = (uint32_t*)dmalloc(n_sig_ * 4ull);
    d_scnt_ = (uint64_t*)dmalloc(n_sig_ * 8ull);
    d_sbase_ = (uint64_t*)dmalloc(n_sig_ * 8ull);
    d_sfirst_ = (uint32_t*)dmalloc(n_sig_ * 4ull);
    d_spart_ = (uint64_t*)dmalloc(scan_part_size(n_sig_) * 8ull);
    d_tail_len_ = (uint32_t*)dmalloc(n_nets_ * 4ull);
    d_tail_off_ = (uint64_t*)dmalloc(n_nets_ * 8ull);
    for (uint64_t** p : {&d_tlen_g_, &d_tsrc_g_, &d_toff_g_}) *p = (uint64_t*)dmalloc(n_gates_ * 8ull);
    d_gpart_ = (uint64_t*)dmalloc(scan_part_size(n_gates_) * 8ull);
    d_saif_chunk_ = (unsigned long long*)dmalloc((size_t)n_nets_ * kSaifWords * 8);
    d_wtot_ = (unsigned long long*)dmalloc(n_nets_ * 8ull);
    d_evjob_ = d_evjob_chunk_ = nullptr;
    if (inst.size() > 1) {  // gate events per instance
      d_evjob_ = (unsigned long long*)dmalloc(inst.size() * 8ull);
      d_evjob_chunk_ = (unsigned long long*)dmalloc(inst.size() * 8ull);
    }
    ensure((void**)&d_tails_, cap_tails_, std::max<size_t>(n_gates_, 1) * 4);
    ensure((void**)&d_cwin_, cap_cwin_, inst.size() * sizeof(ChunkWin));
    stim_words_ = store_words + n_sig_;  // one chunk: every entry and a state entry per signal
  } else {
    stim_words_ = words.size();
  }

  // event pool: a caller-given size is fixed unless pool_grow is set; a
  // growable pool is enlarged between levels and chunks as needed.  Size 0:
  // sized from the waveforms expected (the stimulus plus ten times each
  // instance's mean input events per gate, at least four times the
  // stimulus and 64 MiB), at most 85% of the HBM free or cached; with an
  // arena the largest cached block is taken whole when it holds the least
  // size, so a run reuses the last run's pool.  (Large device allocations
  // and frees can take seconds each: growing a pool step by step from a
  // small start cost more than the simulation on large netlists.)
  pool_auto_ = pool_words == 0 || pool_grow;
  bool pool_taken = false;
  if (pool_words == 0) {
    size_t free_b = 0, total_b = 0;
    HC(hipMemGetInfo(&free_b, &total_b));
    const double room = 0.85 * ((double)free_b + (arena_ ? (double)arena_->cached_bytes() : 0.0)) / 4;
    double est = 0;
    for (const SimInstance& I : inst) {
      uint64_t stim = 0;
      for (auto& w : I.st->waves) stim += w.size();
      const uint64_t n_in = std::max<uint64_t>(1, I.st->waves.size());
      est += (double)stim + 10.0 * (double)I.pg->gates.size() * (double)(stim / n_in + 1);
    }
    // the hard minimum (the two-pass path keeps the stimulus in the pool)
    // applies after the clamp to the room: a taken block is never enlarged
    const uint64_t need = mode == 1 ? stim_words_ + 1024 : 1024;
    uint64_t least = std::min<uint64_t>(std::max<uint64_t>(1ull << 24, 4 * stim_words_), (uint64_t)room);
    least = std::max(least, need);
    const uint64_t want = std::max<uint64_t>(least, (uint64_t)std::min(est, room));
    if (arena_) {
      size_t got = 0;
      if (!arena_->get_large((want + kPoolPad) * sizeof(uint32_t), (least + kPoolPad) * sizeof(uint32_t),
                             (void**)&d_pool_, &got))
        HC(hipErrorOutOfMemory);
      pool_words = got / sizeof(uint32_t) - kPoolPad;  // >= least: got >= (least + kPoolPad) * 4
      pool_taken = true;
    } else {
      pool_words = want;
    }
  }
  if (mode == 1 && pool_words < stim_words_ + 1024) pool_words = stim_words_ + 1024;
  if (pool_words < 1024) pool_words = 1024;
  pool_words_ = pool_words;
  if (!pool_taken && !dev_alloc((void**)&d_pool_, (pool_words_ + kPoolPad) * sizeof(uint32_t)))
    HC(hipErrorOutOfMemory);
  dev_bytes_ += (pool_words_ + kPoolPad) * sizeof(uint32_t);

  if (mode == 1) {
    if (!words.empty())
      HC(hipMemcpyAsync(d_pool_, words.data(), words.size() * 4, hipMemcpyHostToDevice, stream_));
    if (!saif_net_list_.empty()) {
      HC(hipMemcpyAsync(d_saif_nets_, saif_net_list_.data(), saif_net_list_.size() * 4,
                        hipMemcpyHostToDevice, stream_));
      HC(hipMemcpyAsync(d_saif_job_, saif_net_job_.data(), saif_net_job_.size() * 4,
                        hipMemcpyHostToDevice, stream_));
    }
  } else {
    upload_store(src, soff, store_words);
    if (n_sig_) {
      HC(hipMemcpyAsync(d_soff_, soff.data(), n_sig_ * 8ull, hipMemcpyHostToDevice, stream_));
      HC(hipMemcpyAsync(d_slen_, h_slen_.data(), n_sig_ * 4ull, hipMemcpyHostToDevice, stream_));
      HC(hipMemcpyAsync(d_snet_, h_snet_.data(), n_sig_ * 4ull, hipMemcpyHostToDevice, stream_));
      HC(hipMemcpyAsync(d_sjob_, sjob.data(), n_sig_ * 4ull, hipMemcpyHostToDevice, stream_));
    }
  }
  HC(hipMemcpyAsync(d_win_, win_.data(), win_.size() * 8, hipMemcpyHostToDevice, stream_));
  HC(hipStreamSynchronize(stream_));
  cap_ = tile_cap(seg_target ? seg_target : kDefaultSeg);
  level_blocks_ = level_blocks_opt ? (int)level_blocks_opt : level_max_blocks(cap_, stack_mode == 0);
  stats_ = JobStats();
  stats_.level_blocks = (uint64_t)level_blocks_;
  stats_.n_nets = n_nets_;
  stats_.n_gates = n_gates_;
  stats_.n_pins = prog_->n_pins;
  if (trace_on())
    std::fprintf(stderr, "gls: prepared %zu instance(s), pool %.1f GiB, in %.1f ms\n", inst.size(),
                 pool_words_ * 4.0 / 1073741824.0,
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_prep).count());
}

void Job::stop_point() {
  if (should_stop && should_stop()) throw GlsError(GLS_E_CANCELLED, "job cancelled or past its deadline");
}

// Chunk k covering [F, Fn): its look-back start, its relative-time bounds
// and the move of the previous chunk's tails (made in the previous chunk's
// time) to this chunk's time.
Job::Chunk Job::make_chunk(size_t k, uint64_t F, uint64_t Fn, uint64_t prev_lb) const {
  Chunk c;
  c.F = F;
  c.Fn = Fn;
  c.lb = k == 0 ? 0 : F - wmax_;
  c.first = k == 0 ? 1u : 0u;
  c.lo_rel = (uint32_t)(F - c.lb + 1);
  c.hi_rel = Fn == kChunkEnd ? 0xFFFFFFFFu : (uint32_t)(Fn - c.lb + 1);
  c.shift = k == 0 ? 0u : (uint32_t)(c.lb - prev_lb);
  c.slot = 0;
  return c;
}

// The chunk's SAIF bounds per instance, in its relative time, into ChunkWin
// slot c.slot.
void Job::put_chunk_windows(const Chunk& c, bool capture) {
  const size_t n = net_base_.size();
  const uint64_t kMax = 1ull << 62;
  auto rel = [&](uint64_t x) -> uint64_t {  // x >= lb - 1
    const uint64_t r = x - c.lb + 1;
    return (x - c.lb >= kMax) ? kMax : r;
  };
  std::vector<ChunkWin> w(n);
  for (size_t i = 0; i < n; ++i) {
    const uint64_t s = win_[2 * i], e = win_[2 * i + 1];
    const uint64_t lo = std::max(s, c.F), hi = std::min(e, c.Fn);
    ChunkWin& x = w[i];
    if (hi <= lo) {
      x.lo = x.hi = 0;
    } else {
      x.lo = rel(lo);
      x.hi = rel(hi);
    }
    // toggles at t in (s, e]: t_rel > rel(s) and t_rel <= rel(e), where
    // every entry's t_rel is >= 1 (a bound before lb - 1 counts as 0)
    x.tc_lo = s < c.lb ? 0 : rel(s);
    x.tc_hi = e < c.lb ? 0 : rel(e);
  }
  const size_t need = (c.slot + 1) * n * sizeof(ChunkWin);
  if (capture) {
    if (need > cap_cwin_) throw GlsError(GLS_E_STATE, "chunk windows not sized for the capture");
  } else {
    ensure((void**)&d_cwin_, cap_cwin_, need);
  }
  HC(hipMemcpyAsync(d_cwin_ + c.slot * n, w.data(), n * sizeof(ChunkWin), hipMemcpyHostToDevice,
                    stream_));
  HC(hipStreamSynchronize(stream_));
}

// One chunk: its stimulus cut into the pool, then every level.  Per level:
// plan the tiles (k_plan2, two scans), describe them (k_tile_desc), walk
// every tile independently into the level scratch (k_level), join the tiles
// of each gate and finish its SAIF share (k_gate_carry), place the gates in
// the pool behind their tails (scan) and copy the outputs there
// (k_compact).  Plain: one host round trip per level sizes the tile arrays
// and the scratch.  Device-driven (graph capture): no host round trip; the
// arrays keep the capacities learned by an earlier plain run, kernels read
// the level's counts from device memory, and k_guard skips a level that
// would not fit (reported as an error after the run).
//
// Plain runs with an automatically sized overflow stack buffer check, at
// the next host round trip, whether the previous level ran out of it; if so
// the buffer doubles and that level runs again from its starting state
// (bump pointer and event count; its outputs and SAIF shares are
// rewritten).  A plain chunk that cannot fit the pool (or, with automatic
// chunks, is projected not to) returns CHUNK_SHORTER before its results
// are used; *grow_hint then suggests a length factor (0: none).
int Job::run_chunk(const Chunk& c, const Group& G, std::vector<hipEvent_t>& ev_eval,
                   bool device_driven, double* grow_hint) {
  uint64_t* bump = d_scal_ + S_BUMP;
  unsigned long long* evc = (unsigned long long*)(d_scal_ + S_EVC);
  unsigned int* err = (unsigned int*)(d_scal_ + S_ERR);
  const uint32_t seg = seg_target ? seg_target : kDefaultSeg;
  const bool retry = !device_driven && !static_ && (ovf_words_opt == 0 || pool_grow);
  const uint32_t ns = G.s1 - G.s0;  // the group's stimulus signals
  const ChunkWin* cw = d_cwin_ + c.slot * net_base_.size();
  TailArgs T;
  T.tail_len = d_tail_len_;
  T.tail_off = d_tail_off_;
  T.tails = d_tails_;
  T.shift = c.shift;
  T.first = c.first;
  *grow_hint = 0;

  if (d_evjob_chunk_) HC(launch_fill_u64(d_evjob_chunk_, net_base_.size(), 0ull, stream_));
  // the chunk's stimulus at the start of the pool
  HC(launch_stim_count(d_store_, d_soff_ + G.s0, d_slen_ + G.s0, ns, c.lb,
                       c.Fn == kChunkEnd ? ~0ull : c.Fn, d_scnt_, d_sfirst_, stream_));
  if (ns) HC(launch_scan(d_scnt_, d_sbase_, ns, d_spart_, bump, nullptr, stream_));
  else HC(launch_fill_u64((unsigned long long*)bump, 1, 0ull, stream_));
  uint64_t stim_chunk = 0;
  if (!device_driven) {
    HC(hipMemcpyAsync(&stim_chunk, bump, 8, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    if (stim_chunk + 1024 > pool_words_) {
      // nothing in the pool to keep yet
      if (pool_auto_) refresh_pool(std::max<uint64_t>(2 * stim_chunk, 2 * pool_words_));
      if (stim_chunk + 1024 > pool_words_) return CHUNK_SHORTER;
    }
  }
  HC(launch_stim_copy(d_store_, d_soff_ + G.s0, d_snet_ + G.s0, ns, c.lb, d_scnt_, d_sfirst_,
                      d_sbase_, d_pool_, pool_words_, d_woff_, d_wlen_, err, stream_));

  size_t prev = SIZE_MAX;  // last level run
  uint64_t prev_bump = 0, prev_evc = 0, prev_err = 0;
  // true when the previous level overflowed and was rolled back
  auto rolled_back = [&](const uint64_t* h) {
    if (!retry || prev == SIZE_MAX || !(h[S_ERR] & 2u)) return false;
    uint64_t s[8];
    std::memcpy(s, h, sizeof s);
    s[S_BUMP] = prev_bump;
    s[S_EVC] = prev_evc;
    s[S_ERR] = prev_err;
    HC(hipMemcpyAsync(d_scal_, s, sizeof s, hipMemcpyHostToDevice, stream_));
    HC(hipStreamSynchronize(stream_));
    ovf_words_ *= 2;
    ensure((void**)&d_ovf_, cap_ovf_, ovf_words_ * 4);
    if (gexec_) {  // a captured graph holds the old buffer
      HC(hipGraphExecDestroy(gexec_));
      gexec_ = nullptr;
      graph_sig_.clear();
    }
    return true;
  };
  const bool projected = !device_driven && chunk_ps == 0;
  const uint64_t group_gates = G.g1 - G.g0;
  size_t next = G.L0;  // level to run next
  while (next <= G.L1) {
    const size_t l = next++;
    if (!device_driven) stop_point();
    if (l == G.L1) {  // after the last level: did it overflow?
      if (!retry || prev == SIZE_MAX) break;
      uint64_t h[8];
      HC(hipMemcpyAsync(h, d_scal_, sizeof h, hipMemcpyDeviceToHost, stream_));
      HC(hipStreamSynchronize(stream_));
      if (!rolled_back(h)) break;
      next = prev;  // run it again
      prev = SIZE_MAX;
      continue;
    }
    uint32_t g0 = level_start_[l], ng = level_start_[l + 1] - g0;
    if (ng == 0) continue;
    const GateRec* gates = prog_->d_grec + g0;
    HC(launch_plan2(gates, ng, d_wlen_, seg, d_plan_, d_ntile_, d_bound_, err, d_tail_len_, stream_));
    HC(launch_scan(d_ntile_, d_tilebase_, ng, d_part_, d_scal_ + S_TOT_A, nullptr, stream_));
    HC(launch_scan(d_bound_, d_scrbase_, ng, d_part_, d_scal_ + S_TOT_B, nullptr, stream_));
    uint64_t ntiles = tiles_cap_;
    // The level scratch lives in the unused top of the event pool when twice
    // its size is free above the bump (the level's outputs, at most the
    // scratch size, then land below it); otherwise in an allocation of its
    // own.  The pool guard keeps the outputs clear of a scratch in the pool.
    uint32_t* scratch = d_scratch_;
    uint64_t pool_limit = pool_words_;
    if (!device_driven) {
      uint64_t h[8];
      HC(hipMemcpyAsync(h, d_scal_, sizeof h, hipMemcpyDeviceToHost, stream_));
      HC(hipStreamSynchronize(stream_));
      if (rolled_back(h)) {
        next = prev;  // run the previous level again
        prev = SIZE_MAX;
        continue;
      }
      prev = l;
      prev_bump = h[S_BUMP];
      prev_evc = h[S_EVC];
      prev_err = h[S_ERR];
      ntiles = h[S_TOT_A];
      const uint64_t scr = h[S_TOT_B], used = h[S_BUMP];
      if (ntiles >= 0xFFFFFFF0ull) throw GlsError(GLS_E_RANGE, "too many tiles in one level");
      // automatic chunks: a chunk projected (by its gates so far) not to fit
      // the largest pool is cut short before it runs further
      const uint64_t done_g = g0 - G.g0;
      if (projected && done_g >= std::max<uint64_t>(64, group_gates / 32) && used > stim_chunk) {
        const double proj = (double)(used - stim_chunk) * group_gates / done_g;
        const double cap = (double)(pool_auto_ && !static_ ? pool_max_words() : pool_words_);
        if (((double)stim_chunk + proj) * 1.35 > cap) {
          *grow_hint = std::max(0.05, 0.85 * (cap / 1.35 - stim_chunk) / proj);
          return CHUNK_SHORTER;
        }
      }
      if (!try_ensure((void**)&d_tiles_, cap_tiles_, ntiles * sizeof(TileDesc)) ||
          !try_ensure((void**)&d_tout_, cap_tout_, ntiles * sizeof(TileOut)) ||
          !try_ensure((void**)&d_toff_, cap_toff_, ntiles * 4))
        return CHUNK_SHORTER;
      grow_pool(used + 2 * scr, used);
      if (used <= pool_words_ && 2 * scr <= pool_words_ - used) {
        pool_limit = pool_words_ - scr;
        scratch = d_pool_ + pool_limit;
      } else {
        // outputs in the pool, scratch apart
        if (used + scr > pool_words_ && !grow_pool(used + scr, used)) return CHUNK_SHORTER;
        if (!try_ensure((void**)&d_scratch_, cap_scratch_, scr * 4)) return CHUNK_SHORTER;
        scratch = d_scratch_;
      }
      max_tiles_ = std::max(max_tiles_, ntiles);
      max_scratch_ = std::max(max_scratch_, scr);
    } else {
      HC(launch_guard(d_scal_, tiles_cap_, scratch_cap_, stream_));
      if (graph_scratch_in_pool_) {
        pool_limit = pool_words_ - scratch_cap_;
        scratch = d_pool_ + pool_limit;
      }
    }
    HC(launch_tile_desc(gates, ng, d_plan_, d_tilebase_, d_woff_, d_wlen_, d_pool_, prog_->d_lutp,
                        d_tiles_, d_scal_ + S_TOT_A, ntiles, stream_));
    HC(launch_fill_u64((unsigned long long*)(d_scal_ + S_TICKET), 2, 0ull, stream_));  // ticket, ovf
    LevelArgs a;
    a.tiles = d_tiles_;
    a.pool = d_pool_;
    a.dtab = prog_->d_dtab;
    a.luts = prog_->d_luts;
    a.ntiles = d_scal_ + S_TOT_A;
    a.scr_base = d_scrbase_;
    a.scratch = scratch;
    a.tout = d_tout_;
    a.ovf = d_ovf_;
    a.ovf_cap = ovf_words_;
    a.ovf_used = (unsigned long long*)(d_scal_ + S_OVF_USED);
    a.ev_count = evc;
    a.err = err;
    a.cwin = cw;
    a.cap = cap_;
    a.prefetch = prefetch;
    a.ticket = (unsigned long long*)(d_scal_ + S_TICKET);
    a.tiles_walked = (unsigned long long*)(d_scal_ + S_TILES);
    a.stk = nullptr;
    a.lo_rel = c.lo_rel;
    a.hi_rel = c.hi_rel;
    a.first = c.first;
    const unsigned nb = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)std::max(1, level_blocks_));
    if (stack_mode == 1) {
      const uint64_t w = (uint64_t)std::max(1, level_blocks_) * (cap_ + 64);
      if (!device_driven) ensure((void**)&d_stk_, cap_stk_, w * 4);
      a.stk = d_stk_;
    }
    WideArgs wa;
    wa.gates = gates;
    wa.wide = prog_->d_wide + prog_->wide_start[l];
    wa.nwide = prog_->wide_start[l + 1] - prog_->wide_start[l];
    wa.g0 = g0;
    wa.tile_base = d_tilebase_;
    wa.tiles = d_tiles_;
    wa.woff = d_woff_;
    wa.wlen = d_wlen_;
    wa.pool = d_pool_;
    wa.luts = prog_->d_luts;
    wa.dtab = prog_->d_dtab;
    wa.scr_base = d_scrbase_;
    wa.scratch = scratch;
    wa.tout = d_tout_;
    wa.ev_count = evc;
    wa.err = err;
    wa.cwin = cw;
    wa.lo_rel = c.lo_rel;
    wa.hi_rel = c.hi_rel;
    wa.first = c.first;
    if (!device_driven && !static_) {
      hipEvent_t e_a, e_b;
      HC(hipEventCreate(&e_a));
      HC(hipEventCreate(&e_b));
      ev_eval.push_back(e_a);
      ev_eval.push_back(e_b);
      HC(hipEventRecord(e_a, stream_));
      HC(launch_level(a, nb, stream_));
      HC(launch_wide(wa, stream_));
      HC(hipEventRecord(e_b, stream_));
    } else {
      HC(launch_level(a, nb, stream_));
      HC(launch_wide(wa, stream_));
    }
    // gate counts go to d_bound_ (free again once the scratch bases are scanned)
    HC(launch_gate_carry(gates, ng, d_tilebase_, d_ntile_, d_tout_, d_toff_, d_bound_, d_wlen_,
                         d_saif_chunk_, cw, err, T, d_evjob_chunk_, stream_));
    HC(launch_scan(d_bound_, d_outoff_, ng, d_part_, d_scal_ + S_TOT_OUT, bump, stream_));
    HC(launch_pool_guard(d_scal_, pool_limit, stream_));
    HC(launch_compact(d_tiles_, d_scal_ + S_TOT_A, ntiles, d_tout_, d_toff_, d_scrbase_, scratch,
                      d_outoff_, d_pool_, d_woff_, d_wlen_, err, d_saif_chunk_, cw, saif_tc, T,
                      stream_));
    HC(launch_bump(bump, d_scal_ + S_TOT_OUT, stream_));
  }
  return CHUNK_OK;
}

// Gate-output waveforms of a finished chunk to the host: each gate net's
// entries after its tail, in absolute time.
void Job::fetch_chunk_waves(const Chunk& c, const Group& G) {
  std::vector<uint64_t> off(n_nets_);
  std::vector<uint32_t> len(n_nets_), tl(n_nets_);
  uint64_t used = 0;
  HC(hipMemcpyAsync(off.data(), d_woff_, n_nets_ * 8ull, hipMemcpyDeviceToHost, stream_));
  HC(hipMemcpyAsync(len.data(), d_wlen_, n_nets_ * 4ull, hipMemcpyDeviceToHost, stream_));
  HC(hipMemcpyAsync(tl.data(), d_tail_len_, n_nets_ * 4ull, hipMemcpyDeviceToHost, stream_));
  HC(hipMemcpyAsync(&used, d_scal_ + S_BUMP, 8, hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
  std::vector<uint32_t> pool(used);
  if (used) HC(hipMemcpyAsync(pool.data(), d_pool_, used * 4, hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
  if (waves_.size() != n_nets_) waves_.assign(n_nets_, {});
  for (uint32_t gi = G.g0; gi < G.g1; ++gi) {
    const uint32_t n = prog_->h_grec[gi].out_net;
    if (off[n] + len[n] > used) throw GlsError(GLS_E_STATE, "waveform outside the event pool");
    for (uint32_t q = tl[n]; q < len[n]; ++q) {
      const uint32_t w = pool[off[n] + q];
      const uint64_t t = (uint64_t)(w >> 2) + c.lb - 1;  // rel >= 1
      waves_[n].push_back((t << 2) | (w & 3u));
    }
  }
}

// After a chunk: its SAIF shares join the totals, its waveforms go to the
// host (keep_waves, plain runs), and the gate outputs' tails for the next
// chunk are cut from their waveforms.
void Job::end_chunk(const Chunk& c, const Group& G, bool device_driven) {
  const GateRec* gates = prog_->d_grec + G.g0;
  const uint32_t ng = G.g1 - G.g0;
  HC(launch_saif_accum(gates, ng, d_saif_chunk_, d_saif_, d_wlen_, d_tail_len_, d_wtot_, stream_));
  if (d_evjob_) HC(launch_add_u64(d_evjob_, d_evjob_chunk_, net_base_.size(), stream_));
  if (!device_driven && keep_waves) fetch_chunk_waves(c, G);
  if (c.Fn == kChunkEnd || ng == 0) return;
  const uint32_t key = (uint32_t)((c.Fn - wmax_) - c.lb + 1);  // lb of the next chunk
  HC(launch_tail_len(gates, ng, d_woff_, d_wlen_, d_pool_, key, d_tlen_g_, d_tsrc_g_, stream_));
  HC(launch_scan(d_tlen_g_, d_toff_g_, ng, d_gpart_, d_scal_ + S_TAILS, nullptr, stream_));
  if (!device_driven) {
    uint64_t total = 0;
    HC(hipMemcpyAsync(&total, d_scal_ + S_TAILS, 8, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    ensure((void**)&d_tails_, cap_tails_, std::max<uint64_t>(total, 1) * 4);
    max_tails_ = std::max(max_tails_, total);
  }
  HC(launch_tail_copy(gates, ng, d_tlen_g_, d_tsrc_g_, d_toff_g_, d_pool_, d_tails_, cap_tails_ / 4,
                      d_tail_len_, d_tail_off_, (unsigned int*)(d_scal_ + S_ERR), stream_));
}

// A plain run of every chunk.  Automatic chunks: the first chunk is the
// whole window, or shorter when the waveform estimate says it would not fit
// the pool; a chunk projected not to fit the largest pool (or that does not
// fit) is run again shorter, the pool taking the largest size free HBM
// allows, and each next chunk aims at 60% of that pool by the output rate of
// the one before.
void Job::run_chunks(std::vector<hipEvent_t>& ev_eval) {
  HC(launch_set_scal(d_scal_, 0, stream_));
  HC(launch_fill_u64(d_saif_, (uint64_t)n_nets_ * kSaifWords, 0ull, stream_));
  HC(launch_fill_u64(d_wtot_, n_nets_, 0ull, stream_));
  if (d_evjob_) HC(launch_fill_u64(d_evjob_, net_base_.size(), 0ull, stream_));
  HC(launch_saif_stim(d_store_, d_soff_, d_slen_, d_snet_, d_sjob_, n_sig_, d_saif_, d_win_, stream_));
  waves_.clear();
  if (keep_waves) waves_.assign(n_nets_, {});
  sched_.clear();
  gsched_.assign(groups_.size(), {});
  stats_.chunk_retries = 0;
  max_bump_ = 0;
  max_tails_ = 0;
  for (size_t q = 0; q < groups_.size(); ++q) {
    run_group_chunks(groups_[q], ev_eval);
    gsched_[q].assign(sched_.end() - (std::ptrdiff_t)gsched_count_, sched_.end());
  }
  stats_.chunks = sched_.size();
}

// The time chunks of one instance group, from its own time 0: the group's
// gate outputs start from X tails.
void Job::run_group_chunks(const Group& G, std::vector<hipEvent_t>& ev_eval) {
  const uint32_t ng = G.g1 - G.g0;
  ensure((void**)&d_tails_, cap_tails_, std::max<size_t>(ng, 1) * 4);
  HC(launch_tails_init(prog_->d_grec + G.g0, ng, d_tail_len_, d_tail_off_, d_tails_, stream_));
  max_tails_ = std::max<uint64_t>(max_tails_, ng);
  gsched_count_ = 0;
  const uint64_t horizon = G.horizon;
  const uint64_t lmax = kRelSpan - wmax_ - 8;  // relative times of a chunk stay in the event word
  const uint64_t lmin = wmax_ + 1;             // the look-back of chunk 1 starts after t = 0
  uint64_t L;
  if (chunk_ps == kNoChunks) {
    if (horizon >= lmax) throw GlsError(GLS_E_RANGE, "times reach 2^30 - 2^18: the run needs time chunks");
    L = horizon;
  } else if (chunk_ps) {
    if (chunk_ps < lmin)
      throw GlsError(GLS_E_ARG, "chunk length must exceed the largest gate delay");
    L = std::min<uint64_t>(chunk_ps, lmax);
  } else {
    L = std::min<uint64_t>(horizon, lmax);
    // first chunk from the waveform estimate (it runs one to three times
    // high on the generator's netlists): about the pool's worth at 1.2x the
    // estimated rate, the whole window when that covers most of it.  A first
    // chunk that does not fit is found out only after part of it has run
    // (projected from its first levels, then run again shorter): on the
    // 2M-cell workload two such attempts cost 3.6% of the run.
    if (G.est_words > 0) {
      // a growable pool counts at the largest size it may take
      const double pw = (double)(pool_auto_ && !static_ ? std::max(pool_words_, pool_max_words())
                                                        : pool_words_);
      const double l0 = 1.2 * pw * (double)horizon / G.est_words;
      if (l0 < 0.8 * (double)L) L = std::max<uint64_t>(lmin, (uint64_t)l0);
    }
  }
  // a chunk that did not fit caps every later chunk at the length it was
  // run again with: chunks after an empty stretch cannot grow back into a
  // dense one, and the number of retries is bounded (each at least halves)
  uint64_t l_cap = lmax;
  uint64_t F = 0, prev_lb = 0;
  size_t k = 0;
  for (;;) {
    const uint64_t Fn = (F + L >= horizon) ? kChunkEnd : F + L;
    const Chunk c = make_chunk(k, F, Fn, prev_lb);
    uint64_t snap[16];
    HC(hipMemcpyAsync(snap, d_scal_, sizeof snap, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    put_chunk_windows(c, false);
    double hint = 0;
    const auto tc0 = std::chrono::steady_clock::now();
    const uint64_t re0 = pool_reallocs_;
    const int res = run_chunk(c, G, ev_eval, false, &hint);
    if (trace_on()) {
      uint64_t b = 0;
      HC(hipMemcpyAsync(&b, d_scal_ + S_BUMP, 8, hipMemcpyDeviceToHost, stream_));
      HC(hipStreamSynchronize(stream_));
      std::fprintf(stderr, "gls: chunk %zu [%llu, %llu) L=%llu %s bump=%llu pool=%llu reallocs=%llu %.1f ms\n",
                   k, (unsigned long long)F, (unsigned long long)(Fn == kChunkEnd ? 0 : Fn),
                   (unsigned long long)L, res == CHUNK_OK ? "ok" : "shorter", (unsigned long long)b,
                   (unsigned long long)pool_words_, (unsigned long long)(pool_reallocs_ - re0),
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc0).count());
    }
    if (res != CHUNK_OK) {
      if (chunk_ps != 0)
        throw GlsError(GLS_E_OOM, "event pool exhausted (raise the pool size or allow time chunks)");
      HC(hipStreamSynchronize(stream_));
      HC(hipMemcpyAsync(d_scal_, snap, sizeof snap, hipMemcpyHostToDevice, stream_));
      uint64_t nl = hint > 0 ? (uint64_t)((double)L * std::min(0.5, hint)) : L / 2;
      if (L <= lmin) throw GlsError(GLS_E_OOM, "event pool exhausted by the shortest time chunk");
      L = std::max<uint64_t>(nl, lmin);
      l_cap = std::min(l_cap, L);
      ++stats_.chunk_retries;
      refresh_pool(pool_max_words());  // the chunk starts over: nothing to keep
      continue;
    }
    uint64_t used = 0;
    HC(hipMemcpyAsync(&used, d_scal_ + S_BUMP, 8, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    max_bump_ = std::max(max_bump_, used);
    end_chunk(c, G, false);
    sched_.push_back(Fn);
    ++gsched_count_;
    if (Fn == kChunkEnd) break;
    if (chunk_ps == 0) {
      // next length: 60% of the largest pool at this chunk's output rate
      // (the level scratch and the projection's margin need the rest)
      const double target = 0.6 * (double)(pool_auto_ && !static_ ? pool_max_words() : pool_words_);
      const double rate = (double)std::max<uint64_t>(used, 1) / (double)L;
      const double want = target / rate;
      L = (uint64_t)std::min<double>(std::max<double>(want, (double)L / 4), (double)L * 4);
      L = std::max<uint64_t>(std::min<uint64_t>(L, l_cap), lmin);
      if (pool_auto_ && used > pool_words_ / 2) refresh_pool(pool_max_words());
    }
    prev_lb = c.lb;
    F = Fn;
    ++k;
  }
}

void Job::run_two_pass(std::vector<hipEvent_t>& ev_eval) {
  uint64_t* bump = d_scal_ + S_BUMP;
  unsigned long long* evc = (unsigned long long*)(d_scal_ + S_EVC);
  const uint32_t seg = seg_target ? seg_target : 256;
  for (size_t l = 0; l + 1 < level_start_.size(); ++l) {
    stop_point();
    uint32_t g0 = level_start_[l], ng = level_start_[l + 1] - g0;
    if (ng == 0) continue;
    const GateDev* gates = prog_->d_gdev + g0;
    HC(launch_plan(gates, ng, d_wlen_, seg, d_nseg_, d_scr_, d_gnseg_, stream_));
    HC(launch_scan(d_nseg_, d_segbase_, ng, d_part_, d_scal_ + S_TOT_A, nullptr, stream_));
    HC(launch_scan(d_scr_, d_scrbase_, ng, d_part_, d_scal_ + S_TOT_B, nullptr, stream_));
    uint64_t h[3];
    HC(hipMemcpyAsync(h, d_scal_, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    uint64_t nsegs = h[S_TOT_A], nscr = h[S_TOT_B], used = h[S_BUMP];
    ensure((void**)&d_segmap_, cap_segmap_, nsegs * 4);
    ensure((void**)&d_sums_, cap_sums_, nsegs * sizeof(SegSum));
    ensure((void**)&d_carry_, cap_carry_, nsegs * sizeof(SegCarry));
    ensure((void**)&d_scratch_, cap_scratch_, nscr * 4);
    // the level's output is at most its scratch size; refuse before writing
    if (!grow_pool(used + nscr, used))
      throw GlsError(GLS_E_OOM, "event pool exhausted at level " + std::to_string(l) +
                                    " (raise the pool size or shorten the window)");
    HC(launch_segmap(ng, d_segbase_, d_gnseg_, d_segmap_, stream_));
    hipEvent_t e_a, e_b;
    HC(hipEventCreate(&e_a));
    HC(hipEventCreate(&e_b));
    ev_eval.push_back(e_a);
    ev_eval.push_back(e_b);
    HC(hipEventRecord(e_a, stream_));
    HC(launch_eval(gates, d_segbase_, d_scrbase_, d_gnseg_, d_segmap_, nsegs, d_woff_, d_wlen_,
                   d_pool_, prog_->d_luts, d_scratch_, d_sums_, evc, stream_));
    HC(hipEventRecord(e_b, stream_));
    HC(launch_gate_scan(gates, ng, d_segbase_, d_gnseg_, d_sums_, d_carry_, d_outlen_, d_saif_,
                        win_[0], win_[1], stream_));
    HC(launch_scan(d_outlen_, d_outoff_, ng, d_part_, d_scal_ + S_TOT_OUT, bump, stream_));
    HC(launch_emit(gates, d_segbase_, d_segmap_, nsegs, d_sums_, d_carry_, d_outoff_, d_outlen_,
                   d_scratch_, d_pool_, d_woff_, d_wlen_, d_saif_, win_[0], win_[1], stream_));
    HC(launch_bump(bump, d_scal_ + S_TOT_OUT, stream_));
  }
}

void Job::run() {
  HC(hipSetDevice(device_));
  cap_ = tile_cap(seg_target ? seg_target : kDefaultSeg);
  level_blocks_ = level_blocks_opt ? (int)level_blocks_opt : level_max_blocks(cap_, stack_mode == 0);
  stats_.level_blocks = (uint64_t)level_blocks_;
  const std::vector<uint64_t> sig = {cap_, (uint64_t)level_blocks_, seg_target, prefetch,
                                     (uint64_t)mode, stack_mode, chunk_ps, keep_waves};
  last_replay_ = false;
  if (graph && gexec_ && sig == graph_sig_) {
    // replay the captured chunk and level sequence: no host round trip
    auto w0 = std::chrono::steady_clock::now();
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipEventRecord(e0, stream_));
    HC(hipGraphLaunch(gexec_, stream_));
    HC(hipEventRecord(e1, stream_));
    HC(hipEventSynchronize(e1));
    float ms = 0;
    HC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    uint64_t h[16];
    HC(hipMemcpyAsync(h, d_scal_, sizeof h, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    if (h[S_ERR] & 8u)
      throw GlsError(GLS_E_STATE, "device-driven run outgrew the capacities of its first run");
    stats_.eval_ms = ms;
    stats_.eval_launches = 1;
    last_replay_ = true;
    finish_stats(h, ms);
    if (keep_waves) {  // one chunk (checked at capture): the pool holds every waveform
      waves_.assign(n_nets_, {});
      fetch_chunk_waves(make_chunk(0, 0, kChunkEnd, 0), groups_[0]);
    }
    stats_.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    return;
  }
  auto w0 = std::chrono::steady_clock::now();
  allocs_ = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!static_) {
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
  }
  if (mode == 1) {
    uint64_t scal[16] = {0};
    scal[S_BUMP] = stim_words_;
    HC(hipMemcpyAsync(d_woff_, h_woff_init_.data(), n_nets_ * 8ull, hipMemcpyHostToDevice, stream_));
    HC(hipMemcpyAsync(d_wlen_, h_wlen_init_.data(), n_nets_ * 4ull, hipMemcpyHostToDevice, stream_));
    HC(hipMemcpyAsync(d_scal_, scal, sizeof scal, hipMemcpyHostToDevice, stream_));
    HC(launch_fill_u64(d_saif_, (uint64_t)n_nets_ * kSaifWords, 0ull, stream_));
    HC(hipStreamSynchronize(stream_));
  }
  if (e0) HC(hipEventRecord(e0, stream_));
  std::vector<hipEvent_t> ev_eval;
  try {
    if (mode == 1) {
      HC(launch_saif_nets(d_saif_nets_, d_saif_job_, (uint32_t)saif_net_list_.size(), d_woff_,
                          d_wlen_, d_pool_, d_saif_, d_win_, stream_));
      run_two_pass(ev_eval);
    } else {
      run_chunks(ev_eval);
    }
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    for (hipEvent_t e : ev_eval) (void)hipEventDestroy(e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    throw;
  }
  float ms = 0;
  if (e0) {
    HC(hipEventRecord(e1, stream_));
    HC(hipEventSynchronize(e1));
    HC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  } else {
    HC(hipStreamSynchronize(stream_));
    ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  }
  double eval_ms = 0;
  for (size_t i = 0; i + 1 < ev_eval.size(); i += 2) {
    float x = 0;
    HC(hipEventElapsedTime(&x, ev_eval[i], ev_eval[i + 1]));
    eval_ms += x;
  }
  for (hipEvent_t e : ev_eval) (void)hipEventDestroy(e);
  stats_.eval_ms = eval_ms;
  stats_.eval_launches = ev_eval.size() / 2;
  uint64_t h[16];
  HC(hipMemcpyAsync(h, d_scal_, sizeof h, hipMemcpyDeviceToHost, stream_));
  HC(hipStreamSynchronize(stream_));
  finish_stats(h, ms);
  if (mode == 1 && keep_waves) {
    // two-pass: every waveform sits in the pool in absolute time
    std::vector<uint64_t> off(n_nets_);
    std::vector<uint32_t> len(n_nets_), pool(h[S_BUMP]);
    HC(hipMemcpyAsync(off.data(), d_woff_, n_nets_ * 8ull, hipMemcpyDeviceToHost, stream_));
    HC(hipMemcpyAsync(len.data(), d_wlen_, n_nets_ * 4ull, hipMemcpyDeviceToHost, stream_));
    if (!pool.empty())
      HC(hipMemcpyAsync(pool.data(), d_pool_, pool.size() * 4, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    waves_.assign(n_nets_, {});
    for (const GateRec& g : prog_->h_grec)
      for (uint32_t q = 0; q < len[g.out_net]; ++q) {
        const uint32_t w = pool[off[g.out_net] + q];
        waves_[g.out_net].push_back(((uint64_t)(w >> 2) << 2) | (w & 3u));
      }
  }
  stats_.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  stats_.device_allocs = allocs_;
  if (graph && mode == 0 && sig != graph_sig_ && !static_) {
    build_graph();
    graph_sig_ = sig;
  }
}

// ------------------------------------------------ allocation-free run ---

// The scratch of the allocation-free execution: one array per GPU, held
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (the array's name in host
// code is a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[Job::kStaticScratchBytes];

// Device address of this device's `scratch`, looked up once per device with
// that device current.
static unsigned char* static_scratch(int device) {
  static std::mutex mu;
  static std::map<int, unsigned char*> ptrs;
  std::lock_guard<std::mutex> lk(mu);
  auto it = ptrs.find(device);
  if (it != ptrs.end()) return it->second;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw GlsError(GLS_E_DEVICE, "no device address for the static scratch");
  ptrs[device] = (unsigned char*)p;
  return (unsigned char*)p;
}

// Every per-level and per-chunk buffer that run() would allocate on demand
// is carved from the static array: chunk windows, gate-output tails (a
// sixteenth of the array), the overflow stack buffer (2^22 words, or a
// smaller ovf_words), and tile descriptors / summaries / offsets in the
// rest.  The level scratch lives in the pool's free top (a level that does
// not fit there makes its chunk shorter).  None of these buffers is read
// before the run writes it, so nothing is cleared.  The job's own buffers
// are put back afterwards.
void Job::run_static() {
  HC(hipSetDevice(device_));
  if (mode != 0) throw GlsError(GLS_E_UNSUPPORTED, "the allocation-free execution runs the fused path");
  if (graph) throw GlsError(GLS_E_UNSUPPORTED, "the allocation-free execution does not capture graphs");
  if (keep_waves)
    throw GlsError(GLS_E_UNSUPPORTED, "the allocation-free execution does not keep waveforms");
  if (stack_mode) throw GlsError(GLS_E_UNSUPPORTED, "the allocation-free execution keeps stacks in LDS");
  unsigned char* base = static_scratch(device_);
  size_t off = 0;
  auto carve = [&](size_t bytes) -> void* {
    bytes = (bytes + 255) / 256 * 256;
    if (off + bytes > kStaticScratchBytes)
      throw GlsError(GLS_E_RANGE, "the batch needs more than the allocation-free execution's static array");
    void* p = base + off;
    off += bytes;
    return p;
  };
  struct Saved {
    void* p;
    size_t cap;
  };
  void** bufs[] = {(void**)&d_cwin_, (void**)&d_tails_, (void**)&d_ovf_, (void**)&d_tiles_,
                   (void**)&d_tout_, (void**)&d_toff_, (void**)&d_scratch_};
  size_t* caps[] = {&cap_cwin_, &cap_tails_, &cap_ovf_, &cap_tiles_, &cap_tout_, &cap_toff_,
                    &cap_scratch_};
  Saved saved[7];
  for (int i = 0; i < 7; ++i) saved[i] = {*bufs[i], *caps[i]};
  const uint64_t saved_ovf = ovf_words_;
  const size_t cw = net_base_.size() * sizeof(ChunkWin);
  d_cwin_ = (ChunkWin*)carve(cw);
  cap_cwin_ = cw;
  const size_t tb = kStaticScratchBytes / 16;
  d_tails_ = (uint32_t*)carve(tb);
  cap_tails_ = tb;
  ovf_words_ = std::min<uint64_t>(ovf_words_opt ? ovf_words_opt : (1ull << 22), 1ull << 22);
  d_ovf_ = (uint32_t*)carve(ovf_words_ * 4);
  cap_ovf_ = ovf_words_ * 4;
  const size_t per_tile = sizeof(TileDesc) + sizeof(TileOut) + 4;
  const size_t tiles = (kStaticScratchBytes - off - 3 * 256) / per_tile;
  d_tiles_ = (TileDesc*)carve(tiles * sizeof(TileDesc));
  cap_tiles_ = tiles * sizeof(TileDesc);
  d_tout_ = (TileOut*)carve(tiles * sizeof(TileOut));
  cap_tout_ = tiles * sizeof(TileOut);
  d_toff_ = (uint32_t*)carve(tiles * 4);
  cap_toff_ = tiles * 4;
  d_scratch_ = nullptr;  // the level scratch stays in the pool
  cap_scratch_ = 0;
  auto restore = [&] {
    for (int i = 0; i < 7; ++i) {
      *bufs[i] = saved[i].p;
      *caps[i] = saved[i].cap;
    }
    ovf_words_ = saved_ovf;
    static_ = false;
  };
  static_ = true;
  try {
    run();
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    restore();
    throw;
  }
  restore();
}

// Capture one device-driven execution of every chunk and level into a graph
// (the chunk sequence of the plain run that precedes it).
void Job::build_graph() {
  if (gexec_) {
    HC(hipGraphExecDestroy(gexec_));
    gexec_ = nullptr;
  }
  if (keep_waves && sched_.size() > 1)
    throw GlsError(GLS_E_UNSUPPORTED, "graph replay keeps waveforms of one-chunk runs only");
  tiles_cap_ = std::max<uint64_t>(1, max_tiles_);
  scratch_cap_ = std::max<uint64_t>(1, max_scratch_);
  ensure((void**)&d_tiles_, cap_tiles_, tiles_cap_ * sizeof(TileDesc));
  ensure((void**)&d_tout_, cap_tout_, tiles_cap_ * sizeof(TileOut));
  ensure((void**)&d_toff_, cap_toff_, tiles_cap_ * 4);
  ensure((void**)&d_tails_, cap_tails_, std::max<uint64_t>(max_tails_, 1) * 4);
  // the replay's bump sequence is the plain run's: a scratch at the pool top
  // stays clear of every level's outputs when it fits above the largest bump
  graph_scratch_in_pool_ = max_bump_ <= pool_words_ && scratch_cap_ <= pool_words_ - max_bump_;
  if (!graph_scratch_in_pool_) ensure((void**)&d_scratch_, cap_scratch_, scratch_cap_ * 4);
  // every chunk's windows in a slot of its own (the groups' chunks in turn)
  const size_t K = sched_.size();
  ensure((void**)&d_cwin_, cap_cwin_, std::max<size_t>(K, 1) * net_base_.size() * sizeof(ChunkWin));
  std::vector<std::vector<Chunk>> chunks(groups_.size());
  size_t slot = 0;
  for (size_t q = 0; q < groups_.size(); ++q) {
    uint64_t F = 0, prev_lb = 0;
    for (size_t k = 0; k < gsched_[q].size(); ++k) {
      Chunk c = make_chunk(k, F, gsched_[q][k], prev_lb);
      c.slot = slot++;
      put_chunk_windows(c, true);
      chunks[q].push_back(c);
      prev_lb = c.lb;
      F = gsched_[q][k];
    }
  }
  hipGraph_t g = nullptr;
  HC(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
  try {
    HC(launch_set_scal(d_scal_, 0, stream_));
    HC(launch_fill_u64(d_saif_, (uint64_t)n_nets_ * kSaifWords, 0ull, stream_));
    HC(launch_fill_u64(d_wtot_, n_nets_, 0ull, stream_));
    if (d_evjob_) HC(launch_fill_u64(d_evjob_, net_base_.size(), 0ull, stream_));
    HC(launch_saif_stim(d_store_, d_soff_, d_slen_, d_snet_, d_sjob_, n_sig_, d_saif_, d_win_,
                        stream_));
    std::vector<hipEvent_t> none;
    for (size_t q = 0; q < groups_.size(); ++q) {
      const Group& G = groups_[q];
      HC(launch_tails_init(prog_->d_grec + G.g0, G.g1 - G.g0, d_tail_len_, d_tail_off_, d_tails_,
                           stream_));
      for (const Chunk& c : chunks[q]) {
        double hint = 0;
        run_chunk(c, G, none, true, &hint);
        end_chunk(c, G, true);
      }
    }
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HC(hipStreamEndCapture(stream_, &g));
  hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HC(e);
}

void Job::finish_stats(const uint64_t* h, double ms) {
  if (h[S_ERR] & 16u)
    throw GlsError(GLS_E_OOM, "event pool exhausted (raise the pool size or shorten the window)");
  if (h[S_ERR] & 2u)
    throw GlsError(GLS_E_UNSUPPORTED, "a tile exceeded the overflow stack buffer (event density too skewed)");
  if (h[S_ERR] & 4u)
    throw GlsError(GLS_E_RANGE, "a gate's inputs carry 2^30 or more events");
  if (h[S_ERR] & 32u)
    throw GlsError(GLS_E_STATE, "a gate's input waveform was out of time order; results discarded");
  stats_.kernel_ms = ms;
  stats_.n_events = h[S_EVC];
  h_evjob_.assign(net_base_.size(), 0);
  if (d_evjob_) {
    HC(hipMemcpyAsync(h_evjob_.data(), d_evjob_, h_evjob_.size() * 8, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
  } else if (!h_evjob_.empty()) {
    h_evjob_[0] = stats_.n_events;
  }
  stats_.pool_words_used = mode == 1 ? h[S_BUMP] : max_bump_;
  stats_.tiles = h[S_TILES];
  stats_.tiles_hbm = h[S_TILES_HBM];
  stats_.pool_words = pool_words_;
  stats_.chunks = mode == 1 ? 1 : sched_.size();
  // algorithmic bytes: every input waveform read once per reading gate, every
  // output written once (whole-run waveform lengths), 3 u64 SAIF counters per
  // net, CSR metadata
  std::vector<uint64_t> W(n_nets_, 0);
  if (mode == 1) {
    std::vector<uint32_t> wlen(n_nets_);
    HC(hipMemcpyAsync(wlen.data(), d_wlen_, n_nets_ * 4ull, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    for (uint32_t n = 0; n < n_nets_; ++n) W[n] = wlen[n];
  } else {
    HC(hipMemcpyAsync(W.data(), d_wtot_, n_nets_ * 8ull, hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    for (uint32_t s = 0; s < n_sig_; ++s) W[h_snet_[s]] = h_slen_[s];
  }
  uint64_t in_w = 0, out_w = 0;
  for (const GateRec& g : prog_->h_grec) {
    for (int j = 0; j < g.k; ++j) in_w += W[g.in_net[j]];
    out_w += W[g.out_net];
  }
  if (prog_->h_grec.size() != n_gates_) in_w = out_w = 0;
  stats_.in_words = in_w;
  stats_.out_words = out_w;
  stats_.alg_bytes = 4 * (in_w + out_w) + 24ull * n_nets_ + 16ull * n_gates_ + 4ull * stats_.n_pins;
}

void Job::fetch_saif(std::vector<uint64_t>& v) const {
  v.resize((size_t)n_nets_ * kSaifWords);
  hip_check(hipMemcpyAsync(v.data(), d_saif_, v.size() * 8, hipMemcpyDeviceToHost, stream_),
            "saif D2H");
  hip_check(hipStreamSynchronize(stream_), "saif D2H");
}

}  // namespace gls
