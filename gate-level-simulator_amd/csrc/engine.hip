#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cstring>

#include "../../include/gls_status.h"

namespace gls {

// launch helpers (kernels.hip)
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
hipError_t launch_saif_nets(const uint32_t* nets, uint32_t count, const uint64_t* woff,
                            const uint32_t* wlen, const uint32_t* pool, unsigned long long* saif,
                            uint64_t start, uint64_t end, hipStream_t st);

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GLS_E_OOM : GLS_E_DEVICE;
  throw GlsError(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HC(x) hip_check((x), #x)

Job::Job(int device, hipStream_t stream) : device_(device), stream_(stream) {}

Job::~Job() {
  void* ptrs[] = {d_gates_, d_luts_, d_woff_, d_wlen_, d_pool_, d_saif_, d_saif_nets_, d_scal_,
                  d_nseg_, d_scr_, d_segbase_, d_scrbase_, d_outlen_, d_outoff_, d_part_, d_gnseg_,
                  d_segmap_, d_sums_, d_carry_, d_scratch_};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

void Job::ensure(void** p, size_t& cap, size_t need) {
  if (need <= cap && *p) return;
  size_t n = std::max<size_t>(need, cap + cap / 2);
  n = std::max<size_t>(n, 4096);
  if (*p) HC(hipFree(*p));
  *p = nullptr;
  HC(hipMalloc(p, n));
  cap = n;
}

void Job::prepare(const Program& pg, const Stimulus& st, const std::vector<int>& stim_nets,
                  uint64_t start, uint64_t end, uint64_t pool_words) {
  HC(hipSetDevice(device_));
  if (end <= start) throw GlsError(GLS_E_ARG, "dumpoff time must be above dumpon time");
  start_ = start;
  end_ = end;
  n_gates_ = (uint32_t)pg.gates.size();
  level_start_ = pg.level_start;
  n_nets_ = (uint32_t)pg.net_names.size();
  for (int sn : stim_nets) n_nets_ = std::max<uint32_t>(n_nets_, (uint32_t)sn + 1);

  // gates
  std::vector<GateDev> gd(n_gates_);
  gate_in_nets_.assign(n_gates_, {});
  for (uint32_t i = 0; i < n_gates_; ++i) {
    const CompiledGate& g = pg.gates[i];
    GateDelays D = gate_delays(g, st.timescale_ps);
    GateDev& r = gd[i];
    std::memset(&r, 0, sizeof r);
    for (int j = 0; j < 4; ++j) {
      r.in_net[j] = j < g.k ? (uint32_t)g.in_net[j] : 0u;
      r.nrec[j] = D.nrec[j];
      for (int kd = 0; kd < 3; ++kd) {
        r.dly[j][kd][0] = D.v[j][kd][0];
        r.dly[j][kd][1] = D.v[j][kd][1];
      }
      if (j < g.k) gate_in_nets_[i].push_back((uint32_t)g.in_net[j]);
    }
    r.out_net = (uint32_t)g.out_net;
    r.lut = (uint16_t)g.lut;
    r.k = (uint8_t)g.k;
    r.dmax = D.dmax;
    if (g.k < 1) throw GlsError(GLS_E_UNSUPPORTED, "gate without inputs");
  }
  std::vector<uint8_t> luts(pg.luts.size() * 256, 1);
  for (size_t i = 0; i < pg.luts.size(); ++i)
    std::memcpy(&luts[i * 256], pg.luts[i].data(), pg.luts[i].size());

  // initial waveforms: stimulus signals, then constants
  std::vector<uint32_t> words;
  h_woff_init_.assign(n_nets_, 0);
  h_wlen_init_.assign(n_nets_, 0);
  saif_net_list_.clear();
  for (size_t i = 0; i < st.waves.size(); ++i) {
    int net = stim_nets[i];
    const Wave& w = st.waves[i];
    h_woff_init_[net] = words.size();
    h_wlen_init_[net] = (uint32_t)w.size();
    for (size_t q = 0; q < w.size(); ++q) {
      if (w.t[q] >= (1u << 30) - (1u << 17))
        throw GlsError(GLS_E_RANGE, "stimulus time beyond the 30-bit event word");
      words.push_back((w.t[q] << 2) | (w.v[q] & 3u));
    }
    saif_net_list_.push_back((uint32_t)net);
  }
  for (uint32_t n = 0; n < (uint32_t)pg.net_names.size(); ++n) {
    if (pg.role[n] != NR_CONST) continue;
    h_woff_init_[n] = words.size();
    h_wlen_init_[n] = 1;
    words.push_back((0u << 2) | (uint32_t)(pg.const_val[n] & 3));
    saif_net_list_.push_back(n);
  }
  stim_words_ = words.size();

  size_t free_b = 0, total_b = 0;
  HC(hipMemGetInfo(&free_b, &total_b));
  if (pool_words == 0) pool_words = (uint64_t)(free_b * 0.70) / 4;
  if (pool_words < stim_words_ + 1024) pool_words = stim_words_ + 1024;
  pool_words_ = pool_words;

  HC(hipMalloc((void**)&d_gates_, std::max<size_t>(1, gd.size()) * sizeof(GateDev)));
  HC(hipMalloc((void**)&d_luts_, std::max<size_t>(256, luts.size())));
  HC(hipMalloc((void**)&d_woff_, n_nets_ * sizeof(uint64_t)));
  HC(hipMalloc((void**)&d_wlen_, n_nets_ * sizeof(uint32_t)));
  HC(hipMalloc((void**)&d_pool_, pool_words_ * sizeof(uint32_t)));
  HC(hipMalloc((void**)&d_saif_, (size_t)n_nets_ * 3 * sizeof(unsigned long long)));
  HC(hipMalloc((void**)&d_saif_nets_, std::max<size_t>(1, saif_net_list_.size()) * 4));
  HC(hipMalloc((void**)&d_scal_, 16 * sizeof(uint64_t)));
  uint32_t maxw = 1;
  for (size_t l = 0; l + 1 < level_start_.size(); ++l)
    maxw = std::max(maxw, level_start_[l + 1] - level_start_[l]);
  HC(hipMalloc((void**)&d_nseg_, maxw * 8ull));
  HC(hipMalloc((void**)&d_scr_, maxw * 8ull));
  HC(hipMalloc((void**)&d_segbase_, maxw * 8ull));
  HC(hipMalloc((void**)&d_scrbase_, maxw * 8ull));
  HC(hipMalloc((void**)&d_outlen_, maxw * 8ull));
  HC(hipMalloc((void**)&d_outoff_, maxw * 8ull));
  HC(hipMalloc((void**)&d_part_, scan_part_size(maxw) * 8ull));
  HC(hipMalloc((void**)&d_gnseg_, maxw * 4ull));

  if (!gd.empty())
    HC(hipMemcpyAsync(d_gates_, gd.data(), gd.size() * sizeof(GateDev), hipMemcpyHostToDevice, stream_));
  if (!luts.empty())
    HC(hipMemcpyAsync(d_luts_, luts.data(), luts.size(), hipMemcpyHostToDevice, stream_));
  if (!words.empty())
    HC(hipMemcpyAsync(d_pool_, words.data(), words.size() * 4, hipMemcpyHostToDevice, stream_));
  if (!saif_net_list_.empty())
    HC(hipMemcpyAsync(d_saif_nets_, saif_net_list_.data(), saif_net_list_.size() * 4,
                      hipMemcpyHostToDevice, stream_));
  HC(hipStreamSynchronize(stream_));
  stats_ = JobStats();
  stats_.n_nets = n_nets_;
  stats_.n_gates = n_gates_;
  stats_.n_pins = pg.n_pins;
}

void Job::run() {
  HC(hipSetDevice(device_));
  auto w0 = std::chrono::steady_clock::now();
  hipEvent_t e0, e1;
  HC(hipEventCreate(&e0));
  HC(hipEventCreate(&e1));
  uint64_t scal[16] = {0};
  scal[0] = stim_words_;
  HC(hipMemcpyAsync(d_woff_, h_woff_init_.data(), n_nets_ * 8ull, hipMemcpyHostToDevice, stream_));
  HC(hipMemcpyAsync(d_wlen_, h_wlen_init_.data(), n_nets_ * 4ull, hipMemcpyHostToDevice, stream_));
  HC(hipMemcpyAsync(d_scal_, scal, sizeof scal, hipMemcpyHostToDevice, stream_));
  HC(hipMemsetAsync(d_saif_, 0, (size_t)n_nets_ * 3 * 8, stream_));
  HC(hipStreamSynchronize(stream_));

  HC(hipEventRecord(e0, stream_));
  HC(launch_saif_nets(d_saif_nets_, (uint32_t)saif_net_list_.size(), d_woff_, d_wlen_, d_pool_,
                      d_saif_, start_, end_, stream_));
  uint64_t* bump = d_scal_ + 0;
  uint64_t* tot_segs = d_scal_ + 1;
  uint64_t* tot_scr = d_scal_ + 2;
  uint64_t* tot_out = d_scal_ + 3;
  unsigned long long* evc = (unsigned long long*)(d_scal_ + 4);
  // HIP events around every k_eval launch, on the stream it runs on
  std::vector<hipEvent_t> ev_eval;
  for (size_t l = 0; l + 1 < level_start_.size(); ++l) {
    uint32_t g0 = level_start_[l], ng = level_start_[l + 1] - g0;
    if (ng == 0) continue;
    const GateDev* gates = d_gates_ + g0;
    HC(launch_plan(gates, ng, d_wlen_, seg_target, d_nseg_, d_scr_, d_gnseg_, stream_));
    HC(launch_scan(d_nseg_, d_segbase_, ng, d_part_, tot_segs, nullptr, stream_));
    HC(launch_scan(d_scr_, d_scrbase_, ng, d_part_, tot_scr, nullptr, stream_));
    uint64_t h[4];
    HC(hipMemcpyAsync(h, d_scal_, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
    HC(hipStreamSynchronize(stream_));
    uint64_t nsegs = h[1], nscr = h[2], used = h[0];
    ensure((void**)&d_segmap_, cap_segmap_, nsegs * 4);
    ensure((void**)&d_sums_, cap_sums_, nsegs * sizeof(SegSum));
    ensure((void**)&d_carry_, cap_carry_, nsegs * sizeof(SegCarry));
    ensure((void**)&d_scratch_, cap_scratch_, nscr * 4);
    // the level's output is at most its scratch size; refuse before writing
    if (used + nscr > pool_words_)
      throw GlsError(GLS_E_OOM, "event pool exhausted at level " + std::to_string(l) +
                                    " (raise the pool size or shorten the window)");
    HC(launch_segmap(ng, d_segbase_, d_gnseg_, d_segmap_, stream_));
    hipEvent_t a, b;
    HC(hipEventCreate(&a));
    HC(hipEventCreate(&b));
    ev_eval.push_back(a);
    ev_eval.push_back(b);
    HC(hipEventRecord(a, stream_));
    HC(launch_eval(gates, d_segbase_, d_scrbase_, d_gnseg_, d_segmap_, nsegs, d_woff_, d_wlen_,
                   d_pool_, d_luts_, d_scratch_, d_sums_, evc, stream_));
    HC(hipEventRecord(b, stream_));
    HC(launch_gate_scan(gates, ng, d_segbase_, d_gnseg_, d_sums_, d_carry_, d_outlen_, d_saif_,
                        start_, end_, stream_));
    HC(launch_scan(d_outlen_, d_outoff_, ng, d_part_, tot_out, bump, stream_));
    HC(launch_emit(gates, d_segbase_, d_segmap_, nsegs, d_sums_, d_carry_, d_outoff_, d_outlen_,
                   d_scratch_, d_pool_, d_woff_, d_wlen_, d_saif_, start_, end_, stream_));
    HC(launch_bump(bump, tot_out, stream_));
  }
  HC(hipEventRecord(e1, stream_));
  HC(hipEventSynchronize(e1));
  float ms = 0;
  HC(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  double eval_ms = 0;
  for (size_t i = 0; i + 1 < ev_eval.size(); i += 2) {
    float x = 0;
    HC(hipEventElapsedTime(&x, ev_eval[i], ev_eval[i + 1]));
    eval_ms += x;
  }
  for (hipEvent_t e : ev_eval) (void)hipEventDestroy(e);
  stats_.eval_ms = eval_ms;
  stats_.eval_launches = ev_eval.size() / 2;
  uint64_t h[8];
  HC(hipMemcpy(h, d_scal_, sizeof h, hipMemcpyDeviceToHost));
  auto w1 = std::chrono::steady_clock::now();
  stats_.kernel_ms = ms;
  stats_.wall_ms = std::chrono::duration<double, std::milli>(w1 - w0).count();
  stats_.n_events = h[4];
  stats_.pool_words_used = h[0];
  // algorithmic bytes: every input waveform read once per reading gate, every
  // output written once, 3 u64 SAIF counters per net, CSR metadata
  std::vector<uint32_t> wlen(n_nets_);
  HC(hipMemcpy(wlen.data(), d_wlen_, n_nets_ * 4ull, hipMemcpyDeviceToHost));
  uint64_t in_w = 0, out_w = 0;
  std::vector<uint8_t> is_out(n_nets_, 0);
  for (uint32_t g = 0; g < n_gates_; ++g)
    for (uint32_t n : gate_in_nets_[g]) in_w += wlen[n];
  // output nets: every gate's out_net
  std::vector<GateDev> gd(n_gates_);
  if (n_gates_) HC(hipMemcpy(gd.data(), d_gates_, n_gates_ * sizeof(GateDev), hipMemcpyDeviceToHost));
  for (auto& g : gd) out_w += wlen[g.out_net];
  stats_.in_words = in_w;
  stats_.out_words = out_w;
  stats_.alg_bytes = 4 * (in_w + out_w) + 24ull * n_nets_ + 16ull * n_gates_ + 4ull * stats_.n_pins;
}

void Job::fetch_saif(std::vector<uint64_t>& v) const {
  v.resize((size_t)n_nets_ * 3);
  hip_check(hipMemcpy(v.data(), d_saif_, v.size() * 8, hipMemcpyDeviceToHost), "saif D2H");
}

void Job::fetch_waves(std::vector<uint64_t>& off, std::vector<uint32_t>& len,
                      std::vector<uint32_t>& pool) const {
  off.resize(n_nets_);
  len.resize(n_nets_);
  hip_check(hipMemcpy(off.data(), d_woff_, n_nets_ * 8ull, hipMemcpyDeviceToHost), "woff D2H");
  hip_check(hipMemcpy(len.data(), d_wlen_, n_nets_ * 4ull, hipMemcpyDeviceToHost), "wlen D2H");
  uint64_t used = 0;
  hip_check(hipMemcpy(&used, d_scal_, 8, hipMemcpyDeviceToHost), "bump D2H");
  pool.resize(used);
  if (used) hip_check(hipMemcpy(pool.data(), d_pool_, used * 4, hipMemcpyDeviceToHost), "pool D2H");
}

}  // namespace gls
