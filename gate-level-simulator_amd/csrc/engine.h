// Device engine: uploads a compiled program + stimulus into HBM and runs the
// level kernels.  All device memory is owned here.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "csr.h"
#include "device.h"
#include "netlist.h"

namespace gls {

struct JobStats {
  uint64_t n_events = 0;        // gate events (distinct input-event times per output)
  uint64_t in_words = 0;        // sum over gates of input waveform lengths
  uint64_t out_words = 0;       // sum over gates of output waveform lengths
  uint64_t n_nets = 0, n_gates = 0, n_pins = 0;
  uint64_t alg_bytes = 0;       // algorithmic bytes of the level kernels
  double kernel_ms = 0;         // device time of the level loop (hipEvent)
  double wall_ms = 0;           // host wall time of the level loop
  double eval_ms = 0;           // summed device time of the k_eval launches
  uint64_t eval_launches = 0;
  uint64_t level_blocks = 0;    // workgroups per fused level launch
  uint64_t pool_words_used = 0;
};

// A prepared simulation: program, delays, stimulus resident on the device.
class Job {
 public:
  Job(int device, hipStream_t stream);
  ~Job();
  Job(const Job&) = delete;
  Job& operator=(const Job&) = delete;

  // Compile-time inputs; stim_nets[i] is the net id of stimulus signal i.
  void prepare(const Program& pg, const Stimulus& st, const std::vector<int>& stim_nets,
               uint64_t start, uint64_t end, uint64_t pool_words);
  // Run all levels (plain per-level launches). Returns GLS status.
  void run();
  // Results
  void fetch_saif(std::vector<uint64_t>& t0t1tx) const;
  void fetch_waves(std::vector<uint64_t>& off, std::vector<uint32_t>& len,
                   std::vector<uint32_t>& pool_words) const;
  const JobStats& stats() const { return stats_; }
  uint32_t seg_target = 0;  // input events per time segment (0: mode default)
  int mode = 0;             // 0: fused LDS-tiled level kernel, 1: two-pass kernels
  uint32_t level_blocks_opt = 0;  // workgroups per fused launch (0: occupancy limit)
  uint32_t tiles_per_ticket = 0;  // fused tiles handed out per ticket (0: one)
  uint32_t prefetch = 1;          // stage the next tile before storing the current one

 private:
  void ensure(void** p, size_t& cap, size_t need);
  void run_fused(std::vector<hipEvent_t>& ev_eval);
  void run_two_pass(std::vector<hipEvent_t>& ev_eval);
  int device_;
  hipStream_t stream_;
  uint64_t start_ = 0, end_ = 0;
  uint32_t n_nets_ = 0, n_gates_ = 0;
  std::vector<uint32_t> level_start_;
  std::vector<uint32_t> saif_net_list_;
  uint64_t pool_words_ = 0, stim_words_ = 0;
  // device buffers
  GateDev* d_gates_ = nullptr;
  GateRec* d_grec_ = nullptr;
  uint8_t* d_luts_ = nullptr;
  uint64_t* d_woff_ = nullptr;
  uint32_t* d_wlen_ = nullptr;
  uint32_t* d_pool_ = nullptr;
  unsigned long long* d_saif_ = nullptr;
  uint32_t* d_saif_nets_ = nullptr;
  uint64_t* d_scal_ = nullptr;  // [0]=bump [1]=total_segs [2]=total_scr [3]=total_out [4]=ev_count [5]=overflow
  // per-level temporaries
  uint64_t *d_nseg_ = nullptr, *d_scr_ = nullptr, *d_segbase_ = nullptr, *d_scrbase_ = nullptr;
  uint64_t *d_outlen_ = nullptr, *d_outoff_ = nullptr, *d_part_ = nullptr;
  uint32_t* d_gnseg_ = nullptr;
  size_t cap_segmap_ = 0, cap_sums_ = 0, cap_carry_ = 0, cap_scratch_ = 0;
  uint32_t* d_segmap_ = nullptr;
  SegSum* d_sums_ = nullptr;
  SegCarry* d_carry_ = nullptr;
  uint32_t* d_scratch_ = nullptr;
  // fused path
  uint64_t *d_ntile_ = nullptr, *d_bound_ = nullptr, *d_tilebase_ = nullptr;
  size_t cap_tiles_ = 0, cap_tstate_ = 0;
  GatePlan* d_plan_ = nullptr;
  TileDesc* d_tiles_ = nullptr;
  uint32_t* d_lutp_ = nullptr;
  uint16_t* d_dtab_ = nullptr;
  unsigned long long* d_tstate_ = nullptr;
  uint32_t* d_ovf_ = nullptr;
  uint64_t ovf_words_ = 0;
  int level_blocks_ = 0;
  uint32_t cap_ = 0;                          // LDS input words per fused tile
  static constexpr uint32_t kDefaultSeg = 32;  // input events per time segment
  std::vector<uint64_t> h_woff_init_;
  std::vector<uint32_t> h_wlen_init_;
  JobStats stats_;
  std::vector<std::vector<uint32_t>> gate_in_nets_;  // for stats
};

void hip_check(hipError_t e, const char* what);

}  // namespace gls
