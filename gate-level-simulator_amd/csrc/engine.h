// Device engine: uploads compiled programs and stimulus into HBM and runs
// the level kernels.  All device memory is owned here.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
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
  double eval_ms = 0;           // summed device time of the gate-evaluation launches
  uint64_t eval_launches = 0;
  uint64_t level_blocks = 0;    // workgroups per fused level launch
  uint64_t pool_words_used = 0;
  uint64_t pool_words = 0;      // event pool capacity after the run
  uint64_t tiles = 0, tiles_hbm = 0;  // fused tiles walked, and of them on the HBM path
  uint64_t chunks = 0;          // time chunks of the run (fused path)
  uint64_t chunk_retries = 0;   // chunks run again shorter (pool too small)
  uint64_t device_allocs = 0;   // device allocations made by the last execution
};

// Host and device image of one compiled program for one delay time scale:
// gate records (local net ids), delay tables, truth tables.  Shared by every
// job that simulates the program, so a resident netlist is uploaded once.
struct DevProgram {
  int device = -1;
  uint32_t n_gates = 0, n_nets = 0, n_pins = 0, timescale_ps = 1;
  std::vector<uint32_t> level_start;
  std::vector<GateRec> h_grec;
  std::vector<GateDev> h_gdev;
  std::vector<uint16_t> h_dtab;
  std::vector<uint8_t> h_luts;   // 256-byte blocks (a fan-in 5 / 6 table spans 4 / 16)
  std::vector<uint32_t> h_lutp;  // 16 packed words per block
  // wide gates (fan-in 5..6): their indices in h_grec, level by level
  // (level l: h_wide[wide_start[l] .. wide_start[l+1]))
  std::vector<uint32_t> h_wide;
  std::vector<uint32_t> wide_start;
  uint32_t max_k = 0;            // largest fan-in
  // device tables (valid once is_resident(); one allocation, d_block)
  GateRec* d_grec = nullptr;
  GateDev* d_gdev = nullptr;
  uint16_t* d_dtab = nullptr;
  uint8_t* d_luts = nullptr;
  uint32_t* d_lutp = nullptr;
  uint32_t* d_wide = nullptr;
  std::mutex upload_mu;
  DevProgram() = default;
  DevProgram(const DevProgram&) = delete;
  DevProgram& operator=(const DevProgram&) = delete;
  ~DevProgram();
  size_t device_bytes() const;
  bool is_resident() const;
  // No-op when already resident.  All or nothing: on failure nothing is
  // published and a later call uploads again.
  void upload(int device, hipStream_t st);

 private:
  void* d_block = nullptr;
  std::atomic<bool> resident{false};
};

std::shared_ptr<DevProgram> build_program_image(const Program& pg, uint32_t timescale_ps);

// One simulation instance: a program, its stimulus (signal i drives net
// stim_nets[i]) and its SAIF window.
struct SimInstance {
  const Program* pg = nullptr;
  std::shared_ptr<DevProgram> image;  // optional: resident image of pg
  const Stimulus* st = nullptr;
  std::vector<int> stim_nets;
  uint64_t start = 0, end = 0;
};

// Resident program images of one device, by (program, time scale).
class ProgramCache {
 public:
  // The image of pg (built and cached on first use; the cache keeps pg alive).
  std::shared_ptr<DevProgram> get(const std::shared_ptr<const Program>& pg, uint32_t timescale_ps);
  bool resident(const Program* pg, uint32_t timescale_ps) const;
  size_t device_bytes() const;
  size_t size() const;
  void clear();

 private:
  mutable std::mutex mu_;
  std::map<std::pair<const Program*, uint32_t>,
           std::pair<std::shared_ptr<const Program>, std::shared_ptr<DevProgram>>>
      map_;
};

// Caching device allocator shared by the jobs of a serving loop: a job's
// buffers go back to the arena when it ends and later jobs take them again,
// so a stream of jobs neither allocates nor frees device memory once the
// arena is warm (hipFree waits for the whole device, which would stall the
// other running jobs).  A cached block of at least the asked size and at
// most 1.5 times it is reused; when the device is out of memory the cached
// blocks are freed and the allocation is tried again.  The caller returns a
// block only when no device work that uses it is pending.
class DevArena {
 public:
  explicit DevArena(uint64_t max_cached_bytes) : max_cached_(max_cached_bytes) {}
  ~DevArena();
  DevArena(const DevArena&) = delete;
  DevArena& operator=(const DevArena&) = delete;
  // A block of at least `bytes` bytes, or false when the device has no room.
  bool get(size_t bytes, void** p);
  // For a buffer that uses whatever size it gets (an automatically sized
  // event pool): the largest cached block when it holds at least min_bytes,
  // else a new block of `want` bytes, or of less (down to min_bytes) when the
  // device has no room; *got is the block's size.
  bool get_large(size_t want, size_t min_bytes, void** p, size_t* got);
  void put(void* p);
  void trim();  // frees every cached block
  // frees every cached block when the device has less than 4 GiB (or 1.5%
  // of its memory) free: the HIP runtime needs room of its own
  void keep_headroom();
  uint64_t cached_bytes();
  uint64_t hits() const { return hits_; }
  uint64_t misses() const { return misses_; }

 private:
  static size_t size_class(size_t bytes);
  void trim_locked();
  void keep_headroom_locked();
  std::mutex mu_;
  std::multimap<size_t, void*> free_;          // cached blocks by size
  std::map<void*, size_t> live_;               // blocks handed out, with their sizes
  uint64_t cached_ = 0, max_cached_;
  uint64_t hits_ = 0, misses_ = 0;
};

// A prepared simulation of one or more instances, run level by level: level
// l of every instance is evaluated by the same launches.
//
// The fused path runs the window as a sequence of time chunks (device.h,
// ChunkWin): the stimulus stays resident; each chunk cuts its part of the
// stimulus into the event pool, runs every level, adds its SAIF share and
// keeps each gate output's tail for the next chunk.  One chunk covers the
// whole window when it fits the pool; otherwise the chunk length adapts to
// the pool (a chunk that would overflow it is run again shorter), and times
// of any size up to 2^32 ps are simulated by moving each chunk to its own
// time base.
class Job {
 public:
  Job(int device, hipStream_t stream, DevArena* arena = nullptr);
  ~Job();
  Job(const Job&) = delete;
  Job& operator=(const Job&) = delete;

  void prepare(const std::vector<SimInstance>& inst, uint64_t pool_words);
  // Run all levels.  Throws GlsError.
  void run();
  // run() with no device allocation: the per-level scratch is carved from
  // one static device array of kStaticScratchBytes per GPU; the pool cannot
  // grow (time chunks shrink to fit it instead).  Fused path, no graph, no
  // kept waveforms, stacks in LDS.  Calls that overlap on one GPU share the
  // array and race.
  void run_static();
  static constexpr size_t kStaticScratchBytes = 256ull << 20;
  // Results (union net space; instance i's nets start at net_base(i))
  void fetch_saif(std::vector<uint64_t>& t0t1tx) const;
  // Waveforms of the gate outputs (absolute (time << 2 | value) words; other
  // nets empty); needs keep_waves before run().
  const std::vector<std::vector<uint64_t>>& waves() const { return waves_; }
  const JobStats& stats() const { return stats_; }
  size_t n_instances() const { return net_base_.size(); }
  uint32_t net_base(size_t i) const { return net_base_[i]; }
  uint64_t device_bytes() const { return dev_bytes_; }
  // gate events of instance i in the last run
  uint64_t instance_events(size_t i) const { return i < h_evjob_.size() ? h_evjob_[i] : 0; }
  // chunk boundaries of the last run (F_1 .. F_K; the last is kChunkEnd)
  const std::vector<uint64_t>& chunks() const { return sched_; }

  static constexpr uint64_t kChunkEnd = ~0ull;
  static constexpr uint64_t kNoChunks = ~0ull;  // chunk_ps: one chunk, never split

  uint32_t seg_target = 0;        // input events per time segment (0: mode default)
  int mode = 0;                   // 0: fused LDS-tiled level kernel, 1: two-pass kernels
  uint32_t level_blocks_opt = 0;  // workgroups per fused launch (0: occupancy limit)
  uint32_t prefetch = 1;          // stage the next tile before storing the current one
  uint32_t saif_tc = 0;           // count 0<->1 transitions per net (SAIF TC)
  uint64_t ovf_words_opt = 0;     // overflow stack words (0: from 2^24, doubled as needed)
  // 1: after a plain execution, capture the whole level sequence (no host
  // round trips) into a hipGraph and replay it on later executions
  uint32_t graph = 0;
  uint32_t stack_mode = 0;  // 0: preemption stacks in LDS, 1: per-workgroup scratch in HBM
  uint32_t pool_grow = 0;   // 1: given pool and overflow-buffer sizes are starting sizes that grow
  uint32_t keep_waves = 0;  // 1: collect every gate output's waveform on the host
  // time-chunk length in time units (fused path): 0 automatic, kNoChunks one
  // chunk (a pool too small fails with GLS_E_OOM), else a fixed length
  uint64_t chunk_ps = 0;
  // Asked before each level of a plain execution; true stops the run with
  // GLS_E_CANCELLED after the stream has drained (job cancel or deadline).
  std::function<bool()> should_stop;
  bool graph_replayed() const { return last_replay_; }
  static constexpr uint32_t kDefaultSeg = 32;  // input events per time segment

 private:
  struct Chunk {
    uint64_t F, Fn;       // absolute bounds [F, Fn) (Fn = kChunkEnd: to the end)
    uint64_t lb;          // look-back start (abs); relative time = abs - lb + 1
    uint32_t lo_rel, hi_rel, shift, first;
    size_t slot;          // ChunkWin slot of the chunk's windows
  };
  enum ChunkResult { CHUNK_OK = 0, CHUNK_SHORTER = 1 };
  // An instance group: instances [i0, i1), their gates [g0, g1) in levels
  // [L0, L1) of the level table, their stimulus signals [s0, s1), and the
  // time after which none of their output entries lies.
  struct Group {
    uint32_t i0, i1, L0, L1, g0, g1, s0, s1;
    uint64_t horizon;
    double est_words;  // estimated waveform words of the group's whole window
  };
  // one stimulus signal of the resident store: n (time, value) entries
  struct SigSrc {
    const uint32_t* t;
    const uint8_t* v;
    size_t n;
  };
  // words per pinned staging buffer of the stimulus upload (two of them)
  static constexpr uint64_t kStageWords = 1ull << 24;
  void upload_store(const std::vector<SigSrc>& src, const std::vector<uint64_t>& soff,
                    uint64_t total);
  void ensure(void** p, size_t& cap, size_t need);
  bool try_ensure(void** p, size_t& cap, size_t need);
  Chunk make_chunk(size_t k, uint64_t F, uint64_t Fn, uint64_t prev_lb) const;
  void put_chunk_windows(const Chunk& c, bool capture);
  int run_chunk(const Chunk& c, const Group& G, std::vector<hipEvent_t>& ev_eval,
                bool device_driven, double* grow_hint);
  void end_chunk(const Chunk& c, const Group& G, bool device_driven);
  void run_chunks(std::vector<hipEvent_t>& ev_eval);
  void run_group_chunks(const Group& G, std::vector<hipEvent_t>& ev_eval);
  void build_graph();
  void finish_stats(const uint64_t* scal, double ms);
  void run_two_pass(std::vector<hipEvent_t>& ev_eval);
  void stop_point();
  void* dmalloc(size_t bytes);
  bool dev_alloc(void** p, size_t bytes);  // arena or hipMalloc; false: no room
  void dev_free(void* p);                  // arena (after the stream drains) or hipFree
  bool grow_pool(uint64_t need, uint64_t used);
  bool refresh_pool(uint64_t words);
  uint64_t pool_max_words() const;
  void fetch_chunk_waves(const Chunk& c, const Group& G);
  int device_;
  hipStream_t stream_;
  DevArena* arena_ = nullptr;
  std::shared_ptr<DevProgram> prog_;
  std::vector<uint32_t> net_base_;
  std::vector<uint64_t> win_;  // start, end per instance
  uint32_t n_nets_ = 0, n_gates_ = 0;
  std::vector<uint32_t> level_start_;
  std::vector<uint32_t> saif_net_list_, saif_net_job_;
  uint64_t pool_words_ = 0, stim_words_ = 0, dev_bytes_ = 0;
  bool pool_auto_ = false;  // pool sized by the job (grows between levels)
  // device buffers of the job
  uint64_t* d_woff_ = nullptr;
  uint32_t* d_wlen_ = nullptr;
  uint32_t* d_pool_ = nullptr;
  unsigned long long* d_saif_ = nullptr;
  uint32_t* d_saif_nets_ = nullptr;
  uint32_t* d_saif_job_ = nullptr;
  unsigned long long* d_win_ = nullptr;
  uint64_t* d_scal_ = nullptr;
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
  size_t cap_tiles_ = 0, cap_tout_ = 0, cap_toff_ = 0;
  GatePlan* d_plan_ = nullptr;
  TileDesc* d_tiles_ = nullptr;
  TileOut* d_tout_ = nullptr;
  uint32_t* d_toff_ = nullptr;
  uint32_t* d_ovf_ = nullptr;
  uint64_t ovf_words_ = 0;
  size_t cap_ovf_ = 0;
  int level_blocks_ = 0;
  uint32_t cap_ = 0;  // LDS input capacity per tile (words)
  std::vector<void*> owned_;
  // two-pass path: initial waveform tables
  std::vector<uint64_t> h_woff_init_;
  std::vector<uint32_t> h_wlen_init_;
  // fused path: resident stimulus (signals and constants), 64-bit words
  uint32_t n_sig_ = 0;
  uint64_t* d_store_ = nullptr;
  uint64_t* d_soff_ = nullptr;
  uint32_t *d_slen_ = nullptr, *d_snet_ = nullptr, *d_sjob_ = nullptr;
  uint64_t *d_scnt_ = nullptr, *d_sbase_ = nullptr, *d_spart_ = nullptr;
  uint32_t* d_sfirst_ = nullptr;
  std::vector<uint32_t> h_slen_, h_snet_;
  uint64_t t_last_ = 0;   // last stimulus time
  uint32_t wmax_ = 1;     // largest gate delay (the look-back length)
  uint64_t horizon_ = 0;  // no output time reaches it
  // tails of the gate outputs
  uint32_t* d_tail_len_ = nullptr;
  uint64_t* d_tail_off_ = nullptr;
  uint32_t* d_tails_ = nullptr;
  size_t cap_tails_ = 0;
  uint64_t *d_tlen_g_ = nullptr, *d_tsrc_g_ = nullptr, *d_toff_g_ = nullptr, *d_gpart_ = nullptr;
  unsigned long long* d_saif_chunk_ = nullptr;
  // batches: gate events per instance (run total and current chunk)
  unsigned long long *d_evjob_ = nullptr, *d_evjob_chunk_ = nullptr;
  std::vector<uint64_t> h_evjob_;
  unsigned long long* d_wtot_ = nullptr;
  ChunkWin* d_cwin_ = nullptr;
  size_t cap_cwin_ = 0;
  std::vector<uint64_t> sched_;  // F_1 .. F_K of the last run, group after group
  std::vector<Group> groups_;
  std::vector<std::vector<uint64_t>> gsched_;  // per group: its F_1 .. F_K
  size_t gsched_count_ = 0;                     // chunks of the group being run
  uint64_t max_tiles_ = 0, tiles_cap_ = 0, max_scratch_ = 0, scratch_cap_ = 0;
  uint64_t max_tails_ = 0, max_bump_ = 0;
  uint64_t pool_reallocs_ = 0;  // pool allocations made while running
  uint64_t allocs_ = 0;         // device allocations made since the run started
  bool static_ = false;         // run_static(): no allocation, scratch from the static array
  bool graph_scratch_in_pool_ = false;  // captured levels keep their scratch at the pool top
  uint32_t* d_stk_ = nullptr;
  size_t cap_stk_ = 0;
  hipGraphExec_t gexec_ = nullptr;
  std::vector<uint64_t> graph_sig_;
  bool last_replay_ = false;
  std::vector<std::vector<uint64_t>> waves_;
  JobStats stats_;
};

void hip_check(hipError_t e, const char* what);

}  // namespace gls
