// Serving loop: a request queue in front of one device.  Jobs (netlist +
// stimulus + SAIF window) are admitted by their device-memory footprint as
// memory and slots free up, each admitted job runs on its own HIP stream so
// the kernels of concurrent jobs overlap on the device, and compiled
// netlists stay resident for later jobs.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace gls {

// One simulation as the C-ABI sees it: program, names for the SAIF writer,
// stimulus, window.
struct RunInst {
  std::shared_ptr<const Program> pg;
  std::vector<int> stim_nets;
  std::vector<std::string> names;  // all net names incl. stimulus-only signals
  std::vector<uint8_t> in_saif;
  std::vector<uint8_t> is_gate;
  std::unordered_map<std::string, int> by_name;
  const Stimulus* st = nullptr;
  uint64_t start = 0, end = 0;
  int timescale_ps = 1;
};

// SAIF text of one instance; with_tc adds each net's (TC n) toggle count.
std::string inst_saif(const RunInst& in, const std::vector<uint64_t>& saif, uint32_t base,
                      bool with_tc);

struct ServerOptions {
  uint64_t device_budget = 0;  // bytes of HBM the server may use (0: 90% of free)
  uint32_t max_running = 4;    // jobs on the device at once
  uint32_t seg_events = 0;     // fused-kernel segment size (0: default)
  double pool_factor = 1.0;    // scales the event-pool estimate of a job
  // queued jobs admitted together run as one batch of up to this many
  // instances (level l of all of them in one launch); 0 or 1: one job per run
  uint32_t max_batch = 1;
};

struct ServerJobInfo {
  int status = 1;              // 1 queued, 2 running, 0 done, < 0 GLS_E_* failure
  double queue_ms = 0, run_ms = 0, latency_ms = 0;
  uint64_t n_events = 0, device_bytes = 0;
  uint32_t attempts = 0, resident = 0, overlap = 0;
};

struct ServerStats {
  uint64_t submitted = 0, done = 0, failed = 0, events = 0, cancelled = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, wall_s = 0, events_per_s = 0;
  uint32_t peak_running = 0, resident_programs = 0;
  uint64_t peak_bytes = 0;
  uint64_t arena_hits = 0, arena_misses = 0;  // device blocks reused / newly allocated
};

class Server {
 public:
  // arena: the device-buffer cache to share (null: the server's own); the
  // default budget counts its cached blocks as free
  Server(int device, const ServerOptions& o, ProgramCache* cache,
         std::shared_ptr<DevArena> arena = nullptr);
  ~Server();  // finishes every submitted job first
  // deadline_s > 0: the job fails with GLS_E_CANCELLED when it has not
  // finished that many seconds after submission (checked while queued and
  // between levels while running).
  uint64_t submit(RunInst inst, std::shared_ptr<const Stimulus> keep, double deadline_s = 0);
  // A queued job is removed at once; a running one stops before its next
  // level.  Returns the job's status after the call (a finished job keeps
  // its result).
  int cancel(uint64_t id);
  // Wait for a job (timeout_s < 0: forever).  Returns its status (1/2 when
  // still queued/running at the timeout).
  int wait(uint64_t id, double timeout_s);
  ServerJobInfo info(uint64_t id);
  std::string saif(uint64_t id);
  ServerStats stats();
  void drain();

 private:
  struct Rec {
    uint64_t id = 0;
    RunInst inst;
    std::shared_ptr<const Stimulus> keep;
    std::shared_ptr<DevProgram> image;
    ServerJobInfo info;
    uint64_t pool_words = 0, reserve = 0;
    std::vector<uint64_t> saif;
    std::chrono::steady_clock::time_point t_submit, t_start, t_end, deadline;
    bool has_deadline = false;
    std::atomic<bool> stop{false};  // cancel asked for a running job
  };
  bool expired(const Rec& r, std::chrono::steady_clock::time_point now) const {
    return r.stop.load(std::memory_order_relaxed) || (r.has_deadline && now >= r.deadline);
  }
  void fail_queued(Rec* r);  // caller holds mu_
  void worker(int slot);
  // Runs one job (one retry with twice the pool when its estimate was
  // short); fills status, events, device bytes, attempts and SAIF.
  void run_one(Rec* r, hipStream_t stream, int& status, uint64_t& events, uint64_t& bytes,
               uint32_t& attempts, std::vector<uint64_t>& saif);
  // Runs jobs admitted together as one batch; false when the batch failed
  // (its jobs are then run one at a time).
  bool run_group(const std::vector<Rec*>& group, hipStream_t stream);
  void finish(Rec* r, int status, uint64_t events, uint64_t bytes, uint32_t attempts,
              std::vector<uint64_t> saif);
  uint64_t estimate(const Rec& r, bool* resident) const;
  uint64_t pool_cap_words() const;
  int device_;
  ServerOptions opt_;
  ProgramCache* cache_;
  std::shared_ptr<DevArena> arena_;  // the jobs' device buffers, reused from job to job
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<uint64_t> queue_;
  std::map<uint64_t, std::unique_ptr<Rec>> jobs_;
  uint64_t next_id_ = 1, in_use_ = 0, peak_bytes_ = 0;
  uint32_t running_ = 0, peak_running_ = 0;
  bool stop_ = false;
  std::vector<std::thread> workers_;
  std::chrono::steady_clock::time_point t0_;
  bool started_ = false;
};

}  // namespace gls
