// Serving loop (see server.h).
#include "server.h"

#include <algorithm>
#include <cmath>

#include "../../include/gls_status.h"

namespace gls {

using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

Server::Server(int device, const ServerOptions& o, ProgramCache* cache,
               std::shared_ptr<DevArena> arena)
    : device_(device), opt_(o), cache_(cache), arena_(std::move(arena)) {
  hip_check(hipSetDevice(device_), "hipSetDevice");
  if (opt_.device_budget == 0) {
    // free HBM plus the shared arena's cached blocks (freed when an
    // allocation needs their room; freeing them now would make the first
    // jobs allocate again, and large allocations are slow)
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const uint64_t cached = arena_ ? arena_->cached_bytes() : 0;
    opt_.device_budget = (uint64_t)((free_b + cached) * 0.9);
  }
  // an own arena keeps cached blocks within the budget; a cached block is
  // freed when an allocation fails
  if (!arena_) arena_ = std::make_shared<DevArena>(opt_.device_budget);
  if (opt_.max_running == 0) opt_.max_running = 1;
  if (!(opt_.pool_factor > 0)) opt_.pool_factor = 1.0;
  for (uint32_t i = 0; i < opt_.max_running; ++i) workers_.emplace_back(&Server::worker, this, (int)i);
}

Server::~Server() {
  drain();
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  for (auto& t : workers_) t.join();
}

// Event-pool words of a job: its stimulus plus an estimate of the level
// outputs (internal nets toggle a few times more often than the inputs, and
// a gate reserves about fan-in times its inputs' events: ten times the
// inputs' mean events per gate), at most cap_words.  A job whose waveforms
// need more runs in time chunks that fit its pool.
static uint64_t job_pool_words(const RunInst& in, double pool_factor, uint64_t cap_words) {
  uint64_t stim = 0;
  for (auto& w : in.st->waves) stim += w.size();
  const uint64_t n_in = std::max<uint64_t>(1, in.st->waves.size());
  const double est = (double)stim + pool_factor * 10.0 * (double)in.pg->gates.size() *
                                        (double)(stim / n_in + 1) + (double)(1u << 20);
  // rounded up to a power of two (at most the cap): jobs of similar sizes
  // then ask for the same block sizes, which the arena hands back cached
  const uint64_t w = std::max<uint64_t>(1u << 22, (uint64_t)std::min(est, (double)cap_words));
  uint64_t c = 1u << 22;
  while (c < w) c <<= 1;
  return std::min(c, std::max<uint64_t>(w, cap_words));
}

// Largest event pool of one job: 60% of the budget shared by the jobs that
// may run at once (the rest is per-net arrays, level scratch and stacks).
uint64_t Server::pool_cap_words() const {
  const double share = 0.6 * (double)opt_.device_budget / std::max<uint32_t>(1, opt_.max_running);
  return std::max<uint64_t>(1ull << 24, (uint64_t)(share / 4));
}

// Device bytes a job needs: its event pool, its resident stimulus (8 bytes
// an event), per-net arrays, level scratch, overflow stacks, and the program
// image when it is not resident yet.
uint64_t Server::estimate(const Rec& r, bool* resident) const {
  const RunInst& in = r.inst;
  const uint64_t gates = in.pg->gates.size(), nets = in.names.size();
  uint64_t stim = 0;
  for (auto& w : in.st->waves) stim += w.size();
  const uint64_t pool = job_pool_words(in, opt_.pool_factor, pool_cap_words());
  const uint64_t ovf = std::min<uint64_t>(1ull << 24, std::max<uint64_t>(1ull << 18, pool / 16));
  uint64_t bytes = pool * 4 + stim * 8 + ovf * 4 + nets * 48 + gates * 64 + (64ull << 20);
  *resident = cache_->resident(in.pg.get(), (uint32_t)in.timescale_ps);
  if (!*resident) bytes += gates * (sizeof(GateRec) + sizeof(GateDev) + 512);
  return bytes;
}

uint64_t Server::submit(RunInst inst, std::shared_ptr<const Stimulus> keep, double deadline_s) {
  auto r = std::make_unique<Rec>();
  r->inst = std::move(inst);
  r->keep = std::move(keep);
  r->t_submit = Clock::now();
  if (deadline_s > 0) {
    r->has_deadline = true;
    r->deadline = r->t_submit + std::chrono::duration_cast<Clock::duration>(
                                    std::chrono::duration<double>(deadline_s));
  }
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw GlsError(GLS_E_STATE, "server is stopping");
    if (!started_) {
      t0_ = r->t_submit;
      started_ = true;
    }
    id = next_id_++;
    r->id = id;
    jobs_[id] = std::move(r);
    queue_.push_back(id);
  }
  cv_work_.notify_all();
  return id;
}

void Server::worker(int slot) {
  (void)slot;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hipStream_t stream = nullptr;
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  for (;;) {
    std::vector<Rec*> group;
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        if (stop_ && queue_.empty()) break;
        // queued jobs cancelled or past their deadline leave the queue first
        const Clock::time_point now = Clock::now();
        bool wake = false, any_deadline = false;
        Clock::time_point next_deadline = Clock::time_point::max();
        for (auto it = queue_.begin(); it != queue_.end();) {
          Rec* c = jobs_[*it].get();
          if (expired(*c, now)) {
            fail_queued(c);
            it = queue_.erase(it);
            wake = true;
            continue;
          }
          if (c->has_deadline) {
            any_deadline = true;
            next_deadline = std::min(next_deadline, c->deadline);
          }
          ++it;
        }
        if (wake) cv_done_.notify_all();
        if (stop_ && queue_.empty()) break;
        // admission: the oldest queued job whose footprint fits what is free;
        // a job may pass the head of the queue only while the head does not
        // fit.  A job that can never fit fails and the scan goes on past it.
        // With batching, later queued jobs without a deadline that also fit
        // join the first one's run.
        uint64_t taken = 0;
        for (auto it = queue_.begin(); it != queue_.end();) {
          Rec* c = jobs_[*it].get();
          bool resident = false;
          const uint64_t need = c->reserve ? c->reserve : estimate(*c, &resident);
          if (!c->reserve) {
            c->reserve = need;
            c->info.resident = resident ? 1u : 0u;
          }
          if (need > opt_.device_budget) {  // can never run: fail it
            c->info.status = GLS_E_OOM;
            c->t_end = Clock::now();
            c->t_start = c->t_end;
            c->info.queue_ms = ms_between(c->t_submit, c->t_end);
            c->info.latency_ms = c->info.queue_ms;
            it = queue_.erase(it);
            cv_done_.notify_all();
            continue;
          }
          const bool joinable = group.empty() || (!c->has_deadline && !group[0]->has_deadline);
          if (joinable && in_use_ + taken + need <= opt_.device_budget) {
            group.push_back(c);
            taken += need;
            it = queue_.erase(it);
            if (group.size() >= std::max<uint32_t>(1, opt_.max_batch)) break;
            continue;
          }
          if (group.empty()) {
            ++it;
            continue;
          }
          ++it;
        }
        if (!group.empty()) break;
        if (any_deadline) cv_work_.wait_until(lk, next_deadline);
        else cv_work_.wait(lk);
      }
      if (group.empty()) break;
      const Clock::time_point now = Clock::now();
      for (Rec* r : group) {
        in_use_ += r->reserve;
        ++running_;
        r->info.status = 2;
        r->t_start = now;
      }
      peak_bytes_ = std::max(peak_bytes_, in_use_);
      peak_running_ = std::max(peak_running_, running_);
      for (Rec* r : group) r->info.overlap = running_ - 1;
    }
    if (group.size() > 1 && run_group(group, stream)) {
      cv_done_.notify_all();
      cv_work_.notify_all();
      continue;
    }
    for (Rec* r : group) {
      int status = 0;
      uint64_t events = 0, bytes = 0;
      uint32_t attempts = 0;
      std::vector<uint64_t> saif;
      run_one(r, stream, status, events, bytes, attempts, saif);
      finish(r, status, events, bytes, attempts, std::move(saif));
      cv_done_.notify_all();
      cv_work_.notify_all();
    }
  }
  (void)hipStreamDestroy(stream);
}

void Server::run_one(Rec* r, hipStream_t stream, int& status, uint64_t& events, uint64_t& bytes,
                     uint32_t& attempts, std::vector<uint64_t>& saif) {
  status = 0;
  try {
    const RunInst& in = r->inst;
    SimInstance I;
    I.pg = in.pg.get();
    I.image = cache_->get(in.pg, (uint32_t)in.timescale_ps);
    I.st = in.st;
    I.stim_nets = in.stim_nets;
    I.start = in.start;
    I.end = in.end;
    uint64_t pool = job_pool_words(in, opt_.pool_factor, pool_cap_words());
    for (uint32_t attempt = 0; attempt < 2; ++attempt) {
      ++attempts;
      try {
        Job job(device_, stream, arena_.get());
        job.seg_target = opt_.seg_events;
        job.should_stop = [this, r] { return expired(*r, Clock::now()); };
        if (job.should_stop()) throw GlsError(GLS_E_CANCELLED, "job cancelled or past its deadline");
        job.prepare({I}, pool);
        job.run();
        job.fetch_saif(saif);
        events = job.stats().n_events;
        bytes = job.device_bytes();
        status = 0;
        break;
      } catch (const GlsError& e) {
        status = e.code;
        if (e.code != GLS_E_OOM || attempt == 1) break;
        pool *= 2;  // the estimate was short: one retry with a larger pool
        std::lock_guard<std::mutex> lk(mu_);
        const uint64_t more = pool * 4 / 2;
        if (in_use_ + more > opt_.device_budget) break;
        in_use_ += more;
        r->reserve += more;
        peak_bytes_ = std::max(peak_bytes_, in_use_);
      }
    }
  } catch (const GlsError& e) {
    status = e.code;
  } catch (const std::exception&) {
    status = GLS_E_DEVICE;
  }
}

bool Server::run_group(const std::vector<Rec*>& group, hipStream_t stream) {
  std::vector<uint64_t> saif;
  std::vector<uint64_t> events(group.size()), base(group.size());
  uint64_t bytes = 0;
  try {
    std::vector<SimInstance> I(group.size());
    uint64_t pool = 0;
    for (size_t i = 0; i < group.size(); ++i) {
      const RunInst& in = group[i]->inst;
      I[i].pg = in.pg.get();
      I[i].image = cache_->get(in.pg, (uint32_t)in.timescale_ps);
      I[i].st = in.st;
      I[i].stim_nets = in.stim_nets;
      I[i].start = in.start;
      I[i].end = in.end;
      pool += job_pool_words(in, opt_.pool_factor, pool_cap_words());
    }
    Job job(device_, stream, arena_.get());
    job.seg_target = opt_.seg_events;
    // the batch stops early only when every job in it was cancelled
    job.should_stop = [this, &group] {
      const Clock::time_point now = Clock::now();
      for (Rec* r : group)
        if (!expired(*r, now)) return false;
      return true;
    };
    job.prepare(I, pool);
    job.run();
    job.fetch_saif(saif);
    for (size_t i = 0; i < group.size(); ++i) {
      events[i] = job.instance_events(i);
      base[i] = job.net_base(i);
    }
    bytes = job.device_bytes();
  } catch (const GlsError& e) {
    if (e.code == GLS_E_CANCELLED) {
      for (Rec* r : group) finish(r, GLS_E_CANCELLED, 0, 0, 1, {});
      return true;
    }
    return false;  // run the jobs one at a time
  } catch (const std::exception&) {
    return false;
  }
  const Clock::time_point now = Clock::now();
  for (size_t i = 0; i < group.size(); ++i) {
    Rec* r = group[i];
    const size_t n = r->inst.names.size();
    std::vector<uint64_t> mine(saif.begin() + (size_t)kSaifWords * base[i],
                               saif.begin() + (size_t)kSaifWords * (base[i] + n));
    // a job cancelled (or past its deadline) while its batch ran keeps no result
    if (expired(*r, now)) finish(r, GLS_E_CANCELLED, 0, 0, 1, {});
    else finish(r, 0, events[i], bytes / group.size(), 1, std::move(mine));
  }
  return true;
}

void Server::finish(Rec* r, int status, uint64_t events, uint64_t bytes, uint32_t attempts,
                    std::vector<uint64_t> saif) {
  std::lock_guard<std::mutex> lk(mu_);
  r->t_end = Clock::now();
  r->info.status = status;
  r->info.n_events = events;
  r->info.device_bytes = bytes;
  r->info.attempts = attempts;
  r->info.queue_ms = ms_between(r->t_submit, r->t_start);
  r->info.run_ms = ms_between(r->t_start, r->t_end);
  r->info.latency_ms = ms_between(r->t_submit, r->t_end);
  r->saif = std::move(saif);
  in_use_ -= r->reserve;
  --running_;
}

void Server::fail_queued(Rec* r) {
  r->info.status = GLS_E_CANCELLED;
  r->t_end = Clock::now();
  r->t_start = r->t_end;
  r->info.queue_ms = ms_between(r->t_submit, r->t_end);
  r->info.latency_ms = r->info.queue_ms;
}

int Server::cancel(uint64_t id) {
  int st;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) throw GlsError(GLS_E_ARG, "unknown job");
    Rec* r = it->second.get();
    if (r->info.status <= 0) return r->info.status;  // finished: keeps its result
    r->stop.store(true, std::memory_order_relaxed);
    if (r->info.status == 1) {
      auto q = std::find(queue_.begin(), queue_.end(), id);
      if (q != queue_.end()) queue_.erase(q);
      fail_queued(r);
    }
    st = r->info.status;
  }
  cv_done_.notify_all();
  cv_work_.notify_all();
  return st;
}

int Server::wait(uint64_t id, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw GlsError(GLS_E_ARG, "unknown job");
  Rec* r = it->second.get();
  auto done = [&] { return r->info.status <= 0; };
  if (timeout_s < 0) cv_done_.wait(lk, done);
  else cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  return r->info.status;
}

ServerJobInfo Server::info(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw GlsError(GLS_E_ARG, "unknown job");
  return it->second->info;
}

std::string Server::saif(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw GlsError(GLS_E_ARG, "unknown job");
  Rec* r = it->second.get();
  if (r->info.status != 0) throw GlsError(GLS_E_STATE, "job not finished successfully");
  return inst_saif(r->inst, r->saif, 0, false);
}

void Server::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] {
    if (!queue_.empty() || running_) return false;
    return true;
  });
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  ServerStats s;
  std::vector<double> lat;
  Clock::time_point last = t0_;
  for (auto& kv : jobs_) {
    const Rec& r = *kv.second;
    ++s.submitted;
    if (r.info.status == 0) {
      ++s.done;
      s.events += r.info.n_events;
      lat.push_back(r.info.latency_ms);
      last = std::max(last, r.t_end);
    } else if (r.info.status < 0) {
      ++s.failed;
      if (r.info.status == GLS_E_CANCELLED) ++s.cancelled;
    }
  }
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      // nearest rank
      size_t k = (size_t)std::ceil(p / 100.0 * (double)lat.size());
      return lat[std::min(lat.size() - 1, k ? k - 1 : 0)];
    };
    s.p50_ms = pct(50);
    s.p99_ms = pct(99);
    double sum = 0;
    for (double x : lat) sum += x;
    s.mean_ms = sum / (double)lat.size();
  }
  s.wall_s = started_ ? std::chrono::duration<double>(last - t0_).count() : 0;
  s.events_per_s = s.wall_s > 0 ? (double)s.events / s.wall_s : 0;
  s.peak_running = peak_running_;
  s.peak_bytes = peak_bytes_;
  s.resident_programs = (uint32_t)cache_->size();
  s.arena_hits = arena_->hits();
  s.arena_misses = arena_->misses();
  return s;
}

}  // namespace gls
