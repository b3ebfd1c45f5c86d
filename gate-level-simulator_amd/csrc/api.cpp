// C ABI of libgls (include/gls.h).
#include "../../include/gls.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>

#include "csr.h"
#include "engine.h"
#include "formats.h"
#include "gen.h"
#include "interdat.h"
#include "server.h"

using namespace gls;

static thread_local std::string g_err;

struct gls_design {
  Library lib;
  Netlist nl;
  std::string vlib_text, gv_text, sdf_text;  // empty when generated
  // compiled programs by the set of netlist nets the stimulus drives (the
  // VCD decides which nets are driven from outside; its other signals and
  // their order do not change the program)
  std::mutex mu;
  std::map<std::vector<std::string>, std::shared_ptr<const Program>> progs;
};
struct gls_stim {
  Stimulus st;
};
// A HIP stream destroyed with the last of its holders: a context and the
// runs prepared on it (a run may be freed after its context, and its buffers
// go back to the arena only once the stream has drained).
struct StreamRef {
  hipStream_t s = nullptr;
  ~StreamRef() {
    if (s) (void)hipStreamDestroy(s);
  }
};
struct gls_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // stream_ref->s
  std::shared_ptr<StreamRef> stream_ref;
  ProgramCache cache;  // resident program images of this device
  // device buffers of finished runs, reused by later runs and servers of
  // this context (large device allocations and frees can take seconds)
  std::shared_ptr<DevArena> arena;
};
struct gls_run {
  std::shared_ptr<StreamRef> stream_ref;  // outlive the job, which syncs the stream and
  std::shared_ptr<DevArena> arena;        // returns its buffers to the arena
  std::unique_ptr<Job> job;
  std::vector<RunInst> inst;
  bool keep_waves = false;
  bool saif_tc = false;  // SAIF text carries (TC n) records
  bool executed = false;
  std::vector<uint64_t> saif;  // kSaifWords per net of the union net space
};

namespace gls {
static std::vector<std::string> program_key(const Netlist& nl, const std::vector<std::string>& names) {
  std::vector<std::string> k;
  for (auto& n : names)
    if (nl.net(n) >= 0) k.push_back(n);
  std::sort(k.begin(), k.end());
  k.erase(std::unique(k.begin(), k.end()), k.end());
  return k;
}

std::shared_ptr<const Program> design_program(gls_design* d, const std::vector<std::string>& names) {
  auto key = program_key(d->nl, names);
  std::lock_guard<std::mutex> lk(d->mu);
  auto it = d->progs.find(key);
  if (it != d->progs.end()) return it->second;
  auto pg = std::make_shared<const Program>(compile_program(d->nl, d->lib, names));
  d->progs.emplace(std::move(key), pg);
  return pg;
}

RunInst make_inst(gls_design* d, const gls_stim* s, uint64_t dumpon, uint64_t dumpoff) {
  RunInst r;
  r.pg = design_program(d, s->st.names);
  r.names = r.pg->net_names;
  r.in_saif = r.pg->in_saif;
  r.is_gate.assign(r.names.size(), 0);
  for (auto& g : r.pg->gates) r.is_gate[g.out_net] = 1;
  for (size_t i = 0; i < r.names.size(); ++i) r.by_name.emplace(r.names[i], (int)i);
  for (auto& nm : s->st.names) {
    int id = d->nl.net(nm);
    if (id < 0) {  // a VCD signal the netlist does not use: SAIF record only
      auto it = r.by_name.find(nm);
      if (it != r.by_name.end()) id = it->second;
      else {
        id = (int)r.names.size();
        r.names.push_back(nm);
        r.in_saif.push_back(1);
        r.is_gate.push_back(0);
        r.by_name.emplace(nm, id);
      }
    }
    r.stim_nets.push_back(id);
  }
  r.st = &s->st;
  r.start = dumpon;
  r.end = dumpoff;
  r.timescale_ps = s->st.timescale_ps;
  return r;
}

std::string inst_saif(const RunInst& in, const std::vector<uint64_t>& saif, uint32_t base,
                      bool with_tc) {
  std::vector<SaifRecord> recs;
  for (size_t n = 0; n < in.names.size(); ++n) {
    if (!in.in_saif[n]) continue;
    SaifRecord s;
    const uint64_t* c = &saif[(size_t)kSaifWords * (base + n)];
    s.name = in.names[n];
    s.t0 = c[0];
    s.t1 = c[1];
    s.tx = c[2];
    s.tc = c[3];
    recs.push_back(std::move(s));
  }
  return saif_text(std::move(recs), in.start, in.end, in.timescale_ps, with_tc);
}
}  // namespace gls

#define API_BEGIN try {
#define API_END                                  \
  }                                              \
  catch (const GlsError& e) {                    \
    g_err = e.what();                            \
    return e.code;                               \
  }                                              \
  catch (const std::bad_alloc&) {                \
    g_err = "host out of memory";                \
    return GLS_E_OOM;                            \
  }                                              \
  catch (const std::exception& e) {              \
    g_err = e.what();                            \
    return GLS_E_FORMAT;                         \
  }                                              \
  g_err.clear();                                 \
  return GLS_OK;

extern "C" {

const char* gls_last_error(void) { return g_err.c_str(); }
const char* gls_version(void) { return "gls-mi355x 0.1"; }

int gls_design_from_texts_threads(const char* vlib, const char* gv, const char* sdf, int threads,
                                  gls_design** out) {
  API_BEGIN
  if (!vlib || !gv || !out) throw GlsError(GLS_E_ARG, "null argument");
  if (threads < 0) throw GlsError(GLS_E_ARG, "negative thread count");
  auto d = std::make_unique<gls_design>();
  d->vlib_text = vlib;
  d->gv_text = gv;
  d->sdf_text = sdf ? sdf : "";
  d->lib = parse_vlib_text(d->vlib_text);
  d->nl = parse_gv_text(d->gv_text, d->lib, threads);
  if (sdf) parse_sdf_text(d->sdf_text, d->nl, threads);
  *out = d.release();
  API_END
}

int gls_design_from_texts(const char* vlib, const char* gv, const char* sdf, gls_design** out) {
  return gls_design_from_texts_threads(vlib, gv, sdf, 0, out);
}

int gls_design_digest(const gls_design* d, uint64_t* digest) {
  API_BEGIN
  if (!d || !digest) throw GlsError(GLS_E_ARG, "null argument");
  *digest = netlist_digest(d->nl);
  API_END
}

int gls_design_from_files(const char* vlib, const char* gv, const char* sdf, gls_design** out) {
  API_BEGIN
  if (!vlib || !gv || !out) throw GlsError(GLS_E_ARG, "null argument");
  std::string a = read_file(vlib), b = read_file(gv), c = sdf ? read_file(sdf) : std::string();
  int rc = gls_design_from_texts(a.c_str(), b.c_str(), sdf ? c.c_str() : nullptr, out);
  if (rc != GLS_OK) return rc;
  API_END
}

int gls_design_generate(int config, uint64_t seed, int n_cells, int n_levels, int n_pis,
                        gls_design** out) {
  API_BEGIN
  if (!out) throw GlsError(GLS_E_ARG, "null argument");
  auto d = std::make_unique<gls_design>();
  d->lib = parse_vlib_text(generic_vlib_text());
  NetlistSpec sp = config_netlist_spec(config);
  if (seed) sp.seed = seed;
  if (n_cells > 0) sp.n_cells = n_cells;
  if (n_levels > 0) sp.n_levels = n_levels;
  if (n_pis > 0) sp.n_pis = n_pis;
  d->nl = generate_netlist(sp, d->lib);
  *out = d.release();
  API_END
}

static void ensure_texts(const gls_design* d, std::string& v, std::string& g, std::string& s) {
  if (!d->gv_text.empty()) {
    v = d->vlib_text;
    g = d->gv_text;
    s = d->sdf_text;
  } else {
    v = generic_vlib_text();
    g = gv_text(d->nl, d->lib);
    s = sdf_text(d->nl, d->lib);
  }
}

static std::vector<std::string> declared_inputs(const Netlist& nl) {
  std::vector<std::string> r;
  for (size_t n = 0; n < nl.net_names.size(); ++n)
    if (nl.net_is_input[n]) r.push_back(nl.net_names[n]);
  return r;
}

// inter.dat: the `inter` archive of the reference's process step, then the
// library, netlist, program (compiled against the declared inputs) and
// source texts (src/process.cpp:604-606 writes, src/simulate.cu:1835-1838
// reads).
int gls_design_save(const gls_design* d, const char* path) {
  API_BEGIN
  if (!d || !path) throw GlsError(GLS_E_ARG, "null argument");
  auto pg = design_program(const_cast<gls_design*>(d), declared_inputs(d->nl));
  DesignTexts t{d->vlib_text, d->gv_text, d->sdf_text};
  write_file(path, inter_file_text(d->lib, d->nl, *pg, t));
  API_END
}

static int load_text_sections(const std::string& a, gls_design** out) {
  size_t p = 0;
  auto line = [&]() {
    size_t q = a.find('\n', p);
    if (q == std::string::npos) throw GlsError(GLS_E_FORMAT, "inter.dat: truncated");
    std::string r = a.substr(p, q - p);
    p = q + 1;
    return r;
  };
  line();
  std::string parts[3];
  const char* keys[3] = {"vlib", "gv", "sdf"};
  for (int i = 0; i < 3; ++i) {
    std::string h = line();
    size_t sp = h.find(' ');
    if (sp == std::string::npos || h.substr(0, sp) != keys[i])
      throw GlsError(GLS_E_FORMAT, "inter.dat: missing section");
    size_t n = std::stoull(h.substr(sp + 1));
    if (p + n + 1 > a.size()) throw GlsError(GLS_E_FORMAT, "inter.dat: truncated section");
    parts[i] = a.substr(p, n);
    p += n + 1;
  }
  return gls_design_from_texts(parts[0].c_str(), parts[1].c_str(), parts[2].c_str(), out);
}

int gls_design_load(const char* path, gls_design** out) {
  API_BEGIN
  if (!path || !out) throw GlsError(GLS_E_ARG, "null argument");
  std::string a = read_file(path);
  // files of the earlier layout: the three source texts
  if (a.compare(0, 12, "GLS-INTER 1\n") == 0) return load_text_sections(a, out);
  auto d = std::make_unique<gls_design>();
  Program pg;
  DesignTexts t;
  if (!read_inter_section(a, d->lib, d->nl, pg, t))
    throw GlsError(GLS_E_FORMAT, "inter.dat: no compiled section");
  d->vlib_text = std::move(t.vlib);
  d->gv_text = std::move(t.gv);
  d->sdf_text = std::move(t.sdf);
  d->progs.emplace(program_key(d->nl, declared_inputs(d->nl)),
                   std::make_shared<const Program>(std::move(pg)));
  *out = d.release();
  API_END
}

int gls_design_write(const gls_design* d, const char* vlib, const char* gv, const char* sdf) {
  API_BEGIN
  if (!d) throw GlsError(GLS_E_ARG, "null argument");
  std::string v, g, s;
  ensure_texts(d, v, g, s);
  if (vlib) write_file(vlib, v);
  if (gv) write_file(gv, g);
  if (sdf) write_file(sdf, s);
  API_END
}

int gls_design_info_get(const gls_design* d, gls_design_info* info) {
  API_BEGIN
  if (!d || !info) throw GlsError(GLS_E_ARG, "null argument");
  const Program& pg = *design_program(const_cast<gls_design*>(d), declared_inputs(d->nl));
  info->n_instances = d->nl.insts.size();
  info->n_nets = d->nl.net_names.size();
  info->n_inputs = declared_inputs(d->nl).size();
  info->n_gates = pg.gates.size();
  info->n_levels = (uint64_t)pg.n_levels();
  info->n_pins = pg.n_pins;
  info->program_digest = program_digest(pg);
  API_END
}

void gls_design_free(gls_design* d) { delete d; }

int gls_stim_from_vcd(const char* path, gls_stim** out) {
  API_BEGIN
  if (!path || !out) throw GlsError(GLS_E_ARG, "null argument");
  auto s = std::make_unique<gls_stim>();
  s->st = parse_vcd_file(path);
  *out = s.release();
  API_END
}

int gls_stim_from_vcd_threads(const char* path, int threads, gls_stim** out) {
  API_BEGIN
  if (!path || !out) throw GlsError(GLS_E_ARG, "null argument");
  if (threads < 0) throw GlsError(GLS_E_ARG, "negative thread count");
  auto s = std::make_unique<gls_stim>();
  s->st = parse_vcd_file(path, threads);
  *out = s.release();
  API_END
}

int gls_stim_generate(const gls_design* d, uint64_t seed, uint32_t t_end, double mean_gap,
                      double x_prob, gls_stim** out) {
  API_BEGIN
  if (!d || !out) throw GlsError(GLS_E_ARG, "null argument");
  if (!(mean_gap >= 1.0)) throw GlsError(GLS_E_ARG, "mean_gap must be >= 1");
  StimulusSpec sp;
  sp.seed = seed;
  sp.t_end = t_end;
  sp.mean_gap = mean_gap;
  sp.x_prob = x_prob;
  auto s = std::make_unique<gls_stim>();
  s->st = generate_stimulus(d->nl, sp);
  *out = s.release();
  API_END
}

int gls_stim_write_vcd(const gls_stim* s, const char* path) {
  API_BEGIN
  if (!s || !path) throw GlsError(GLS_E_ARG, "null argument");
  write_file(path, vcd_text(s->st));
  API_END
}

int gls_stim_info(const gls_stim* s, uint64_t* n_signals, uint64_t* n_events) {
  API_BEGIN
  if (!s) throw GlsError(GLS_E_ARG, "null argument");
  uint64_t ev = 0;
  for (auto& w : s->st.waves) ev += w.size();
  if (n_signals) *n_signals = s->st.names.size();
  if (n_events) *n_events = ev;
  API_END
}

void gls_stim_free(gls_stim* s) { delete s; }

int gls_ctx_mem_info(gls_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes) {
  API_BEGIN
  if (!c || !free_bytes || !total_bytes) throw GlsError(GLS_E_ARG, "null argument");
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  size_t f = 0, t = 0;
  hip_check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
  *free_bytes = f + (c->arena ? c->arena->cached_bytes() : 0);  // cached blocks are reusable
  *total_bytes = t;
  API_END
}

int gls_ctx_create(int device, gls_ctx** out) {
  API_BEGIN
  if (!out) throw GlsError(GLS_E_ARG, "null argument");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) throw GlsError(GLS_E_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) throw GlsError(GLS_E_ARG, "device index out of range");
  auto c = std::make_unique<gls_ctx>();
  c->device = device;
  hip_check(hipSetDevice(device), "hipSetDevice");
  c->stream_ref = std::make_shared<StreamRef>();
  hip_check(hipStreamCreateWithFlags(&c->stream_ref->s, hipStreamNonBlocking), "hipStreamCreate");
  c->stream = c->stream_ref->s;
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  c->arena = std::make_shared<DevArena>((uint64_t)total_b);
  *out = c.release();
  API_END
}

void gls_ctx_destroy(gls_ctx* c) {
  if (!c) return;
  // the stream goes with the last run prepared on this context
  delete c;
}

static void apply_opts(Job& job, const gls_sim_opts* opts) {
  if (!opts) return;
  if (opts->path > GLS_PATH_TWO_PASS) throw GlsError(GLS_E_ARG, "unknown path");
  if (opts->seg_events) job.seg_target = opts->seg_events;
  job.mode = (int)opts->path;
  job.level_blocks_opt = opts->level_blocks;
  job.prefetch = opts->no_prefetch ? 0u : 1u;
  if (opts->graph && opts->path != GLS_PATH_FUSED)
    throw GlsError(GLS_E_UNSUPPORTED, "graph replay needs the fused path");
  job.graph = opts->graph ? 1u : 0u;
  job.stack_mode = opts->hbm_stacks ? 1u : 0u;
  job.pool_grow = opts->pool_grow ? 1u : 0u;
  job.ovf_words_opt = opts->ovf_words;
  job.saif_tc = opts->saif_tc ? 1u : 0u;
  job.keep_waves = opts->keep_waves ? 1u : 0u;
  if (opts->path == GLS_PATH_TWO_PASS && opts->chunk_ps != 0 && opts->chunk_ps != GLS_NO_CHUNKS)
    throw GlsError(GLS_E_UNSUPPORTED, "time chunks need the fused path");
  job.chunk_ps = opts->chunk_ps;
}

int gls_prepare_batch(gls_ctx* c, size_t n, const gls_design* const* designs,
                      const gls_stim* const* stims, const uint64_t* dumpon, const uint64_t* dumpoff,
                      const gls_sim_opts* opts, gls_run** out) {
  API_BEGIN
  if (!c || !designs || !stims || !dumpon || !dumpoff || !out || n == 0)
    throw GlsError(GLS_E_ARG, "null argument");
  auto r = std::make_unique<gls_run>();
  std::vector<SimInstance> inst;
  for (size_t i = 0; i < n; ++i) {
    if (!designs[i] || !stims[i]) throw GlsError(GLS_E_ARG, "null design or stimulus");
    r->inst.push_back(make_inst(const_cast<gls_design*>(designs[i]), stims[i], dumpon[i], dumpoff[i]));
  }
  for (auto& in : r->inst) {
    SimInstance I;
    I.pg = in.pg.get();
    // the context's image of the program: built once per netlist, so the
    // instances of one netlist share its truth and delay tables in the batch
    I.image = c->cache.get(in.pg, (uint32_t)in.timescale_ps);
    I.st = in.st;
    I.stim_nets = in.stim_nets;
    I.start = in.start;
    I.end = in.end;
    inst.push_back(std::move(I));
  }
  r->keep_waves = opts && opts->keep_waves;
  r->saif_tc = opts && opts->saif_tc;
  r->stream_ref = c->stream_ref;
  r->arena = c->arena;
  r->job = std::make_unique<Job>(c->device, c->stream, c->arena.get());
  apply_opts(*r->job, opts);
  r->job->prepare(inst, opts ? opts->pool_bytes / 4 : 0);
  // sized once here: executions only copy into it
  r->saif.resize((size_t)r->job->stats().n_nets * kSaifWords);
  *out = r.release();
  API_END
}

int gls_prepare(gls_ctx* c, const gls_design* d, const gls_stim* s, uint64_t dumpon,
                uint64_t dumpoff, const gls_sim_opts* opts, gls_run** out) {
  return gls_prepare_batch(c, 1, &d, &s, &dumpon, &dumpoff, opts, out);
}

int gls_run_execute(gls_run* r) {
  API_BEGIN
  if (!r) throw GlsError(GLS_E_ARG, "null argument");
  r->job->run();
  r->job->fetch_saif(r->saif);
  r->executed = true;
  API_END
}

int gls_run_execute_static(gls_run* r) {
  API_BEGIN
  if (!r) throw GlsError(GLS_E_ARG, "null argument");
  r->job->run_static();
  r->job->fetch_saif(r->saif);
  r->executed = true;
  API_END
}

int gls_simulate(gls_ctx* c, const gls_design* d, const gls_stim* s, uint64_t dumpon,
                 uint64_t dumpoff, const gls_sim_opts* opts, gls_run** out) {
  int rc = gls_prepare(c, d, s, dumpon, dumpoff, opts, out);
  if (rc != GLS_OK) return rc;
  rc = gls_run_execute(*out);
  if (rc != GLS_OK) {
    std::string msg = g_err;
    gls_run_free(*out);
    *out = nullptr;
    g_err = msg;
  }
  return rc;
}

int gls_run_stats_get(const gls_run* r, gls_run_stats* st) {
  API_BEGIN
  if (!r || !st) throw GlsError(GLS_E_ARG, "null argument");
  const JobStats& j = r->job->stats();
  st->n_events = j.n_events;
  st->in_words = j.in_words;
  st->out_words = j.out_words;
  st->alg_bytes = j.alg_bytes;
  st->n_nets = j.n_nets;
  st->n_gates = j.n_gates;
  st->n_pins = j.n_pins;
  st->kernel_ms = j.kernel_ms;
  st->wall_ms = j.wall_ms;
  st->eval_ms = j.eval_ms;
  st->eval_launches = j.eval_launches;
  st->level_blocks = j.level_blocks;
  st->pool_bytes = j.pool_words * 4;
  st->tiles = j.tiles;
  st->tiles_hbm = j.tiles_hbm;
  st->pool_used_bytes = j.pool_words_used * 4;
  st->chunks = j.chunks;
  st->chunk_retries = j.chunk_retries;
  st->device_allocs = j.device_allocs;
  API_END
}

static const RunInst& inst_of(const gls_run* r, size_t i) {
  if (i >= r->inst.size()) throw GlsError(GLS_E_ARG, "instance index out of range");
  return r->inst[i];
}

static std::string make_saif(const gls_run* r, size_t i) {
  if (!r->executed) throw GlsError(GLS_E_STATE, "run not executed");
  return inst_saif(inst_of(r, i), r->saif, r->job->net_base(i), r->saif_tc);
}

int gls_run_instances(const gls_run* r, size_t* n) {
  API_BEGIN
  if (!r || !n) throw GlsError(GLS_E_ARG, "null argument");
  *n = r->inst.size();
  API_END
}

int gls_run_write_saif(const gls_run* r, const char* path) {
  API_BEGIN
  if (!r || !path) throw GlsError(GLS_E_ARG, "null argument");
  write_file(path, make_saif(r, 0));
  API_END
}

static void copy_text(const std::string& s, char* buf, size_t cap, size_t* len) {
  *len = s.size();
  if (buf) {
    if (cap < s.size()) throw GlsError(GLS_E_ARG, "buffer too small");
    std::memcpy(buf, s.data(), s.size());
  }
}

int gls_run_saif_text(const gls_run* r, char* buf, size_t cap, size_t* len) {
  return gls_run_instance_saif_text(r, 0, buf, cap, len);
}

int gls_run_instance_saif_text(const gls_run* r, size_t i, char* buf, size_t cap, size_t* len) {
  API_BEGIN
  if (!r || !len) throw GlsError(GLS_E_ARG, "null argument");
  copy_text(make_saif(r, i), buf, cap, len);
  API_END
}

int gls_run_net_saif(const gls_run* r, const char* net, uint64_t v[3]) {
  API_BEGIN
  if (!r || !net || !v) throw GlsError(GLS_E_ARG, "null argument");
  if (!r->executed) throw GlsError(GLS_E_STATE, "run not executed");
  const RunInst& in = inst_of(r, 0);
  auto it = in.by_name.find(net);
  if (it == in.by_name.end()) throw GlsError(GLS_E_ARG, std::string("unknown net ") + net);
  for (int i = 0; i < 3; ++i) v[i] = r->saif[(size_t)kSaifWords * it->second + i];
  API_END
}

int gls_run_net_toggles(const gls_run* r, const char* net, uint64_t* tc) {
  API_BEGIN
  if (!r || !net || !tc) throw GlsError(GLS_E_ARG, "null argument");
  if (!r->executed) throw GlsError(GLS_E_STATE, "run not executed");
  if (!r->saif_tc) throw GlsError(GLS_E_STATE, "toggle counts need the saif_tc option");
  const RunInst& in = inst_of(r, 0);
  auto it = in.by_name.find(net);
  if (it == in.by_name.end()) throw GlsError(GLS_E_ARG, std::string("unknown net ") + net);
  *tc = r->saif[(size_t)kSaifWords * it->second + 3];
  API_END
}

// Waveform of net `id` of instance i as absolute (time << 2 | value) words:
// gate outputs from the run, stimulus signals and constants from their
// inputs.
static void net_words(const gls_run* r, size_t i, int id, std::vector<uint64_t>& out) {
  const RunInst& in = inst_of(r, i);
  out.clear();
  if (in.is_gate[id]) {
    const auto& W = r->job->waves();
    const size_t n = r->job->net_base(i) + (size_t)id;
    if (n < W.size()) out = W[n];
    return;
  }
  for (size_t s = 0; s < in.stim_nets.size(); ++s)
    if (in.stim_nets[s] == id) {
      const Wave& w = in.st->waves[s];
      for (size_t q = 0; q < w.size(); ++q) out.push_back(((uint64_t)w.t[q] << 2) | (w.v[q] & 3u));
      return;
    }
  if (id < (int)in.pg->net_names.size() && in.pg->role[id] == NR_CONST)
    out.push_back((uint64_t)(in.pg->const_val[id] & 3));
}

static int net_id(const gls_run* r, const char* net) {
  if (!r->executed || !r->keep_waves) throw GlsError(GLS_E_STATE, "waveforms not kept");
  const RunInst& in = inst_of(r, 0);
  auto it = in.by_name.find(net);
  if (it == in.by_name.end()) throw GlsError(GLS_E_ARG, std::string("unknown net ") + net);
  return it->second;
}

int gls_run_waveform(const gls_run* r, const char* net, uint32_t* words, size_t cap, size_t* n) {
  API_BEGIN
  if (!r || !net || !n) throw GlsError(GLS_E_ARG, "null argument");
  std::vector<uint64_t> w;
  net_words(r, 0, net_id(r, net), w);
  *n = w.size();
  if (words) {
    if (cap < w.size()) throw GlsError(GLS_E_ARG, "buffer too small");
    for (size_t q = 0; q < w.size(); ++q) {
      if ((w[q] >> 2) >= (1ull << 30))
        throw GlsError(GLS_E_RANGE, "time of 2^30 or more: use gls_run_waveform64");
      words[q] = (uint32_t)w[q];
    }
  }
  API_END
}

int gls_run_waveform64(const gls_run* r, const char* net, uint64_t* words, size_t cap, size_t* n) {
  API_BEGIN
  if (!r || !net || !n) throw GlsError(GLS_E_ARG, "null argument");
  std::vector<uint64_t> w;
  net_words(r, 0, net_id(r, net), w);
  *n = w.size();
  if (words) {
    if (cap < w.size()) throw GlsError(GLS_E_ARG, "buffer too small");
    std::memcpy(words, w.data(), w.size() * 8);
  }
  API_END
}

// Per-net FNV-1a over the name bytes then the little-endian 32-bit event
// words ((time << 2 | value) truncated to 32 bits), finalised with
// SplitMix64 and summed over gate-output nets.
int gls_run_instance_events(const gls_run* r, size_t i, uint64_t* n) {
  API_BEGIN
  if (!r || !n) throw GlsError(GLS_E_ARG, "null argument");
  if (!r->executed) throw GlsError(GLS_E_STATE, "run not executed");
  inst_of(r, i);  // range check
  *n = r->job->instance_events(i);
  API_END
}

int gls_run_instance_waves_hash(const gls_run* r, size_t i, uint64_t* hash) {
  API_BEGIN
  if (!r || !hash) throw GlsError(GLS_E_ARG, "null argument");
  if (!r->executed || !r->keep_waves) throw GlsError(GLS_E_STATE, "waveforms not kept");
  const RunInst& in = inst_of(r, i);
  const auto& W = r->job->waves();
  const uint32_t base = r->job->net_base(i);
  uint64_t total = 0;
  static const std::vector<uint64_t> empty;
  for (size_t n = 0; n < in.names.size(); ++n) {
    if (!in.is_gate[n]) continue;
    uint64_t h = 1469598103934665603ull;
    for (unsigned char ch : in.names[n]) h = (h ^ ch) * 1099511628211ull;
    const std::vector<uint64_t>& w = base + n < W.size() ? W[base + n] : empty;
    for (uint64_t e : w) {
      const uint32_t x = (uint32_t)e;
      for (int b = 0; b < 4; ++b) h = (h ^ ((x >> (8 * b)) & 0xFF)) * 1099511628211ull;
    }
    uint64_t z = h + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    total += z ^ (z >> 31);
  }
  *hash = total;
  API_END
}

int gls_run_waves_hash(const gls_run* r, uint64_t* hash) {
  return gls_run_instance_waves_hash(r, 0, hash);
}

// Autotuner: runs the prepared simulation once per candidate setting and
// keeps the fastest.  Every candidate computes identical results, so the
// run's results stay valid.
int gls_run_autotune(gls_run* r, gls_tune_report* rep) {
  API_BEGIN
  if (!r) throw GlsError(GLS_E_ARG, "null argument");
  if (r->job->mode != 0) throw GlsError(GLS_E_UNSUPPORTED, "the autotuner tunes the fused path");
  Job& j = *r->job;
  gls_tune_report R;
  std::memset(&R, 0, sizeof R);
  auto trial = [&](uint32_t seg, uint32_t blocks, uint32_t pf) {
    j.seg_target = seg;
    j.level_blocks_opt = blocks;
    j.prefetch = pf;
    j.run();
    const double ms = j.stats().kernel_ms;
    if (R.n_tried < GLS_TUNE_MAX) {
      R.tried_seg[R.n_tried] = seg;
      R.tried_blocks[R.n_tried] = (uint32_t)j.stats().level_blocks;
      R.tried_prefetch[R.n_tried] = pf;
      R.tried_ms[R.n_tried] = ms;
      ++R.n_tried;
    }
    return ms;
  };
  // LDS tile size (events per segment), then launch width, then staging
  uint32_t best_seg = 32;
  double best = 1e300;
  for (uint32_t seg : {24u, 32u, 48u, 64u}) {
    double ms = trial(seg, 0, 1);
    if (ms < best) best = ms, best_seg = seg;
  }
  j.seg_target = best_seg;
  j.level_blocks_opt = 0;
  j.run();
  const uint32_t occ = (uint32_t)j.stats().level_blocks;
  uint32_t best_blocks = 0;
  for (uint32_t b : {occ * 3 / 4, occ / 2}) {
    if (b == 0) continue;
    double ms = trial(best_seg, b, 1);
    if (ms < best) best = ms, best_blocks = b;
  }
  uint32_t best_pf = 1;
  {
    double ms = trial(best_seg, best_blocks, 0);
    if (ms < best) best = ms, best_pf = 0;
  }
  j.seg_target = best_seg;
  j.level_blocks_opt = best_blocks;
  j.prefetch = best_pf;
  j.run();
  r->job->fetch_saif(r->saif);
  r->executed = true;
  R.seg_events = best_seg;
  R.level_blocks = (uint32_t)j.stats().level_blocks;
  R.prefetch = best_pf;
  R.best_ms = j.stats().kernel_ms;
  if (rep) *rep = R;
  API_END
}

void gls_run_free(gls_run* r) { delete r; }

// ------------------------------------------------------------- serving ---

struct gls_server {
  std::unique_ptr<Server> srv;
  std::mutex mu;
  std::vector<std::shared_ptr<const Stimulus>> stims;
};

int gls_server_create(gls_ctx* c, const gls_server_opts* o, gls_server** out) {
  API_BEGIN
  if (!c || !out) throw GlsError(GLS_E_ARG, "null argument");
  ServerOptions so;
  if (o) {
    so.device_budget = o->device_budget;
    so.max_running = o->max_running ? o->max_running : 4;
    so.seg_events = o->seg_events;
    so.pool_factor = o->pool_factor > 0 ? o->pool_factor : 1.0;
    so.max_batch = o->max_batch ? o->max_batch : 1;
  }
  auto s = std::make_unique<gls_server>();
  s->srv = std::make_unique<Server>(c->device, so, &c->cache, c->arena);
  *out = s.release();
  API_END
}

int gls_server_submit(gls_server* s, const gls_design* d, const gls_stim* st, uint64_t dumpon,
                      uint64_t dumpoff, uint64_t* job) {
  return gls_server_submit_deadline(s, d, st, dumpon, dumpoff, 0.0, job);
}

int gls_server_submit_deadline(gls_server* s, const gls_design* d, const gls_stim* st,
                               uint64_t dumpon, uint64_t dumpoff, double deadline_s,
                               uint64_t* job) {
  API_BEGIN
  if (!s || !d || !st || !job) throw GlsError(GLS_E_ARG, "null argument");
  if (dumpoff <= dumpon) throw GlsError(GLS_E_ARG, "dumpoff time must be above dumpon time");
  // the server keeps its own copy of the stimulus: the caller may free theirs
  auto keep = std::make_shared<const Stimulus>(st->st);
  RunInst in = make_inst(const_cast<gls_design*>(d), st, dumpon, dumpoff);
  in.st = keep.get();
  *job = s->srv->submit(std::move(in), keep, deadline_s > 0 ? deadline_s : 0.0);
  API_END
}

int gls_server_cancel(gls_server* s, uint64_t job, int* status) {
  API_BEGIN
  if (!s) throw GlsError(GLS_E_ARG, "null argument");
  const int st = s->srv->cancel(job);
  if (status) *status = st;
  API_END
}

int gls_server_wait(gls_server* s, uint64_t job, double timeout_s, int* status) {
  API_BEGIN
  if (!s || !status) throw GlsError(GLS_E_ARG, "null argument");
  *status = s->srv->wait(job, timeout_s);
  API_END
}

int gls_server_job_info(gls_server* s, uint64_t job, gls_job_info* info) {
  API_BEGIN
  if (!s || !info) throw GlsError(GLS_E_ARG, "null argument");
  ServerJobInfo j = s->srv->info(job);
  info->status = j.status;
  info->queue_ms = j.queue_ms;
  info->run_ms = j.run_ms;
  info->latency_ms = j.latency_ms;
  info->n_events = j.n_events;
  info->device_bytes = j.device_bytes;
  info->attempts = j.attempts;
  info->resident = j.resident;
  info->overlap = j.overlap;
  API_END
}

int gls_server_job_saif_text(gls_server* s, uint64_t job, char* buf, size_t cap, size_t* len) {
  API_BEGIN
  if (!s || !len) throw GlsError(GLS_E_ARG, "null argument");
  copy_text(s->srv->saif(job), buf, cap, len);
  API_END
}

int gls_server_drain(gls_server* s) {
  API_BEGIN
  if (!s) throw GlsError(GLS_E_ARG, "null argument");
  s->srv->drain();
  API_END
}

int gls_server_stats_get(gls_server* s, gls_server_stats* st) {
  API_BEGIN
  if (!s || !st) throw GlsError(GLS_E_ARG, "null argument");
  ServerStats x = s->srv->stats();
  st->submitted = x.submitted;
  st->done = x.done;
  st->failed = x.failed;
  st->events = x.events;
  st->p50_ms = x.p50_ms;
  st->p99_ms = x.p99_ms;
  st->mean_ms = x.mean_ms;
  st->wall_s = x.wall_s;
  st->events_per_s = x.events_per_s;
  st->peak_running = x.peak_running;
  st->resident_programs = x.resident_programs;
  st->peak_bytes = x.peak_bytes;
  st->cancelled = x.cancelled;
  st->arena_hits = x.arena_hits;
  st->arena_misses = x.arena_misses;
  API_END
}

void gls_server_destroy(gls_server* s) { delete s; }

}  // extern "C"
