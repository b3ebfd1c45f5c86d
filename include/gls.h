/* libgls — C ABI of the MI355X gate-level waveform simulator.
 *
 * The reference exposes no library interface: its boundary is the pair of
 * command-line tools
 *     process  vlib gv sdf inter.dat               (src/process.cpp:178-194)
 *     simulate inter.dat vcd dumpon dumpoff saif    (src/simulate.cu:1786-1812)
 * with the kernels called straight from main (src/simulate.cu:2027-3254).
 * Each entry point below replaces one step of those tools; the file-level
 * mapping is given per function.  All functions return GLS_OK (0) or a
 * negative GLS_E_* status (gls_status.h) and never exit the process; the
 * message of the last failure on the calling thread is gls_last_error().
 *
 * Handles are opaque.  All device memory is owned by the library.
 */
#ifndef GLS_H
#define GLS_H

#include <stddef.h>
#include <stdint.h>

#include "gls_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gls_design gls_design; /* parsed + levelized netlist */
typedef struct gls_stim gls_stim;     /* stimulus waveforms */
typedef struct gls_ctx gls_ctx;       /* one device */
typedef struct gls_run gls_run;       /* one prepared / finished simulation */

/* Message of the last failing call on this thread ("" if none). */
const char* gls_last_error(void);
/* Library version string. */
const char* gls_version(void);

/* ------------------------------------------------------------ designs --- */
/* Parse vlib + gv + SDF (PreProcess::parse_vlib/parse_gv/parse_sdf,
 * src/PreProcess.cpp:42,650,1276) — the first half of `process`. */
int gls_design_from_files(const char* vlib_path, const char* gv_path, const char* sdf_path,
                          gls_design** out);
/* Same, from in-memory texts. */
int gls_design_from_texts(const char* vlib_text, const char* gv_text, const char* sdf_text,
                          gls_design** out);
/* Same with an explicit thread count for the gv and SDF readers (0: up to
 * 16; 1: one thread; texts under 8 MiB are read on one thread).  The design
 * does not depend on it. */
int gls_design_from_texts_threads(const char* vlib_text, const char* gv_text,
                                  const char* sdf_text, int threads, gls_design** out);
/* 64-bit digest of the parsed design: nets and their flags, instances with
 * their pins and every delay bit for bit.  Equal digests simulate alike. */
int gls_design_digest(const gls_design* d, uint64_t* digest);
/* Seeded synthetic netlist of one benchmark configuration (0, 1, 2; 3 is
 * configuration 1 with the 5- and 6-input cells GEN_AO221_D1 and
 * GEN_AO222_D1 in the mix) with the generic cell library;
 * n_cells/n_levels/n_pis override when > 0. */
int gls_design_generate(int config, uint64_t seed, int n_cells, int n_levels, int n_pis,
                        gls_design** out);
/* inter.dat writer / reader: the hand-off between `process` and `simulate`
 * (src/process.cpp:581-608 writes the Boost text archive of `inter`,
 * src/simulate.cu:1835-1838 reads it).  The file holds that object in
 * text-archive syntax and member order, then the library, netlist and
 * levelized program, which the reader loads without parsing again.  Files of
 * the earlier three-text layout ("GLS-INTER 1") are still read. */
int gls_design_save(const gls_design* d, const char* inter_path);
int gls_design_load(const char* inter_path, gls_design** out);
/* Write the design as vlib / gv / SDF text files (generator output). */
int gls_design_write(const gls_design* d, const char* vlib_path, const char* gv_path,
                     const char* sdf_path);
typedef struct {
  uint64_t n_instances, n_nets, n_inputs; /* netlist */
  uint64_t n_gates, n_levels, n_pins;     /* compiled program (after levelization) */
  uint64_t program_digest;                /* digest of the compiled program (gates, levels,
                                             truth tables, delays) */
} gls_design_info;
/* Counts; levelizes against the design's declared inputs.  The compiler
 * runs on up to 16 host threads (GLS_HOST_THREADS caps them, as it caps the
 * readers); its program does not depend on the count. */
int gls_design_info_get(const gls_design* d, gls_design_info* info);
void gls_design_free(gls_design* d);

/* ---------------------------------------------------------- stimulus --- */
/* Parse a VCD (VCDParser::init/parse/process, src/VCDParser.cpp:244-717).
 * A large body is read in parallel stretches cut at line-initial timestamps,
 * as the reference's threaded parser cuts it (src/VCDParser.cpp:668-683). */
int gls_stim_from_vcd(const char* vcd_path, gls_stim** out);
/* Same with an explicit thread count (0: up to 16; 1: one thread); the
 * result does not depend on it. */
int gls_stim_from_vcd_threads(const char* vcd_path, int threads, gls_stim** out);
/* Seeded Poisson stimulus for the design's primary inputs: initial value
 * U{0,1} at t = 0, mean toggle gap `mean_gap` ps, toggles in (0, t_end],
 * a toggle goes to X with probability x_prob. */
int gls_stim_generate(const gls_design* d, uint64_t seed, uint32_t t_end, double mean_gap,
                      double x_prob, gls_stim** out);
int gls_stim_write_vcd(const gls_stim* s, const char* vcd_path);
/* Number of signals and total events. */
int gls_stim_info(const gls_stim* s, uint64_t* n_signals, uint64_t* n_events);
void gls_stim_free(gls_stim* s);

/* ------------------------------------------------------------ device --- */
/* A context keeps the device buffers of finished runs cached for its later
   runs and servers; they are freed when an allocation fails and by
   gls_ctx_destroy (a server's default budget counts them as free).  Runs
   may be freed before or after their context; servers must be destroyed
   before it. */
int gls_ctx_create(int device, gls_ctx** out);
void gls_ctx_destroy(gls_ctx* c);
/* Free and total device memory of the context's device (hipMemGetInfo; free
 * includes the context's cached blocks), for callers that size batches to
 * fit. */
int gls_ctx_mem_info(gls_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes);

typedef struct {
  uint64_t pool_bytes;   /* event pool size, fixed (0: sized from the expected waveforms,
                            up to 85% of the free HBM, taking the context's cached pool
                            when there is one, and grown between levels as needed) */
  uint32_t seg_events;   /* target input events per time segment (0: path default) */
  uint32_t keep_waves;   /* 1: keep waveforms for gls_run_waveform */
  uint32_t path;         /* GLS_PATH_FUSED (default) or GLS_PATH_TWO_PASS */
  uint32_t level_blocks; /* workgroups per fused level launch (0: occupancy limit) */
  uint32_t no_prefetch;  /* 1: fused kernel loads the next tile only after storing the current */
  uint32_t graph;        /* 1: fused path only; after the first execution the whole level
                            sequence is captured into a hipGraph with no host round trip
                            between levels, and later executions replay it */
  uint32_t hbm_stacks;   /* 1: the fused kernel keeps its preemption stacks in a per-workgroup
                            HBM scratch instead of LDS (half the LDS, more resident workgroups) */
  uint32_t pool_grow;    /* 1: pool_bytes and ovf_words are only starting sizes; both grow as
                            with 0 */
  uint32_t ovf_words;    /* fused path: words of the stack buffer for tiles too long for LDS, per
                            level (0: starts at 2^24 and doubles when a level runs out, the
                            level is then run again); a fixed size that a level outgrows fails
                            the run with GLS_E_UNSUPPORTED */
  uint32_t saif_tc;      /* 1: SAIF records also carry (TC n), the number of 0<->1 transitions
                            in (dumpon, dumpoff].  0 (default): the reference's (T0)(T1)(TX)
                            records */
  uint32_t pad0;
  uint64_t chunk_ps;     /* fused path: time-chunk length in VCD time units.  0 (default):
                            automatic -- one chunk when the run fits HBM, else chunks sized to
                            the event pool; GLS_NO_CHUNKS: one chunk, a pool too small fails
                            with GLS_E_OOM; else a fixed length (above the largest gate
                            delay).  Results do not depend on it. */
} gls_sim_opts;

/* chunk_ps value: never split the window. */
#define GLS_NO_CHUNKS (~0ull)

/* Level-kernel variants: one fused LDS-tiled kernel per level, or the
 * two-pass form (evaluation into scratch, then compaction). */
#define GLS_PATH_FUSED 0
#define GLS_PATH_TWO_PASS 1

/* Upload program + stimulus into HBM (the per-level H2D copies of
 * src/simulate.cu:2434-2482 happen once here).  dumpon/dumpoff are the SAIF
 * window (src/simulate.cu:1811-1812). */
int gls_prepare(gls_ctx* c, const gls_design* d, const gls_stim* s, uint64_t dumpon,
                uint64_t dumpoff, const gls_sim_opts* opts, gls_run** out);
/* Run every level on the device (the level loop of src/simulate.cu:2027-3254
 * and the SAIF reduction of src/simulate.cu:1563-1784).  May be called again
 * on the same run; each call recomputes from the resident stimulus. */
int gls_run_execute(gls_run* r);
/* Allocation-free variant of gls_run_execute (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * The run must have been prepared (gls_prepare / gls_prepare_batch hold the
 * inputs and outputs in device memory); this call makes no hipMalloc,
 * hipFree or other allocating call.  Its per-level scratch (tile
 * descriptors and summaries, overflow stacks, time-chunk tails and windows)
 * comes from one static device array of GLS_STATIC_SCRATCH_BYTES per GPU,
 * held until the process ends; the event pool does not grow (time chunks
 * shrink to fit it, a level whose scratch does not fit the pool's free top
 * fails with GLS_E_OOM at the shortest chunk) and a run whose tiles or tails
 * exceed the array fails with GLS_E_RANGE.  Fused path only, without graph
 * capture, kept waveforms or HBM stacks (GLS_E_UNSUPPORTED).  Calls that
 * overlap on one GPU share the array and race: serialize them. */
#define GLS_STATIC_SCRATCH_BYTES (256ull << 20)
int gls_run_execute_static(gls_run* r);
/* prepare + execute */
int gls_simulate(gls_ctx* c, const gls_design* d, const gls_stim* s, uint64_t dumpon,
                 uint64_t dumpoff, const gls_sim_opts* opts, gls_run** out);

typedef struct {
  uint64_t n_events;   /* gate events: distinct input-event times per output */
  uint64_t in_words;   /* input waveform words read by gates */
  uint64_t out_words;  /* output waveform words written */
  uint64_t alg_bytes;  /* algorithmic HBM bytes of the level kernels */
  uint64_t n_nets, n_gates, n_pins;
  double kernel_ms;    /* device time of the level loop */
  double wall_ms;      /* host wall time of the level loop */
  double eval_ms;      /* summed device time of the gate-evaluation kernel */
  uint64_t eval_launches;
  uint64_t level_blocks; /* workgroups per fused level launch */
  uint64_t pool_bytes;      /* event pool capacity after the run */
  uint64_t pool_used_bytes; /* pool bytes holding waveforms at the end of the run */
  uint64_t tiles;           /* fused path: tiles walked (a re-run level counts twice) */
  uint64_t tiles_hbm;       /* of them too long for LDS (inputs read from HBM) */
  uint64_t chunks;          /* time chunks of the run (1: the whole window at once) */
  uint64_t chunk_retries;   /* chunks run again shorter because they did not fit the pool */
  uint64_t device_allocs;   /* device allocations made by the last execution */
} gls_run_stats;
int gls_run_stats_get(const gls_run* r, gls_run_stats* st);
/* SAIF writer (DumpSaif, src/simulate.cu:113-481 / src/simulate.cpp:37-380). */
int gls_run_write_saif(const gls_run* r, const char* saif_path);
/* SAIF text into a caller buffer; *len receives the size (call with
 * buf == NULL to query). */
int gls_run_saif_text(const gls_run* r, char* buf, size_t cap, size_t* len);
/* T0/T1/TX durations of one net. */
int gls_run_net_saif(const gls_run* r, const char* net, uint64_t t0t1tx[3]);
/* Toggle count of one net: 0<->1 transitions in (dumpon, dumpoff] (no counterpart in the
 * reference, whose SAIF has no TC field; needs the saif_tc option). */
int gls_run_net_toggles(const gls_run* r, const char* net, uint64_t* tc);
/* Waveform of one net as (time << 2 | value) words (needs keep_waves);
 * GLS_E_RANGE when a time is 2^30 or more. */
int gls_run_waveform(const gls_run* r, const char* net, uint32_t* words, size_t cap, size_t* n);
/* Same with 64-bit words, for any time. */
int gls_run_waveform64(const gls_run* r, const char* net, uint64_t* words, size_t cap, size_t* n);
/* Order-independent 64-bit hash over every gate output's waveform (needs
 * keep_waves); each event enters as the low 32 bits of (time << 2 | value). */
int gls_run_waves_hash(const gls_run* r, uint64_t* hash);
void gls_run_free(gls_run* r);

/* ------------------------------------------------------------ batches --- */
/* Batched form (no counterpart in the reference, which simulates one
 * stimulus per process): n instances — netlists of any size and depth, each
 * with its own stimulus and SAIF window — prepared together and evaluated
 * level by level by the same kernel launches (level l of every instance in
 * one launch).  Results per instance are bit-identical to a plain run of that
 * instance.  Fused path only when n > 1.  The single-run accessors above
 * address instance 0. */
int gls_prepare_batch(gls_ctx* c, size_t n, const gls_design* const* designs,
                      const gls_stim* const* stims, const uint64_t* dumpon,
                      const uint64_t* dumpoff, const gls_sim_opts* opts, gls_run** out);
int gls_run_instances(const gls_run* r, size_t* n);
int gls_run_instance_saif_text(const gls_run* r, size_t i, char* buf, size_t cap, size_t* len);
int gls_run_instance_waves_hash(const gls_run* r, size_t i, uint64_t* hash);
/* Gate events of instance i in the last execution (the run's total for a
 * one-instance run). */
int gls_run_instance_events(const gls_run* r, size_t i, uint64_t* n);

/* Autotuner: executes the prepared run once per candidate of the fused
 * kernel's events per segment (LDS tile size), workgroups per launch and
 * tile staging, keeps the fastest for later executions, and leaves the run
 * executed.  The report lists every trial. */
#define GLS_TUNE_MAX 16
typedef struct {
  uint32_t seg_events, level_blocks, prefetch, n_tried;
  double best_ms;
  uint32_t tried_seg[GLS_TUNE_MAX], tried_blocks[GLS_TUNE_MAX], tried_prefetch[GLS_TUNE_MAX];
  double tried_ms[GLS_TUNE_MAX];
} gls_tune_report;
int gls_run_autotune(gls_run* r, gls_tune_report* rep);

/* ------------------------------------------------------------ serving --- */
/* Serving loop (no counterpart in the reference): a request queue in front
 * of one device.  Jobs are admitted by their estimated device-memory
 * footprint as memory and slots free up (continuous batching: a job that
 * arrives while others run starts as soon as it fits), each running job has
 * its own stream so concurrent jobs overlap on the device, and compiled
 * netlists stay resident in HBM for later jobs.  Submit and wait are
 * thread-safe. */
typedef struct gls_server gls_server;
typedef struct {
  uint64_t device_budget; /* HBM bytes the server may use (0: 90% of free) */
  uint32_t max_running;   /* jobs on the device at once (0: 4) */
  uint32_t seg_events;    /* fused-kernel events per segment (0: default) */
  double pool_factor;     /* scales a job's event-pool estimate (0: 1) */
  uint32_t max_batch;     /* queued jobs admitted together run as one batch of up
                             to this many (level l of all in one launch); 0, 1: none */
  uint32_t reserved0;
} gls_server_opts;
typedef struct {
  int status;             /* 1 queued, 2 running, 0 done, < 0 GLS_E_* failure */
  double queue_ms, run_ms, latency_ms;
  uint64_t n_events, device_bytes;
  uint32_t attempts, resident, overlap; /* resident: netlist was already in HBM;
                                           overlap: jobs running when it started */
} gls_job_info;
typedef struct {
  uint64_t submitted, done, failed, events;
  double p50_ms, p99_ms, mean_ms, wall_s, events_per_s;
  uint32_t peak_running, resident_programs;
  uint64_t peak_bytes;
  uint64_t cancelled;     /* failed jobs that were cancelled or missed their deadline */
  uint64_t arena_hits;    /* device buffers taken from the context's cache */
  uint64_t arena_misses;  /* device buffers newly allocated */
} gls_server_stats;
/* The server uses the context's program cache and device-buffer cache:
   destroy it before the context. */
int gls_server_create(gls_ctx* c, const gls_server_opts* o, gls_server** out);
/* The server copies the stimulus; the design must outlive the server. */
int gls_server_submit(gls_server* s, const gls_design* d, const gls_stim* st, uint64_t dumpon,
                      uint64_t dumpoff, uint64_t* job);
/* As gls_server_submit, with a deadline: a job not finished deadline_s
 * seconds after submission ends with GLS_E_CANCELLED (checked while it is
 * queued and before each level while it runs; deadline_s <= 0: none). */
int gls_server_submit_deadline(gls_server* s, const gls_design* d, const gls_stim* st,
                               uint64_t dumpon, uint64_t dumpoff, double deadline_s,
                               uint64_t* job);
/* Cancel a job: a queued job leaves the queue at once, a running job stops
 * before its next level (its device memory is freed when it has stopped),
 * a finished job keeps its result.  *status (may be NULL) is the job's
 * status after the call. */
int gls_server_cancel(gls_server* s, uint64_t job, int* status);
/* timeout_s < 0 waits forever; *status as in gls_job_info. */
int gls_server_wait(gls_server* s, uint64_t job, double timeout_s, int* status);
int gls_server_job_info(gls_server* s, uint64_t job, gls_job_info* info);
int gls_server_job_saif_text(gls_server* s, uint64_t job, char* buf, size_t cap, size_t* len);
int gls_server_drain(gls_server* s);
int gls_server_stats_get(gls_server* s, gls_server_stats* st);
/* Finishes every submitted job, then frees the server. */
void gls_server_destroy(gls_server* s);

#ifdef __cplusplus
}
#endif
#endif
