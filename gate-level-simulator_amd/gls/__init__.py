"""Python binding of libgls (ctypes over the C ABI in include/gls.h).

The simulator itself is native: parsing, levelization and the HIP kernels
all live in libgls.so.  This module only marshals handles and buffers so the
tests and the benchmark can drive the same entry points a C/C++ caller uses.
Loading fails loudly when the library has not been built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "build", "libgls.so")

GLS_OK = 0
STATUS = {
    -1: "GLS_E_ARG", -2: "GLS_E_FORMAT", -3: "GLS_E_OOM", -4: "GLS_E_DEVICE",
    -5: "GLS_E_UNSUPPORTED", -6: "GLS_E_CANCELLED", -7: "GLS_E_RANGE",
    -8: "GLS_E_STATE", -9: "GLS_E_IO", -10: "GLS_E_CYCLE",
}


class GlsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code
        self.status = STATUS.get(code, str(code))


class DesignInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("n_instances", "n_nets", "n_inputs", "n_gates", "n_levels", "n_pins",
                 "program_digest")]


class SimOpts(ctypes.Structure):
    _fields_ = [("pool_bytes", ctypes.c_uint64), ("seg_events", ctypes.c_uint32),
                ("keep_waves", ctypes.c_uint32), ("path", ctypes.c_uint32),
                ("level_blocks", ctypes.c_uint32), ("no_prefetch", ctypes.c_uint32),
                ("graph", ctypes.c_uint32), ("hbm_stacks", ctypes.c_uint32),
                ("pool_grow", ctypes.c_uint32), ("ovf_words", ctypes.c_uint32),
                ("saif_tc", ctypes.c_uint32), ("pad0", ctypes.c_uint32),
                ("chunk_ps", ctypes.c_uint64)]


NO_CHUNKS = (1 << 64) - 1  # chunk_ps: one time chunk, never split
PATH_FUSED = 0
PATH_TWO_PASS = 1


class RunStats(ctypes.Structure):
    _fields_ = [("n_events", ctypes.c_uint64), ("in_words", ctypes.c_uint64),
                ("out_words", ctypes.c_uint64), ("alg_bytes", ctypes.c_uint64),
                ("n_nets", ctypes.c_uint64), ("n_gates", ctypes.c_uint64),
                ("n_pins", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double), ("eval_ms", ctypes.c_double),
                ("eval_launches", ctypes.c_uint64), ("level_blocks", ctypes.c_uint64),
                ("pool_bytes", ctypes.c_uint64), ("pool_used_bytes", ctypes.c_uint64),
                ("tiles", ctypes.c_uint64), ("tiles_hbm", ctypes.c_uint64),
                ("chunks", ctypes.c_uint64), ("chunk_retries", ctypes.c_uint64),
                ("device_allocs", ctypes.c_uint64)]


TUNE_MAX = 16


class TuneReport(ctypes.Structure):
    _fields_ = [("seg_events", ctypes.c_uint32), ("level_blocks", ctypes.c_uint32),
                ("prefetch", ctypes.c_uint32), ("n_tried", ctypes.c_uint32),
                ("best_ms", ctypes.c_double),
                ("tried_seg", ctypes.c_uint32 * TUNE_MAX),
                ("tried_blocks", ctypes.c_uint32 * TUNE_MAX),
                ("tried_prefetch", ctypes.c_uint32 * TUNE_MAX),
                ("tried_ms", ctypes.c_double * TUNE_MAX)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("device_budget", ctypes.c_uint64), ("max_running", ctypes.c_uint32),
                ("seg_events", ctypes.c_uint32), ("pool_factor", ctypes.c_double),
                ("max_batch", ctypes.c_uint32), ("reserved0", ctypes.c_uint32)]


class JobInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("n_events", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("attempts", ctypes.c_uint32), ("resident", ctypes.c_uint32),
                ("overlap", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("done", ctypes.c_uint64),
                ("failed", ctypes.c_uint64), ("events", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("wall_s", ctypes.c_double),
                ("events_per_s", ctypes.c_double), ("peak_running", ctypes.c_uint32),
                ("resident_programs", ctypes.c_uint32), ("peak_bytes", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("arena_hits", ctypes.c_uint64),
                ("arena_misses", ctypes.c_uint64)]


# Every symbol include/gls.h declares: (name, restype, argtypes)
_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_S = ctypes.c_char_p
_U64 = ctypes.c_uint64
_I = ctypes.c_int
SIGNATURES = [
    ("gls_last_error", _S, []),
    ("gls_version", _S, []),
    ("gls_design_from_files", _I, [_S, _S, _S, _PP]),
    ("gls_design_from_texts", _I, [_S, _S, _S, _PP]),
    ("gls_design_from_texts_threads", _I, [_S, _S, _S, ctypes.c_int, _PP]),
    ("gls_design_digest", _I, [_P, ctypes.POINTER(_U64)]),
    ("gls_design_generate", _I, [_I, _U64, _I, _I, _I, _PP]),
    ("gls_design_save", _I, [_P, _S]),
    ("gls_design_load", _I, [_S, _PP]),
    ("gls_design_write", _I, [_P, _S, _S, _S]),
    ("gls_design_info_get", _I, [_P, ctypes.POINTER(DesignInfo)]),
    ("gls_design_free", None, [_P]),
    ("gls_stim_from_vcd", _I, [_S, _PP]),
    ("gls_stim_from_vcd_threads", _I, [_S, ctypes.c_int, _PP]),
    ("gls_run_net_toggles", _I, [_P, _S, ctypes.POINTER(_U64)]),
    ("gls_stim_generate", _I, [_P, _U64, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, _PP]),
    ("gls_stim_write_vcd", _I, [_P, _S]),
    ("gls_stim_info", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    ("gls_stim_free", None, [_P]),
    ("gls_ctx_create", _I, [_I, _PP]),
    ("gls_ctx_mem_info", _I, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    ("gls_ctx_destroy", None, [_P]),
    ("gls_prepare", _I, [_P, _P, _P, _U64, _U64, ctypes.POINTER(SimOpts), _PP]),
    ("gls_run_execute", _I, [_P]),
    ("gls_run_execute_static", _I, [_P]),
    ("gls_simulate", _I, [_P, _P, _P, _U64, _U64, ctypes.POINTER(SimOpts), _PP]),
    ("gls_run_stats_get", _I, [_P, ctypes.POINTER(RunStats)]),
    ("gls_run_write_saif", _I, [_P, _S]),
    ("gls_run_saif_text", _I, [_P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("gls_run_net_saif", _I, [_P, _S, ctypes.POINTER(_U64)]),
    ("gls_run_waveform", _I, [_P, _S, ctypes.POINTER(ctypes.c_uint32), ctypes.c_size_t,
                              ctypes.POINTER(ctypes.c_size_t)]),
    ("gls_run_waveform64", _I, [_P, _S, ctypes.POINTER(_U64), ctypes.c_size_t,
                                ctypes.POINTER(ctypes.c_size_t)]),
    ("gls_run_waves_hash", _I, [_P, ctypes.POINTER(_U64)]),
    ("gls_run_free", None, [_P]),
    ("gls_prepare_batch", _I, [_P, ctypes.c_size_t, _PP, _PP, ctypes.POINTER(_U64),
                               ctypes.POINTER(_U64), ctypes.POINTER(SimOpts), _PP]),
    ("gls_run_instances", _I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    ("gls_run_instance_events", _I, [_P, ctypes.c_size_t, ctypes.POINTER(_U64)]),
    ("gls_run_instance_saif_text", _I, [_P, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]),
    ("gls_run_instance_waves_hash", _I, [_P, ctypes.c_size_t, ctypes.POINTER(_U64)]),
    ("gls_run_autotune", _I, [_P, ctypes.POINTER(TuneReport)]),
    ("gls_server_create", _I, [_P, ctypes.POINTER(ServerOpts), _PP]),
    ("gls_server_submit", _I, [_P, _P, _P, _U64, _U64, ctypes.POINTER(_U64)]),
    ("gls_server_submit_deadline", _I, [_P, _P, _P, _U64, _U64, ctypes.c_double,
                                        ctypes.POINTER(_U64)]),
    ("gls_server_cancel", _I, [_P, _U64, ctypes.POINTER(ctypes.c_int)]),
    ("gls_server_wait", _I, [_P, _U64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    ("gls_server_job_info", _I, [_P, _U64, ctypes.POINTER(JobInfo)]),
    ("gls_server_job_saif_text", _I, [_P, _U64, ctypes.c_char_p, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t)]),
    ("gls_server_drain", _I, [_P]),
    ("gls_server_stats_get", _I, [_P, ctypes.POINTER(ServerStats)]),
    ("gls_server_destroy", None, [_P]),
]

_lib = None


def load(path=None):
    """Load libgls.so (raises OSError if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"libgls.so not built at {p}; run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _b(s):
    return None if s is None else (s.encode() if isinstance(s, str) else s)


def _check(rc):
    if rc != GLS_OK:
        raise GlsError(rc, load().gls_last_error().decode(errors="replace"))


class _Handle:
    _free = None

    def __init__(self, ptr):
        self._ptr = ptr

    @property
    def ptr(self):
        if not self._ptr:
            raise GlsError(-8, "handle already freed")
        return self._ptr

    def close(self):
        if self._ptr and self._free:
            getattr(load(), self._free)(self._ptr)
        self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Design(_Handle):
    _free = "gls_design_free"

    @classmethod
    def from_files(cls, vlib, gv, sdf):
        h = ctypes.c_void_p()
        _check(load().gls_design_from_files(_b(vlib), _b(gv), _b(sdf), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_texts(cls, vlib, gv, sdf, threads=None):
        """threads: None = the library's choice (parallel for large texts),
        1 = one thread, n = up to n threads; the design does not depend on it."""
        h = ctypes.c_void_p()
        if threads is None:
            _check(load().gls_design_from_texts(_b(vlib), _b(gv), _b(sdf), ctypes.byref(h)))
        else:
            _check(load().gls_design_from_texts_threads(_b(vlib), _b(gv), _b(sdf), int(threads),
                                                        ctypes.byref(h)))
        return cls(h.value)

    def digest(self):
        """64-bit digest of the parsed netlist (nets, pins, delays bit for bit)."""
        v = _U64()
        _check(load().gls_design_digest(self.ptr, ctypes.byref(v)))
        return v.value

    @classmethod
    def generate(cls, config, seed=0, n_cells=0, n_levels=0, n_pis=0):
        h = ctypes.c_void_p()
        _check(load().gls_design_generate(config, seed, n_cells, n_levels, n_pis, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def load_inter(cls, path):
        h = ctypes.c_void_p()
        _check(load().gls_design_load(_b(path), ctypes.byref(h)))
        return cls(h.value)

    def save_inter(self, path):
        _check(load().gls_design_save(self.ptr, _b(path)))

    def write(self, vlib, gv, sdf):
        _check(load().gls_design_write(self.ptr, _b(vlib), _b(gv), _b(sdf)))

    def info(self):
        i = DesignInfo()
        _check(load().gls_design_info_get(self.ptr, ctypes.byref(i)))
        return {n: getattr(i, n) for n, _ in DesignInfo._fields_}


class Stimulus(_Handle):
    _free = "gls_stim_free"

    @classmethod
    def from_vcd(cls, path, threads=None):
        """threads: None = the library's choice (parallel for a large body),
        1 = one thread, n = up to n threads; the result does not depend on it."""
        h = ctypes.c_void_p()
        if threads is None:
            _check(load().gls_stim_from_vcd(_b(path), ctypes.byref(h)))
        else:
            _check(load().gls_stim_from_vcd_threads(_b(path), int(threads), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, design, seed, t_end, mean_gap, x_prob=0.01):
        h = ctypes.c_void_p()
        _check(load().gls_stim_generate(design.ptr, seed, t_end, mean_gap, x_prob, ctypes.byref(h)))
        return cls(h.value)

    def write_vcd(self, path):
        _check(load().gls_stim_write_vcd(self.ptr, _b(path)))

    def info(self):
        a, b = _U64(), _U64()
        _check(load().gls_stim_info(self.ptr, ctypes.byref(a), ctypes.byref(b)))
        return {"n_signals": a.value, "n_events": b.value}


class Run(_Handle):
    _free = "gls_run_free"

    def execute(self):
        _check(load().gls_run_execute(self.ptr))
        return self

    def execute_static(self):
        """gls_run_execute_static: the same execution with no device
        allocation (scratch from the library's static device array)."""
        _check(load().gls_run_execute_static(self.ptr))
        return self

    def stats(self):
        s = RunStats()
        _check(load().gls_run_stats_get(self.ptr, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in RunStats._fields_}

    def write_saif(self, path):
        _check(load().gls_run_write_saif(self.ptr, _b(path)))

    def saif_text(self):
        n = ctypes.c_size_t()
        _check(load().gls_run_saif_text(self.ptr, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(load().gls_run_saif_text(self.ptr, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value].decode()

    def net_saif(self, net):
        v = (_U64 * 3)()
        _check(load().gls_run_net_saif(self.ptr, _b(net), v))
        return tuple(v)

    def net_toggles(self, net):
        """0<->1 transitions of a net in (dumpon, dumpoff]."""
        v = _U64()
        _check(load().gls_run_net_toggles(self.ptr, _b(net), ctypes.byref(v)))
        return v.value

    def waveform(self, net):
        n = ctypes.c_size_t()
        _check(load().gls_run_waveform(self.ptr, _b(net), None, 0, ctypes.byref(n)))
        buf = (ctypes.c_uint32 * max(1, n.value))()
        _check(load().gls_run_waveform(self.ptr, _b(net), buf, n.value, ctypes.byref(n)))
        return [(w >> 2, w & 3) for w in buf[:n.value]]

    def waveform64(self, net):
        """Waveform with times of any size: [(time, value code)]."""
        n = ctypes.c_size_t()
        _check(load().gls_run_waveform64(self.ptr, _b(net), None, 0, ctypes.byref(n)))
        buf = (_U64 * max(1, n.value))()
        _check(load().gls_run_waveform64(self.ptr, _b(net), buf, n.value, ctypes.byref(n)))
        return [(w >> 2, w & 3) for w in buf[:n.value]]

    def waves_hash(self):
        h = _U64()
        _check(load().gls_run_waves_hash(self.ptr, ctypes.byref(h)))
        return h.value

    def instances(self):
        n = ctypes.c_size_t()
        _check(load().gls_run_instances(self.ptr, ctypes.byref(n)))
        return n.value

    def instance_saif_text(self, i):
        n = ctypes.c_size_t()
        _check(load().gls_run_instance_saif_text(self.ptr, i, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(load().gls_run_instance_saif_text(self.ptr, i, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value].decode()

    def instance_events(self, i):
        n = _U64()
        _check(load().gls_run_instance_events(self.ptr, i, ctypes.byref(n)))
        return n.value

    def instance_waves_hash(self, i):
        h = _U64()
        _check(load().gls_run_instance_waves_hash(self.ptr, i, ctypes.byref(h)))
        return h.value

    def autotune(self):
        r = TuneReport()
        _check(load().gls_run_autotune(self.ptr, ctypes.byref(r)))
        tried = [dict(seg_events=r.tried_seg[i], level_blocks=r.tried_blocks[i],
                      prefetch=r.tried_prefetch[i], ms=r.tried_ms[i]) for i in range(r.n_tried)]
        return dict(seg_events=r.seg_events, level_blocks=r.level_blocks, prefetch=r.prefetch,
                    best_ms=r.best_ms, tried=tried)


class Server(_Handle):
    """Serving loop over one device (see gls_server_* in include/gls.h)."""
    _free = "gls_server_destroy"

    def __init__(self, ctx, device_budget=0, max_running=4, seg_events=0, pool_factor=1.0,
                 max_batch=1):
        h = ctypes.c_void_p()
        o = ServerOpts(device_budget, max_running, seg_events, pool_factor, max_batch, 0)
        _check(load().gls_server_create(ctx.ptr, ctypes.byref(o), ctypes.byref(h)))
        self._keep = [ctx]
        super().__init__(h.value)

    def submit(self, design, stim, dumpon, dumpoff, deadline_s=None):
        """Queue a job; with deadline_s it fails with GLS_E_CANCELLED when not
        finished that many seconds after submission."""
        j = _U64()
        if deadline_s is None:
            _check(load().gls_server_submit(self.ptr, design.ptr, stim.ptr, dumpon, dumpoff,
                                            ctypes.byref(j)))
        else:
            _check(load().gls_server_submit_deadline(self.ptr, design.ptr, stim.ptr, dumpon,
                                                     dumpoff, float(deadline_s),
                                                     ctypes.byref(j)))
        self._keep.append(design)
        return j.value

    def cancel(self, job):
        """Cancel a job; returns its status afterwards (a finished job keeps
        its result, a queued one is GLS_E_CANCELLED at once, a running one
        stops before its next level)."""
        st = ctypes.c_int()
        _check(load().gls_server_cancel(self.ptr, job, ctypes.byref(st)))
        return st.value

    def wait(self, job, timeout_s=-1.0):
        st = ctypes.c_int()
        _check(load().gls_server_wait(self.ptr, job, timeout_s, ctypes.byref(st)))
        return st.value

    def info(self, job):
        i = JobInfo()
        _check(load().gls_server_job_info(self.ptr, job, ctypes.byref(i)))
        return {n: getattr(i, n) for n, _ in JobInfo._fields_}

    def saif_text(self, job):
        n = ctypes.c_size_t()
        _check(load().gls_server_job_saif_text(self.ptr, job, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(load().gls_server_job_saif_text(self.ptr, job, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value].decode()

    def drain(self):
        _check(load().gls_server_drain(self.ptr))

    def stats(self):
        s = ServerStats()
        _check(load().gls_server_stats_get(self.ptr, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in ServerStats._fields_}


class Context(_Handle):
    _free = "gls_ctx_destroy"

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(load().gls_ctx_create(device, ctypes.byref(h)))
        super().__init__(h.value)

    def mem_info(self):
        """(free, total) bytes of the device's memory."""
        f, t = _U64(), _U64()
        _check(load().gls_ctx_mem_info(self.ptr, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    @staticmethod
    def _opts(pool_bytes=0, seg_events=0, keep_waves=False, path=PATH_FUSED, level_blocks=0,
              prefetch=True, graph=False, hbm_stacks=False, pool_grow=False, ovf_words=0,
              saif_tc=False, chunk_ps=0):
        return SimOpts(pool_bytes, seg_events, 1 if keep_waves else 0, path, level_blocks,
                       0 if prefetch else 1, 1 if graph else 0, 1 if hbm_stacks else 0,
                       1 if pool_grow else 0, ovf_words, 1 if saif_tc else 0, 0, chunk_ps)

    def prepare(self, design, stim, dumpon, dumpoff, **opts):
        """Options: pool_bytes, seg_events, keep_waves, path, level_blocks,
        prefetch, graph, hbm_stacks, pool_grow, ovf_words, saif_tc, chunk_ps
        (gls_sim_opts)."""
        h = ctypes.c_void_p()
        o = self._opts(**opts)
        _check(load().gls_prepare(self.ptr, design.ptr, stim.ptr, dumpon, dumpoff, ctypes.byref(o),
                                  ctypes.byref(h)))
        return Run(h.value)

    def prepare_batch(self, designs, stims, dumpons, dumpoffs, **opts):
        n = len(designs)
        assert len(stims) == n and len(dumpons) == n and len(dumpoffs) == n
        D = (ctypes.c_void_p * n)(*[d.ptr for d in designs])
        S = (ctypes.c_void_p * n)(*[s.ptr for s in stims])
        on = (_U64 * n)(*dumpons)
        off = (_U64 * n)(*dumpoffs)
        h = ctypes.c_void_p()
        if opts.get("path", PATH_FUSED) != PATH_FUSED:
            raise ValueError("batches run on the fused path")
        o = self._opts(**opts)
        _check(load().gls_prepare_batch(self.ptr, n, D, S, on, off, ctypes.byref(o),
                                        ctypes.byref(h)))
        return Run(h.value)

    def simulate(self, design, stim, dumpon, dumpoff, **opts):
        h = ctypes.c_void_p()
        o = self._opts(**opts)
        _check(load().gls_simulate(self.ptr, design.ptr, stim.ptr, dumpon, dumpoff, ctypes.byref(o),
                                   ctypes.byref(h)))
        return Run(h.value)
