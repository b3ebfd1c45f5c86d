"""The C-ABI library loads and exports every symbol include/gls.h declares;
host-only entry points work without a GPU and report errors as status codes."""
import os
import re

import pytest

import gls

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gls.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gls_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = gls.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    bound = {n for n, _, _ in gls.SIGNATURES}
    assert set(declared_symbols()) == bound


def test_version_and_error_string():
    lib = gls.load()
    assert lib.gls_version().startswith(b"gls-mi355x")
    with pytest.raises(gls.GlsError) as e:
        gls.Design.from_files("/nonexistent.vlib", "/nonexistent.gv", None)
    assert e.value.status == "GLS_E_IO"
    assert "nonexistent" in str(e.value)


def test_no_device_is_an_error_not_a_crash():
    import conftest
    if conftest.gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(gls.GlsError) as e:
        gls.Context(0)
    assert e.value.status in ("GLS_E_DEVICE", "GLS_E_ARG")


def test_binding_keyword_options():
    import ctypes
    # every option of gls_sim_opts reaches the struct through each entry point
    import inspect
    fields = [n for n, _ in gls.SimOpts._fields_]
    assert fields == ["pool_bytes", "seg_events", "keep_waves", "path", "level_blocks",
                      "no_prefetch", "graph", "hbm_stacks", "pool_grow", "ovf_words", "saif_tc",
                      "pad0", "chunk_ps"]
    # the C layout: chunk_ps is 8-byte aligned after the eleven 4-byte fields
    assert gls.SimOpts.chunk_ps.offset == 56 and ctypes.sizeof(gls.SimOpts) == 64
    o = gls.Context._opts(pool_bytes=8, seg_events=2, keep_waves=True, path=gls.PATH_FUSED,
                          level_blocks=3, prefetch=False, graph=True, hbm_stacks=True,
                          pool_grow=True, ovf_words=77, saif_tc=True, chunk_ps=12345)
    assert (o.pool_bytes, o.seg_events, o.keep_waves, o.level_blocks, o.no_prefetch, o.graph,
            o.hbm_stacks, o.pool_grow, o.ovf_words, o.saif_tc, o.chunk_ps) == \
        (8, 2, 1, 3, 1, 1, 1, 1, 77, 1, 12345)
    # every keyword reaches the struct through each entry point (they share _opts)
    params = inspect.signature(gls.Context._opts).parameters
    for k in ("seg_events", "level_blocks", "prefetch", "graph", "hbm_stacks", "pool_grow",
              "ovf_words", "saif_tc", "chunk_ps"):
        assert k in params, k
    for fn in (gls.Context.prepare, gls.Context.simulate, gls.Context.prepare_batch):
        assert "opts" in inspect.signature(fn).parameters, fn.__name__


def test_null_handles_and_bad_arguments_are_status_codes():
    import ctypes
    lib = gls.load()
    st = ctypes.c_int()
    h = ctypes.c_uint64()
    assert lib.gls_server_cancel(None, 1, ctypes.byref(st)) == -1  # GLS_E_ARG
    assert lib.gls_server_submit_deadline(None, None, None, 0, 1, 1.0, ctypes.byref(h)) == -1
    assert lib.gls_design_digest(None, ctypes.byref(h)) == -1
    d = ctypes.c_void_p()
    assert lib.gls_design_from_texts_threads(b"", b"", None, -1, ctypes.byref(d)) == -1
    assert lib.gls_design_from_texts_threads(None, b"", None, 0, ctypes.byref(d)) == -1


def test_digests_follow_the_design():
    import golden_cases
    a, _ = golden_cases.make("c0_small")
    b, _ = golden_cases.make("c0_small")
    c = gls.Design.generate(0, 12345, 400, 20, 64)
    assert a.digest() == b.digest() != c.digest()
    ia, ib, ic = a.info(), b.info(), c.info()
    assert ia["program_digest"] == ib["program_digest"] != ic["program_digest"]
