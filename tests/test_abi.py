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
