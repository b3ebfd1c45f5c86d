"""The multi-threaded host code (parallel gv / SDF / VCD readers, generator,
parallel netlist compiler, inter.dat writer and reader) built with the
address + undefined-behaviour sanitizers and with the thread sanitizer, and
run on a netlist that takes every parallel path: no report, and every
threaded result equals its one-thread result."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gate-level-simulator_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_sanitizers():
    r = subprocess.run(["make", "-C", PKG, "-j2", "sancheck"], capture_output=True, text=True,
                       timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert out.count("hostcheck passed") == 2, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "WARNING: ThreadSanitizer" not in out
    assert "runtime error:" not in out
