"""ctypes binding of the oracle (oracle/cpu_ref.cpp) for the tests."""
import ctypes
import hashlib
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libglsoracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise OSError("oracle not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(ORACLE_SO)
        u64p = ctypes.POINTER(ctypes.c_ulonglong)
        _lib.oracle_run.restype = ctypes.c_int
        _lib.oracle_run.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_ulonglong] * 2 + \
            [ctypes.c_char_p] * 2 + [ctypes.c_ulonglong, u64p, u64p, ctypes.POINTER(ctypes.c_double)]
        _lib.oracle_calc.restype = ctypes.c_int
        _lib.oracle_calc.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        _lib.oracle_check_delay.restype = ctypes.c_int
        _lib.oracle_check_delay.argtypes = [ctypes.c_int, ctypes.c_int]
    return _lib


def run(vlib, gv, sdf, vcd, start, end, tmpdir, waves=False, max_events=0):
    """Run the oracle on files; returns dict(saif, n_events, waves_hash, waves?)."""
    saif = os.path.join(tmpdir, "oracle.saif")
    wv = os.path.join(tmpdir, "oracle.waves") if waves else None
    n, h, secs = ctypes.c_ulonglong(), ctypes.c_ulonglong(), ctypes.c_double()
    lib().oracle_run(vlib.encode(), gv.encode(), sdf.encode(), vcd.encode(), start, end,
                     saif.encode(), wv.encode() if wv else None, max_events,
                     ctypes.byref(n), ctypes.byref(h), ctypes.byref(secs))
    out = {"saif": open(saif).read(), "n_events": n.value, "waves_hash": h.value,
           "sim_seconds": secs.value}
    if waves:
        out["waves"] = {}
        for line in open(wv):
            parts = line.split()
            out["waves"][parts[0]] = [tuple(int(x) for x in p.split(":")) for p in parts[1:]]
    return out


def calc(name, vals):
    arr = (ctypes.c_int * len(vals))(*vals)
    return lib().oracle_calc(name.encode(), arr, len(vals))


def check_delay(pre, cur):
    return lib().oracle_check_delay(pre, cur)


def sha256(text):
    return hashlib.sha256(text.encode()).hexdigest()
