"""Writes tests/golden/oracle_golden.json: the oracle's digests (gate events,
waveform hash, SAIF sha256) for the named generated workloads.  Run from the
repository root after building:  python tests/golden/make_golden.py"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "gate-level-simulator_amd"))

import golden_cases  # noqa: E402
import oracle_bind  # noqa: E402

out = {}
for name, c in golden_cases.CASES.items():
    with tempfile.TemporaryDirectory() as d:
        paths = golden_cases.write_inputs(name, d)
        r = oracle_bind.run(*paths, c["window"][0], c["window"][1], d)
        out[name] = {"n_events": r["n_events"], "waves_hash": "%016x" % r["waves_hash"],
                     "saif_sha256": oracle_bind.sha256(r["saif"]), "slow": bool(c.get("slow"))}
        print(name, out[name], "%.1fs" % r["sim_seconds"], flush=True)
with open(os.path.join(HERE, "oracle_golden.json"), "w") as f:
    json.dump(out, f, indent=1, sort_keys=True)
