"""Named workloads whose oracle digests are committed in
tests/golden/oracle_golden.json.  Inputs are regenerated with the seeded
generator (deterministic across hosts), so only digests are stored."""
import os

import gls

CASES = {
    # config index, cells, levels, PIs, window end, mean gap, SAIF window
    "c0_small": dict(config=0, cells=400, levels=20, pis=64, t_end=200000, gap=1000.0,
                     window=(0, 200001)),
    "c1_small": dict(config=1, cells=6000, levels=40, pis=128, t_end=100000, gap=500.0,
                     window=(10000, 100001)),
    "c0_full": dict(config=0, cells=0, levels=0, pis=0, t_end=2000000, gap=1000.0,
                    window=(0, 2000001), slow=True),
}


def seed(name):
    return 0x5EED0000 + CASES[name]["config"]


def make(name):
    c = CASES[name]
    d = gls.Design.generate(c["config"], seed(name), c["cells"], c["levels"], c["pis"])
    s = gls.Stimulus.generate(d, seed(name), c["t_end"], c["gap"], 0.01)
    return d, s


def write_inputs(name, directory):
    d, s = make(name)
    paths = [os.path.join(directory, n) for n in ("lib.vlib", "netlist.gv", "netlist.sdf", "stim.vcd")]
    d.write(*paths[:3])
    s.write_vcd(paths[3])
    return paths
