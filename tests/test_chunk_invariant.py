"""The invariant that makes a time-chunked simulation exact, on the oracle.

Under the preemption rule (a candidate is kept iff its output time is
strictly earlier than that of every later candidate) an input event at time
t only makes candidates with output time >= t, so every net's entries at
times below a frontier F depend only on the stimulus below F.  A chunked run
(all levels over [F_k, F_k+1), carrying each gate's state) relies on exactly
this.  Checked here by running the oracle on the full stimulus and on the
stimulus cut at F: every net's waveform below F, and the SAIF of the window
[0, F], must be identical.
"""
import os

import pytest

import golden_cases
import oracle_bind


def cut_vcd(text, frontier):
    """The VCD with every value change at time >= frontier removed."""
    head, body = text.split("$enddefinitions $end", 1)
    out = [head + "$enddefinitions $end"]
    for line in body.split("\n"):
        if line.startswith("#") and int(line[1:]) >= frontier:
            break
        out.append(line)
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("name,frontier", [("c0_small", 73_457), ("c0_small", 150_001),
                                           ("c1_small", 41_000)])
def test_entries_below_frontier_are_final(tmp_path, name, frontier):
    vlib, gv, sdf, vcd = golden_cases.write_inputs(name, str(tmp_path))
    cut = str(tmp_path / "cut.vcd")
    with open(cut, "w") as f:
        f.write(cut_vcd(open(vcd).read(), frontier))
    full_dir, cut_dir = tmp_path / "full", tmp_path / "cut"
    os.makedirs(full_dir)
    os.makedirs(cut_dir)
    full = oracle_bind.run(vlib, gv, sdf, vcd, 0, frontier, str(full_dir), waves=True)
    part = oracle_bind.run(vlib, gv, sdf, cut, 0, frontier, str(cut_dir), waves=True)
    assert full["waves"].keys() == part["waves"].keys()
    changed = 0
    for net, w in full["waves"].items():
        below = [e for e in w if e[0] < frontier]
        assert [e for e in part["waves"][net] if e[0] < frontier] == below, net
        changed += len(below) != len(w)
    assert changed > 0  # the cut did remove later entries
    assert part["saif"] == full["saif"]
