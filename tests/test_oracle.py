"""Pins of the oracle (oracle/cpu_ref.cpp) on the CPU.

* truth tables: calc() against the Verilog primitive definitions over
  {0, 1, X} and against the udp_mux2 table of the reference's cell library
  (examples/GENERIC_STD_CELL.vlib:54-66);
* CheckDelay against its rule text;
* hand-computed known-answer waveforms, one rule per case;
* random single-gate cases against a closed-form restatement (suffix-min
  filter + dedup instead of the reference's pop/push loop);
* SAIF durations against the closed form "each entry holds its value until
  the next one";
* golden digests of generated workloads (tests/golden/oracle_golden.json).
"""
import itertools
import json
import os
import random

import pytest

import kat_cases
import oracle_bind
import spec_py

V0, VX, VZ, V1 = 0, 1, 2, 3
ALL3 = [V0, VX, V1]


def verilog_and(a, b):
    if V0 in (a, b):
        return V0
    return V1 if (a, b) == (V1, V1) else VX


def verilog_or(a, b):
    if V1 in (a, b):
        return V1
    return V0 if (a, b) == (V0, V0) else VX


def verilog_xor(a, b):
    if VX in (a, b):
        return VX
    return V1 if a != b else V0


def inv(v):
    return {V0: V1, V1: V0}.get(v, VX)


@pytest.mark.parametrize("name,fn", [
    ("and", verilog_and), ("or", verilog_or), ("xor", verilog_xor),
    ("nand", lambda a, b: inv(verilog_and(a, b))), ("nor", lambda a, b: inv(verilog_or(a, b))),
    ("xnor", lambda a, b: inv(verilog_xor(a, b))),
])
def test_two_input_primitives(name, fn):
    for a, b in itertools.product(ALL3, ALL3):
        assert oracle_bind.calc(name, [a, b]) == fn(a, b), (name, a, b)


@pytest.mark.parametrize("name,fn", [("and", verilog_and), ("or", verilog_or), ("xor", verilog_xor)])
def test_three_input_fold(name, fn):
    for a, b, c in itertools.product(ALL3, ALL3, ALL3):
        assert oracle_bind.calc(name, [a, b, c]) == fn(fn(a, b), c), (name, a, b, c)


def test_buf_not():
    assert [oracle_bind.calc("buf", [v]) for v in (V0, VX, VZ, V1)] == [V0, VX, VX, V1]
    assert [oracle_bind.calc("not", [v]) for v in (V0, VX, VZ, V1)] == [V1, VX, VX, V0]


def test_z_inputs_follow_reference_case_analysis():
    # Z acts as X on and/or inputs; xor of two Z is 0 (src/simulate.cpp:545-551)
    assert oracle_bind.calc("and", [VZ, V0]) == V0
    assert oracle_bind.calc("and", [VZ, V1]) == VX
    assert oracle_bind.calc("or", [VZ, V1]) == V1
    assert oracle_bind.calc("xor", [VZ, VZ]) == V0
    assert oracle_bind.calc("xor", [VZ, V0]) == VX


# udp_mux2 table rows of the reference's cell library (i0 i1 s : z); rows not
# matched give x (UDP default).  '?' is any of 0/1/x.
UDP_MUX2_ROWS = [("1", "?", "0", "1"), ("0", "?", "0", "0"), ("?", "1", "1", "1"),
                 ("?", "0", "1", "0"), ("0", "0", "x", "0"), ("1", "1", "x", "1")]
SYM = {"0": V0, "1": V1, "x": VX}


def udp_mux2(i0, i1, s):
    for r in UDP_MUX2_ROWS:
        if all(p == "?" or SYM[p] == v for p, v in zip(r[:3], (i0, i1, s))):
            return SYM[r[3]]
    return VX


def test_udp_mux2_matches_library_table():
    for i0, i1, s in itertools.product(ALL3, ALL3, ALL3):
        assert oracle_bind.calc("udp_mux2", [i0, i1, s]) == udp_mux2(i0, i1, s), (i0, i1, s)


def test_check_delay_rule():
    # CheckDelay's edge classes written out pair by pair (1 rise: leaving 0 or
    # reaching 1; 2 fall: leaving 1 or reaching 0; 3: X -> Z takes the larger
    # of rise and fall; 4: Z -> X the smaller; 0: no change), rows previous
    # value, columns current value, in the order 0, X, Z, 1
    table = {
        V0: {V0: 0, VX: 1, VZ: 1, V1: 1},
        VX: {V0: 2, VX: 0, VZ: 3, V1: 1},
        VZ: {V0: 2, VX: 4, VZ: 0, V1: 1},
        V1: {V0: 2, VX: 2, VZ: 2, V1: 0},
    }
    for pre, row in table.items():
        for cur, want in row.items():
            assert oracle_bind.check_delay(pre, cur) == want, (pre, cur)


@pytest.mark.parametrize("case", kat_cases.CASES, ids=[c["name"] for c in kat_cases.CASES])
def test_known_answers(case, tmp_path):
    paths = kat_cases.write_case(case, str(tmp_path))
    out = oracle_bind.run(*paths, 0, 1000, str(tmp_path), waves=True)
    for net, want in case["expect"].items():
        assert out["waves"][net] == want, net


CELL_SPEC = {  # generic-library cell, pins, edge-specific pins
    "INV": ("GEN_INV_D1", ["i"], "zn", []),
    "BUF": ("GEN_BUF_D1", ["i"], "z", []),
    "AND2": ("GEN_AND2_D1", ["a1", "a2"], "z", []),
    "XOR2": ("GEN_XOR2_D1", ["a1", "a2"], "z", ["a1", "a2"]),
    "MUX2": ("GEN_MUX2_D1", ["i0", "i1", "s"], "z", ["s"]),
    "AOI21": ("GEN_AOI21_D1", ["a1", "a2", "b"], "zn", []),
}


def random_case(rng, cell):
    mod, pins, out, edge_pins = CELL_SPEC[cell]
    stim, delays, sdf = {}, [], []
    nets = [f"p{j}" for j in range(len(pins))]
    for j, n in enumerate(nets):
        t, w = 0, []
        v = rng.choice([V0, V1, VX]) if rng.random() < 0.2 else rng.choice([V0, V1])
        if rng.random() < 0.85:
            w.append((0, v))
        for _ in range(rng.randint(0, 12)):
            t += rng.randint(1, 40)
            nv = rng.choice([x for x in (V0, V1, VX) if x != v])
            w.append((t, nv))
            v = nv
        stim[n] = w
    for j, p in enumerate(pins):
        if rng.random() < 0.1:
            continue  # no IOPATH: delay 1
        if p in edge_pins:
            for e, kw in ((0, "posedge"), (1, "negedge")):
                r, f = rng.randint(1, 30), rng.randint(1, 30)
                delays.append((j, e, r, f))
                sdf.append(f"(IOPATH ({kw} {p}) {out} ({r}:{r}:{r}) ({f}:{f}:{f}))")
        else:
            r, f = rng.randint(1, 30), rng.randint(1, 30)
            delays.append((j, 2, r, f))
            sdf.append(f"(IOPATH {p} {out} ({r}:{r}:{r}) ({f}:{f}:{f}))")
    conn = ", ".join(f".{p}({n})" for p, n in zip(pins, nets))
    inv = {0: "0", 1: "x", 2: "z", 3: "1"}
    case = dict(name="rand", insts=[f"{mod} U0 ({conn}, .{out}(y));"], sdf={"U0": sdf},
                stim={n: [(t, inv[v]) for t, v in w] for n, w in stim.items()}, expect={})
    return case, [stim[n] for n in nets], delays


def test_random_gates_against_closed_form(tmp_path):
    rng = random.Random(1234)
    cells = list(CELL_SPEC)
    for i in range(150):
        cell = cells[i % len(cells)]
        case, waves, delays = random_case(rng, cell)
        if all(len(w) == 0 for w in waves):
            continue
        d = tmp_path / f"c{i}"
        d.mkdir()
        vlib = kat_cases.VLIB
        _, gv, sdf, vcd = kat_cases.texts(case)
        paths = []
        for nm, t in (("lib.vlib", vlib), ("top.gv", gv), ("top.sdf", sdf), ("stim.vcd", vcd)):
            (d / nm).write_text(t)
            paths.append(str(d / nm))
        out = oracle_bind.run(*paths, 0, 2000, str(d), waves=True)
        want = spec_py.gate_wave(cell, waves, delays)
        assert out["waves"]["y"] == want, (i, cell, waves, delays)


def test_saif_closed_form(tmp_path):
    # a PI square wave: 0 for 100, 1 for 100, ... then X; windows clipped both sides
    case = dict(name="sq", insts=["GEN_BUF_D1 U0 (.i(a), .z(y));"],
                sdf={"U0": ["(IOPATH i z (1:1:1) (1:1:1))"]},
                stim={"a": [(50, 0), (150, 1), (250, 0), (350, 1), (450, "x")]}, expect={})
    paths = kat_cases.write_case(case, str(tmp_path))
    for start, end in ((0, 600), (100, 300), (150, 250), (151, 449), (0, 50), (450, 900)):
        out = oracle_bind.run(*paths, start, end, str(tmp_path), waves=True)
        recs = parse_saif(out["saif"])
        for net in ("a", "y"):
            w = out["waves"].get(net) or [(t, kat_cases.V(v)) for t, v in case["stim"]["a"]]
            t0, t1, tx = spec_py.saif(w, start, end)
            assert recs[net] == (t0, t1, tx), (net, start, end)
            assert sum(recs[net]) == end - start


def parse_saif(text):
    recs, name = {}, None
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("(T0"):
            p = s.replace("(", " ").replace(")", " ").split()
            recs[name] = (int(p[1]), int(p[3]), int(p[5]))
        elif s.startswith("(") and not s.startswith("(INSTANCE") and not s.startswith("(NET") \
                and not s.startswith("(TIMESCALE") and not s.startswith("(DURATION"):
            name = s[1:]
    return recs


GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "oracle_golden.json")


def test_golden_digests(tmp_path):
    import golden_cases
    with open(GOLDEN) as f:
        gold = json.load(f)
    for name, want in gold.items():
        if want.get("slow"):
            continue
        paths = golden_cases.write_inputs(name, str(tmp_path))
        start, end = golden_cases.CASES[name]["window"]
        out = oracle_bind.run(*paths, start, end, str(tmp_path))
        assert out["n_events"] == want["n_events"], name
        assert "%016x" % out["waves_hash"] == want["waves_hash"], name
        assert oracle_bind.sha256(out["saif"]) == want["saif_sha256"], name
