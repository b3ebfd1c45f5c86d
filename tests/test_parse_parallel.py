"""The gv and SDF readers and the netlist compiler give the same result on
one thread and on many.

Large texts (over 8 MiB) are read in parallel: SDF in stretches that start
after a line closing a block, gv as a serial statement scan, parallel
tokenising and a serial numbering pass.  These tests compare the design
digest (every net, pin and delay bit for bit) and the first reported error of
a one-thread read with threaded reads, on a 200k-cell netlist and on edited
copies that put the awkward cases at many places: TIMESCALE changes in the
middle of the SDF (recognised and unrecognised spellings, inside and outside
blocks), blocks for unknown instances, repeated blocks, statements over
several lines, constants, bus bits first seen on a pin, wires declared late,
`assign`, unknown modules, concatenations, and planted errors.
"""
import os
import random
import re
import subprocess
import sys

import pytest

import gls

THREADS = (5, 16)


@pytest.fixture(scope="module")
def texts(tmp_path_factory):
    d = gls.Design.generate(1, 0x5EED0001, 200000, 40, 1024)
    p = tmp_path_factory.mktemp("parse")
    paths = [str(p / n) for n in ("lib.vlib", "netlist.gv", "netlist.sdf")]
    d.write(*paths)
    out = [open(x).read() for x in paths]
    assert len(out[1]) > (8 << 20) and len(out[2]) > (8 << 20)  # both take the parallel path
    return out


def outcome(vlib, gv, sdf, threads):
    try:
        return ("ok", gls.Design.from_texts(vlib, gv, sdf, threads=threads).digest())
    except gls.GlsError as e:
        return ("error", e.code, str(e))


def same_for_all_threads(vlib, gv, sdf):
    one = outcome(vlib, gv, sdf, 1)
    for t in THREADS:
        assert outcome(vlib, gv, sdf, t) == one, t
    return one


def edit_sdf(sdf, seed, bad=()):
    rng = random.Random(seed)
    lines = sdf.split("\n")
    out, seen, cur = [], [], None
    spell = ["1ns", "1ps", "10ps", "1us", "1ms", "1s", "weird"]
    for i, l in enumerate(lines):
        out.append(l)
        s = l.strip()
        if s.startswith("(INSTANCE"):
            cur = s[len("(INSTANCE "):-1]
        elif s.startswith("(IOPATH") and cur:
            seen.append((cur, s))
            cur = None
        if l == "  )" and rng.random() < 0.0005:
            out.append("  (TIMESCALE %s)" % rng.choice(spell))
        if s.startswith("(IOPATH") and rng.random() < 0.0005:
            out.append("        (TIMESCALE 1us)")  # inside a block: no effect
        if l == "  )" and rng.random() < 0.001:
            out += ["  (CELL", '    (CELLTYPE "X")', "    (INSTANCE Unknown%d)" % i, "    (DELAY",
                    "      (ABSOLUTE", "        (IOPATH a zn (1:2:3) (4:5:6))", "      )", "    )",
                    "  )"]
        if l == "  )" and seen and rng.random() < 0.001:
            name, io = rng.choice(seen)
            io = io.replace(" ", "\t", 1)
            head = io[:io.rfind("(", 0, io.rfind("("))]
            out += ["  (CELL", '\t(CELLTYPE "X")', "    (INSTANCE %s)" % name, "", "    (DELAY",
                    "      (ABSOLUTE", "        " + head + "(0.%d::7) (3))" % rng.randrange(100),
                    "      )", "    )", "  )"]
        if l == "      )" and rng.random() < 0.0005:
            out.append("      )   ")
    for frac in bad:  # a delay value that does not parse
        j = int(len(out) * frac)
        while not out[j].strip().startswith("(IOPATH"):
            j += 1
        a = out[j].index(":")
        b = out[j].index(":", a + 1)
        out[j] = out[j][:a + 1] + "bad%d" % j + out[j][b:]
    return "\n".join(out)


def edit_gv(gv, seed, errors=()):
    rng = random.Random(seed)
    out, late, k = [], [], 0
    pin = r"\.(a1|i)\((\w+)\)"
    for l in gv.split("\n"):
        if l.strip().startswith("GEN_") and rng.random() < 0.01:
            kind = rng.randrange(8)
            if kind == 0:  # a statement over several lines
                parts = l.split(", ")
                out += [p + ("," if i + 1 < len(parts) else "") for i, p in enumerate(parts)]
                continue
            if kind == 1:
                l = re.sub(pin, lambda m: ".%s(1'b%d)" % (m.group(1), rng.randrange(2)), l, count=1)
            if kind == 2:  # a bus bit first seen on a pin
                l = re.sub(pin, lambda m: ".%s(nbus%d[%d])" % (m.group(1), rng.randrange(50),
                                                                 rng.randrange(8)), l, count=1)
            if kind == 3:
                out += ["  wire late%d;" % k, "  assign late%d = n%d;" % (k, rng.randrange(100))]
                late.append(k)
                k += 1
            if kind == 4 and late:
                l = re.sub(pin, lambda m: ".%s(late%d)" % (m.group(1), rng.choice(late)), l, count=1)
            if kind == 5:
                out += ["  assign cst%d = 1'b%d;" % (k, rng.randrange(2)),
                        "  wire [3:0] bus%d, busb%d;" % (k, k)]
                k += 1
            if kind == 6:
                out.append("  FOO_CELL ufoo%d (.a(n1), .zz(n2));" % k)
                k += 1
            if kind == 7:
                l = re.sub(pin, lambda m: ".%s({%s,1'b0,nb%d[3:1]})" % (m.group(1), m.group(2),
                                                                       rng.randrange(9)), l, count=1)
        out.append(l)
    insts = [j for j, l in enumerate(out) if l.startswith("  GEN_") and l.endswith(";")]
    for kind, frac in errors:
        j = insts[int(len(insts) * frac)]
        if kind == "undeclared":
            out[j] = re.sub(r"\((n\d+|pi\d+)\)", "(nosuchnet)", out[j], count=1)
        elif kind == "pin":
            out[j] = out[j].replace(".zn(", ".qq(").replace(".z(", ".qq(")
        elif kind == "both":  # undeclared net on the first pin, unknown pin later
            out[j] = re.sub(r"\((n\d+|pi\d+)\)", "(nosuchnet2)", out[j], count=1)
            out[j] = out[j].replace(".zn(", ".qq(").replace(".z(", ".qq(")
        elif kind == "used_early":
            out[j] = re.sub(r"\((n\d+|pi\d+)\)", "(declared_later)", out[j], count=1)
            out.insert(j + 5, "  wire declared_later;")
    return "\n".join(out)


def test_generated_texts(texts):
    vlib, gv, sdf = texts
    r = same_for_all_threads(vlib, gv, sdf)
    assert r[0] == "ok"
    d = gls.Design.generate(1, 0x5EED0001, 200000, 40, 1024)
    assert gls.Design.from_texts(vlib, gv, sdf).info() == d.info()


@pytest.mark.parametrize("seed", [1, 2])
def test_sdf_edits(texts, seed):
    vlib, gv, sdf = texts
    edited = edit_sdf(sdf, seed)
    r = same_for_all_threads(vlib, gv, edited)
    assert r[0] == "ok"
    assert r != outcome(vlib, gv, sdf, 1)  # the edits changed some delays


def test_sdf_first_error_wins(texts):
    vlib, gv, sdf = texts
    r = same_for_all_threads(vlib, gv, edit_sdf(sdf, 3, bad=(0.3, 0.7)))
    assert r[0] == "error" and r[1] == -2 and "bad delay value" in r[2]
    # the earlier of the two planted values is the one reported
    line = int(re.search(r"bad(\d+)", r[2]).group(1))
    assert line < len(sdf.split("\n")) * 0.5


@pytest.mark.parametrize("seed", [1, 2])
def test_gv_edits(texts, seed):
    vlib, gv, sdf = texts
    r = same_for_all_threads(vlib, edit_gv(gv, seed), sdf)
    assert r[0] == "ok"


@pytest.mark.parametrize("errors,expect", [
    ((("pin", 0.3), ("undeclared", 0.7)), "has no pin qq"),
    ((("undeclared", 0.2), ("pin", 0.6)), "undeclared net nosuchnet "),
    ((("both", 0.5),), "undeclared net nosuchnet2"),
    ((("used_early", 0.4),), "undeclared net declared_later"),
])
def test_gv_first_error_wins(texts, errors, expect):
    vlib, gv, sdf = texts
    r = same_for_all_threads(vlib, edit_gv(gv, 5, errors), sdf)
    assert r[0] == "error" and r[1] == -2 and expect in r[2], r


def edit_for_compile(gv, seed):
    """Unconnected outputs (pseudo nets) throughout, and late in the file
    instances with one net on two input pins (truth tables first seen there)."""
    rng = random.Random(seed)
    lines = gv.split("\n")
    count = {}
    for l in lines:
        if l.startswith("  GEN_"):
            for n in re.findall(r"\((\w+)\)", l):
                count[n] = count.get(n, 0) + 1
    insts = [i for i, l in enumerate(lines) if l.startswith("  GEN_")]
    for pos, i in enumerate(insts):
        l = lines[i]
        m = re.search(r", \.(zn|z)\((\w+)\)\);", l)
        if m and count.get(m.group(2), 0) == 1 and rng.random() < 0.3:
            l = l.replace(m.group(0), ");")
        if pos > len(insts) * 0.95 and rng.random() < 0.05:
            ins = [p for p in re.findall(r"\.(\w+)\((\w+)\)", l) if p[0] not in ("z", "zn")]
            if len(ins) >= 2:
                l = l.replace(".%s(%s)" % ins[1], ".%s(%s)" % (ins[1][0], ins[0][1]), 1)
        lines[i] = l
    return "\n".join(lines)


def program_digest_one_thread(paths):
    code = ("import sys, gls; d = gls.Design.from_files(*sys.argv[1:4]); "
            "print(d.info()['program_digest'])")
    env = dict(os.environ, GLS_HOST_THREADS="1",
               PYTHONPATH=os.pathsep.join(p for p in sys.path if p))
    out = subprocess.run([sys.executable, "-c", code] + list(paths), env=env, check=True,
                         capture_output=True, text=True, timeout=600)
    return int(out.stdout.strip().split()[-1])


def test_compile_same_on_one_thread(texts, tmp_path):
    vlib, gv, sdf = texts
    cases = [(gv, sdf), (edit_for_compile(gv, 1), "")]
    for k, (g, s) in enumerate(cases):
        paths = [str(tmp_path / ("%d.%s" % (k, e))) for e in ("vlib", "gv", "sdf")]
        for p, t in zip(paths, (vlib, g, s)):
            with open(p, "w") as f:
                f.write(t)
        info = gls.Design.from_files(*paths).info()  # compiled on several threads
        assert info["n_gates"] > 0
        assert program_digest_one_thread(paths) == info["program_digest"], k


def test_inter_roundtrip_keeps_the_generated_design(tmp_path):
    """Generator -> inter.dat writer -> loader: the same netlist and program,
    bit for bit, as the generator's own."""
    d = gls.Design.generate(1, 0x5EED0001, 200000, 40, 1024)
    p = str(tmp_path / "inter.dat")
    d.save_inter(p)
    assert os.path.getsize(p) > (16 << 20)
    d2 = gls.Design.load_inter(p)
    assert d2.digest() == d.digest()
    assert d2.info() == d.info()
