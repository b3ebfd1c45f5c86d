"""Host logic: readers/writers, generator determinism, levelization, error
codes for malformed inputs, inter.dat round trip."""
import hashlib
import os

import pytest

import gls
import golden_cases
import kat_cases


def test_generator_is_deterministic(tmp_path):
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    pa = golden_cases.write_inputs("c0_small", str(a))
    pb = golden_cases.write_inputs("c0_small", str(b))
    for x, y in zip(pa, pb):
        assert open(x, "rb").read() == open(y, "rb").read()


def test_generator_digest_is_pinned(tmp_path):
    # byte-exact generator output across hosts (no libm in the draws)
    paths = golden_cases.write_inputs("c0_small", str(tmp_path))
    h = hashlib.sha256()
    for p in paths[1:]:
        h.update(open(p, "rb").read())
    assert h.hexdigest()[:16] == GEN_DIGEST


GEN_DIGEST = open(os.path.join(os.path.dirname(__file__), "golden", "gen_digest.txt")).read().strip()


def test_config_shapes():
    d = gls.Design.generate(0)
    i = d.info()
    assert i["n_instances"] == 2000 and i["n_inputs"] == 64 and i["n_levels"] == 20
    d1 = gls.Design.generate(1, 0, 20000, 40, 1024)
    i1 = d1.info()
    assert i1["n_levels"] == 40 and i1["n_inputs"] == 1024


def test_roundtrip_files_and_inter(tmp_path):
    paths = golden_cases.write_inputs("c0_small", str(tmp_path))
    d = gls.Design.from_files(*paths[:3])
    g = gls.Design.generate(0, golden_cases.seed("c0_small"), 400, 20, 64)
    assert d.info() == g.info()
    inter = str(tmp_path / "inter.dat")
    d.save_inter(inter)
    d2 = gls.Design.load_inter(inter)
    assert d2.info() == d.info()
    s = gls.Stimulus.from_vcd(paths[3])
    s2 = gls.Stimulus.generate(g, golden_cases.seed("c0_small"), 200000, 1000.0)
    assert s.info() == s2.info()


def test_inter_archive_and_compiled_section(tmp_path):
    # inter.dat opens with the `inter` object in text-archive syntax (the
    # reference's process step writes it, src/process.cpp:604-606), then the
    # compiled section; loading takes the program from that section
    # (program digest equal without compiling again), a damaged section is
    # GLS_E_FORMAT, and files of the three-text layout still load
    vlib, gv, sdf, _ = kat_cases.texts(kat_cases.CASES[0])
    d = gls.Design.from_texts(vlib, gv, sdf)
    p = str(tmp_path / "inter.dat")
    d.save_inter(p)
    raw = open(p, "rb").read()
    assert raw.startswith(b"22 serialization::archive 17 ")
    at = raw.index(b"\nGLS-CSR 1\n")
    text = raw[:at].decode()
    assert " 2 U0 " in text and " 10 GEN_INV_D1 " in text and " 3 top " in text
    assert raw.endswith(b"\nGLS-CSR-AT %020d\n" % at)
    d2 = gls.Design.load_inter(p)
    assert d2.info() == d.info() and d2.digest() == d.digest()
    bad = str(tmp_path / "bad.dat")
    open(bad, "wb").write(raw[:at + 40] + raw[at + 41:])
    with pytest.raises(gls.GlsError) as e:
        gls.Design.load_inter(bad)
    assert e.value.status == "GLS_E_FORMAT"
    open(bad, "wb").write(raw[:at])
    with pytest.raises(gls.GlsError) as e:
        gls.Design.load_inter(bad)
    assert e.value.status == "GLS_E_FORMAT"
    old = str(tmp_path / "old.dat")
    body = "GLS-INTER 1\n" + "".join("%s %d\n%s\n" % (k, len(t), t)
                                      for k, t in (("vlib", vlib), ("gv", gv), ("sdf", sdf)))
    open(old, "w").write(body)
    assert gls.Design.load_inter(old).info() == d.info()


def test_kat_designs_parse(tmp_path):
    for case in kat_cases.CASES:
        vlib, gv, sdf, vcd = kat_cases.texts(case)
        d = gls.Design.from_texts(vlib, gv, sdf)
        assert d.info()["n_instances"] == len(case["insts"])


@pytest.mark.parametrize("bad,status", [
    ("module top ( );\n  input a;\n  GEN_INV_D1 U0 (.i(a), .zn(undeclared));\nendmodule\n", "GLS_E_FORMAT"),
    ("module top ( );\n  input a;\n  wire y;\n  GEN_INV_D1 U0 (.q(a), .zn(y));\nendmodule\n", "GLS_E_FORMAT"),
])
def test_malformed_gv(bad, status):
    with pytest.raises(gls.GlsError) as e:
        gls.Design.from_texts(kat_cases.VLIB, bad, "")
    assert e.value.status == status


def test_combinational_loop_is_reported():
    gv = ("module top ( );\n  input a;\n  wire x, y;\n"
          "  GEN_AND2_D1 U0 (.a1(a), .a2(y), .z(x));\n  GEN_INV_D1 U1 (.i(x), .zn(y));\nendmodule\n")
    d = gls.Design.from_texts(kat_cases.VLIB, gv, "")
    with pytest.raises(gls.GlsError) as e:
        d.info()
    assert e.value.status == "GLS_E_CYCLE"


def test_undriven_net_is_reported():
    gv = "module top ( );\n  input a;\n  wire f, y;\n  GEN_AND2_D1 U0 (.a1(a), .a2(f), .z(y));\nendmodule\n"
    d = gls.Design.from_texts(kat_cases.VLIB, gv, "")
    with pytest.raises(gls.GlsError) as e:
        d.info()
    assert e.value.status == "GLS_E_CYCLE"


def test_malformed_vcd(tmp_path):
    p = tmp_path / "bad.vcd"
    p.write_text("$timescale 1ps $end\n$scope module top $end\n$var wire 1 ! a $end\n"
                 "$upscope $end\n$enddefinitions $end\n#0\n0!\n#5\n7!\n")
    with pytest.raises(gls.GlsError) as e:
        gls.Stimulus.from_vcd(str(p))
    assert e.value.status == "GLS_E_FORMAT"
    p.write_text("$timescale 1ps $end\n$enddefinitions $end\n#5000000000\n")  # past 2^32 - 1
    with pytest.raises(gls.GlsError) as e:
        gls.Stimulus.from_vcd(str(p))
    assert e.value.status == "GLS_E_RANGE"


def test_truncated_vcd_header(tmp_path):
    p = tmp_path / "t.vcd"
    p.write_text("$timescale 1ps $end\n$scope module top $end\n$var wire 1 ! a")
    with pytest.raises(gls.GlsError):
        gls.Stimulus.from_vcd(str(p))


def test_vcd_vectors_and_dedup(tmp_path):
    p = tmp_path / "v.vcd"
    p.write_text("$timescale 1ps $end\n$scope module top $end\n$var wire 4 ! bus [3:0] $end\n"
                 "$var wire 1 \" s $end\n$upscope $end\n$enddefinitions $end\n"
                 "#0\nb0011 !\n0\"\n#10\nb0111 !\n0\"\n#20\nbz !\n")
    s = gls.Stimulus.from_vcd(str(p))
    info = s.info()
    assert info["n_signals"] == 5
    # bus bits: [3]=0,0,z  [2]=0,1,z  [1]=1,1,z  [0]=1,1,z  -> deduped from an initial Z
    # scalars keep repeated values: s has 2 entries
    assert info["n_events"] == (2 + 3 + 2 + 2) + 2


def test_reference_library_cells_match_generic(tmp_path):
    """The generator's cell definitions agree with the reference library."""
    ref = "/root/reference/examples/GENERIC_STD_CELL.vlib"
    if not os.path.exists(ref):
        pytest.skip("reference library not mounted")
    paths = golden_cases.write_inputs("c0_small", str(tmp_path))
    d_ref = gls.Design.from_files(ref, paths[1], paths[2])
    d_gen = gls.Design.from_files(*paths[:3])
    assert d_ref.info() == d_gen.info()


def _big_vcd(path, n_ts=480000, seed=5, backwards_at=None, comment_at=None):
    """A VCD whose body is over 8 MiB (read in parallel stretches): a 6-bit
    bus and a 3-bit ascending bus whose values repeat (de-duplicated per
    bit), scalars with repeated values, several changes of one signal at one
    timestamp, and timestamps repeated on consecutive lines."""
    import random
    rng = random.Random(seed)
    out = ["$timescale 1ps $end", "$scope module top $end",
           "$var wire 6 ! bus [5:0] $end", '$var wire 1 " a $end', "$var wire 1 # b $end",
           "$var wire 3 % nib [0:2] $end", "$upscope $end", "$enddefinitions $end",
           "$dumpvars", "b000000 !", '0"', "x#", "b101 %", "$end"]
    t = 0
    for i in range(n_ts):
        if backwards_at is not None and i == backwards_at:
            t -= 3
        elif rng.random() > 0.05:
            t += rng.randint(1, 40)
        out.append("#%d" % t)
        if comment_at is not None and i == comment_at:
            out.append("$comment a note $end")
        for _ in range(rng.randint(1, 3)):
            k = rng.random()
            if k < 0.35:
                out.append("b%s !" % "".join(rng.choice("01xz") for _ in range(rng.randint(1, 6))))
            elif k < 0.6:
                out.append(rng.choice("01xz") + '"')
            elif k < 0.8:
                out.append(rng.choice("01x") + "#")
            else:
                out.append("b%s %%" % "".join(rng.choice("01") for _ in range(3)))
    path.write_text("\n".join(out) + "\n")
    return path


def _vcd_back(stim, path):
    stim.write_vcd(str(path))
    return path.read_text()


def test_parallel_vcd_read_equals_one_thread(tmp_path):
    p = _big_vcd(tmp_path / "big.vcd")
    assert p.stat().st_size > (8 << 20)
    one = gls.Stimulus.from_vcd(str(p), threads=1)
    ref = _vcd_back(one, tmp_path / "one.vcd")
    for th in (2, 7, 16, None):
        many = gls.Stimulus.from_vcd(str(p), threads=th)
        assert many.info() == one.info(), th
        assert _vcd_back(many, tmp_path / "many.vcd") == ref, th


def test_parallel_vcd_read_special_bodies(tmp_path):
    # a $comment in the body is read by one thread (a comment could hold a
    # line-initial '#'); the result is the same either way
    p = _big_vcd(tmp_path / "c.vcd", comment_at=200000)
    one = gls.Stimulus.from_vcd(str(p), threads=1)
    many = gls.Stimulus.from_vcd(str(p), threads=8)
    assert _vcd_back(many, tmp_path / "m.vcd") == _vcd_back(one, tmp_path / "o.vcd")
    # timestamps going backwards late in the body: the same error as one thread
    p = _big_vcd(tmp_path / "b.vcd", backwards_at=400000)
    for th in (1, 8):
        with pytest.raises(gls.GlsError) as e:
            gls.Stimulus.from_vcd(str(p), threads=th)
        assert e.value.status == "GLS_E_FORMAT", th
