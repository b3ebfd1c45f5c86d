"""Known-answer cases: tiny netlists whose output waveforms are worked out by
hand from the reference's event rules (src/simulate.cpp:919-1201).  Each case
is (name, gv instances, SDF IOPATH lines, stimulus, expected waveforms).

Values: 0, 1, 'x', 'z'; waveforms are lists of (time, value) with value codes
0 -> 0, X -> 1, Z -> 2, 1 -> 3.
"""

VLIB = """`timescale 1ps/1ps
primitive udp_mux2 (z, i0, i1, s);
   output z;
   input  i0, i1, s;
   table
    1  ?  0 :  1;
    0  ?  0 :  0;
    ?  1  1 :  1;
    ?  0  1 :  0;
    0  0  x :  0;
    1  1  x :  1;
   endtable
endprimitive
module GEN_INV_D1 (i,zn);
  input i;
  output zn;
  not (zn, i);
endmodule
module GEN_BUF_D1 (i,z);
  input i;
  output z;
  buf (z, i);
endmodule
module GEN_AND2_D1 (a1,a2,z);
  input a1;
  input a2;
  output z;
  and (z, a1, a2);
endmodule
module GEN_XOR2_D1 (a1,a2,z);
  input a1;
  input a2;
  output z;
  xor (z, a1, a2);
endmodule
module GEN_MUX2_D1 (i0,i1,s,z);
  input i0;
  input i1;
  input s;
  output z;
  udp_mux2 udpi0 (z, i0, i1, s);
endmodule
module GEN_AOI21_D1 (a1,a2,b,zn);
  input a1;
  input a2;
  input b;
  output zn;
  wire net0, net1;
  and (net0, a1, a2);
  or (net1, net0, b);
  not (zn, net1);
endmodule
module GEN_TIE_D1 (a,z);
  input a;
  output z;
  and (z, a, 1'b1);
endmodule
module GEN_AO221_D1 (a1,a2,b1,b2,c,z);
  input a1;
  input a2;
  input b1;
  input b2;
  input c;
  output z;
  wire net0, net1;
  and (net0, a1, a2);
  and (net1, b1, b2);
  or (z, net0, net1, c);
endmodule
module GEN_AO222_D1 (a1,a2,b1,b2,c1,c2,z);
  input a1;
  input a2;
  input b1;
  input b2;
  input c1;
  input c2;
  output z;
  wire net0, net1, net2;
  and (net0, c1, c2);
  and (net1, a1, a2);
  and (net2, b1, b2);
  or (z, net0, net1, net2);
endmodule
module GEN_HA_D1 (a,b,s,co);
  input a;
  input b;
  output s;
  output co;
  xor (s, a, b);
  and (co, a, b);
endmodule
"""

C = {"0": 0, "x": 1, "z": 2, "1": 3}
X, Z = 1, 2


def V(c):
    return C[str(c)]


CASES = [
    # INV, rise 10 / fall 20: every input change gives one output change.
    dict(name="inv_basic",
         insts=["GEN_INV_D1 U0 (.i(a), .zn(y));"],
         sdf={"U0": ["(IOPATH i zn (10:10:10) (20:20:20))"]},
         stim={"a": [(0, 0), (100, 1), (200, 0)]},
         expect={"y": [(0, X), (10, 3), (120, 0), (210, 3)]}),
    # BUF rise 50 / fall 5: a 20 ps high pulse is swallowed (fall at 125
    # pre-empts the rise scheduled at 150; the value equals the queue tail).
    dict(name="buf_preempt",
         insts=["GEN_BUF_D1 U0 (.i(a), .z(y));"],
         sdf={"U0": ["(IOPATH i z (50:50:50) (5:5:5))"]},
         stim={"a": [(0, 0), (100, 1), (120, 0)]},
         expect={"y": [(0, X), (5, 0)]}),
    # BUF rise 5 / fall 50: the same pulse passes.
    dict(name="buf_pass",
         insts=["GEN_BUF_D1 U0 (.i(a), .z(y));"],
         sdf={"U0": ["(IOPATH i z (5:5:5) (50:50:50))"]},
         stim={"a": [(0, 0), (100, 1), (120, 0)]},
         expect={"y": [(0, X), (50, 0), (105, 3), (170, 0)]}),
    # AND2, both inputs switch together: delay = min over the related IOPATHs.
    dict(name="and2_min",
         insts=["GEN_AND2_D1 U0 (.a1(a), .a2(b), .z(y));"],
         sdf={"U0": ["(IOPATH a1 z (10:10:10) (30:30:30))", "(IOPATH a2 z (20:20:20) (15:15:15))"]},
         stim={"a": [(0, 0), (100, 1)], "b": [(0, 0), (100, 1)]},
         expect={"y": [(0, X), (15, 0), (110, 3)]}),
    # AND2 with an X on one input while the other is 0: no output event;
    # then the other input rises: X out (rise class from 0), back to 1.
    dict(name="and2_x",
         insts=["GEN_AND2_D1 U0 (.a1(a), .a2(b), .z(y));"],
         sdf={"U0": ["(IOPATH a1 z (10:10:10) (30:30:30))", "(IOPATH a2 z (20:20:20) (15:15:15))"]},
         stim={"a": [(0, 0), (50, "x"), (200, 1)], "b": [(0, 0), (100, 1)]},
         # t0: f=0 fall, num_in=2 -> min(30,15)=15.  t50: a->X, b=0: f=0, no event.
         # t100: b->1, a=X: f=X, 0->X is class 1 (rise), num_in=1, one IOPATH -> 20 -> (120,X).
         # t200: a->1: f=1, X->1 rise, num_in=1 -> a1 rise 10 -> (210,1).
         expect={"y": [(0, X), (15, 0), (120, X), (210, 3)]}),
    # XOR2 with posedge/negedge IOPATHs per input.
    dict(name="xor2_edges",
         insts=["GEN_XOR2_D1 U0 (.a1(a), .a2(b), .z(y));"],
         sdf={"U0": ["(IOPATH (posedge a1) z (10:10:10) (11:11:11))",
                     "(IOPATH (negedge a1) z (12:12:12) (13:13:13))",
                     "(IOPATH (posedge a2) z (20:20:20) (21:21:21))",
                     "(IOPATH (negedge a2) z (22:22:22) (23:23:23))"]},
         stim={"a": [(0, 0), (100, 1), (300, 0)], "b": [(0, 1), (200, 0)]},
         # t0: f=1 rise; a: X->0 is a fall -> negedge a1 (rise 12); b: X->1 rise -> posedge a2 (20); min 12
         # t100: a 0->1 (posedge) f=0 fall -> 11.  t200: b 1->0 (negedge) f=1 rise -> 22.
         # t300: a 1->0 (negedge) f=0 fall -> 13.
         expect={"y": [(0, X), (12, 3), (111, 0), (222, 3), (313, 0)]}),
    # MUX2 select going X with different data: output X; an X select edge
    # counts as posedge.
    dict(name="mux2_xsel",
         insts=["GEN_MUX2_D1 U0 (.i0(d0), .i1(d1), .s(s), .z(y));"],
         sdf={"U0": ["(IOPATH i0 z (30:30:30) (31:31:31))",
                     "(IOPATH i1 z (40:40:40) (41:41:41))",
                     "(IOPATH (posedge s) z (7:7:7) (8:8:8))",
                     "(IOPATH (negedge s) z (9:9:9) (10:10:10))"]},
         stim={"d0": [(0, 0)], "d1": [(0, 1)], "s": [(0, 0), (100, "x"), (200, 1), (300, "x")]},
         # t0: f=d0=0, class fall, num_in=3: i0 31, i1 41, s X->0 fall -> negedge fall 10 -> 10
         # t100: s 0->X: f=X, class rise (0->X); num_in=1, two IOPATHs, input edge 0->X = rise -> posedge rise 7
         # t200: s X->1: f=1, class rise; input edge X->1 rise -> posedge rise 7 -> 207
         # t300: s 1->X: f=X, class fall (1->X); input edge 1->X fall -> negedge fall 10 -> 310
         expect={"y": [(0, X), (10, 0), (107, X), (207, 3), (310, X)]}),
    # MUX2 with select X at t = 0 and equal data: f = data; the select's
    # X->X "edge" is regarded as posedge (num_in = 3, min over i0/i1/posedge s).
    dict(name="mux2_x_at_zero",
         insts=["GEN_MUX2_D1 U0 (.i0(d0), .i1(d1), .s(s), .z(y));"],
         sdf={"U0": ["(IOPATH i0 z (30:30:30) (31:31:31))",
                     "(IOPATH i1 z (40:40:40) (41:41:41))",
                     "(IOPATH (posedge s) z (7:7:7) (8:8:8))",
                     "(IOPATH (negedge s) z (9:9:9) (10:10:10))"]},
         stim={"d0": [(0, 1)], "d1": [(0, 1)], "s": [(0, "x")]},
         expect={"y": [(0, X), (7, 3)]}),
    # AOI21 chain with internal wires; an input with no IOPATH uses delay 1.
    dict(name="aoi21_default_delay",
         insts=["GEN_AOI21_D1 U0 (.a1(a), .a2(b), .b(c), .zn(y));"],
         sdf={"U0": ["(IOPATH a1 zn (10:10:10) (20:20:20))", "(IOPATH a2 zn (11:11:11) (21:21:21))"]},
         stim={"a": [(0, 1)], "b": [(0, 1)], "c": [(0, 0), (100, 1), (200, 0)]},
         # t0: a1&a2=1 -> zn=0, class fall; num_in=3, related a1,a2 -> min(20,21)=20
         # t100: c->1: zn stays 0: no event.  t200: no event.
         expect={"y": [(0, X), (20, 0)]}),
    dict(name="aoi21_no_iopath",
         insts=["GEN_AOI21_D1 U0 (.a1(a), .a2(b), .b(c), .zn(y));"],
         sdf={"U0": ["(IOPATH a1 zn (10:10:10) (20:20:20))"]},
         stim={"a": [(0, 0)], "b": [(0, 0)], "c": [(0, 0), (100, 1), (200, 0)]},
         # t0: zn = 1 rise, num_in=3, related {a1} -> 10.  t100: c->1: zn=0 fall,
         # num_in=1, no related IOPATH -> delay 1 -> (101,0).  t200: c->0 -> (201,1).
         expect={"y": [(0, X), (10, 3), (101, 0), (201, 3)]}),
    # Cell with a supply tie inside (and with 1'b1): z = a.
    dict(name="tie_supply",
         insts=["GEN_TIE_D1 U0 (.a(a), .z(y));"],
         sdf={"U0": ["(IOPATH a z (3:3:3) (4:4:4))"]},
         stim={"a": [(0, 1), (50, 0)]},
         expect={"y": [(0, X), (3, 3), (54, 0)]}),
    # Two levels: INV feeding BUF; the second level sees the first's (0,X).
    dict(name="two_levels",
         insts=["GEN_INV_D1 U0 (.i(a), .zn(m));", "GEN_BUF_D1 U1 (.i(m), .z(y));"],
         sdf={"U0": ["(IOPATH i zn (10:10:10) (20:20:20))"], "U1": ["(IOPATH i z (5:5:5) (6:6:6))"]},
         stim={"a": [(0, 0), (100, 1)]},
         # m: (0,X),(10,1),(120,0).  y: t0 f=X (buf of X) no event; t10: f=1 rise 5 -> 15;
         # t120: f=0 fall 6 -> 126
         expect={"m": [(0, X), (10, 3), (120, 0)], "y": [(0, X), (15, 3), (126, 0)]}),
    # A cell with two outputs: each output is its own gate with its own
    # IOPATHs; the chain is evaluated up to the function driving that output.
    dict(name="half_adder_two_outputs",
         insts=["GEN_HA_D1 U0 (.a(a), .b(b), .s(y), .co(c));"],
         sdf={"U0": ["(IOPATH a s (10:10:10) (11:11:11))", "(IOPATH b s (12:12:12) (13:13:13))",
                     "(IOPATH a co (20:20:20) (21:21:21))", "(IOPATH b co (22:22:22) (23:23:23))"]},
         stim={"a": [(0, 0), (100, 1), (300, 0)], "b": [(0, 0), (200, 1)]},
         # t0: both inputs, s=0 and co=0 fall, num_in=2 -> s min(11,13)=11, co min(21,23)=21.
         # t100: a->1: s=1 rise, one related IOPATH a->s -> 10 -> (110,1); co stays 0.
         # t200: b->1: s=0 fall via b->s 13 -> (213,0); co=1 rise via b->co 22 -> (222,1).
         # t300: a->0: s=1 rise via a->s 10 -> (310,1); co=0 fall via a->co 21 -> (321,0).
         expect={"y": [(0, X), (11, 0), (110, 3), (213, 0), (310, 3)],
                 "c": [(0, X), (21, 0), (222, 3), (321, 0)]}),
    # Gate tied to a constant through the netlist (1'b0 on an input pin).
    dict(name="const_input",
         insts=["GEN_AND2_D1 U0 (.a1(a), .a2(1'b1), .z(y));"],
         sdf={"U0": ["(IOPATH a1 z (10:10:10) (20:20:20))", "(IOPATH a2 z (1:1:1) (2:2:2))"]},
         stim={"a": [(0, 0), (100, 1)]},
         # t0: a and const both have entries -> num_in=2, f=0 fall: min(20, 2)=2
         # t100: f=1 rise, num_in=1 -> a1 rise 10 -> 110
         expect={"y": [(0, X), (2, 0), (110, 3)]}),
    # AO221 (5 distinct inputs, beyond the lane walk's fan-in of 4): the
    # 3-input OR of two ANDs and c.
    dict(name="ao221_five_inputs",
         insts=["GEN_AO221_D1 U0 (.a1(a), .a2(b), .b1(c), .b2(d), .c(e), .z(y));"],
         sdf={"U0": ["(IOPATH a1 z (10:10:10) (11:11:11))", "(IOPATH a2 z (12:12:12) (13:13:13))",
                     "(IOPATH b1 z (14:14:14) (15:15:15))", "(IOPATH b2 z (16:16:16) (17:17:17))",
                     "(IOPATH c z (18:18:18) (19:19:19))"]},
         stim={"a": [(0, 0), (200, 1), (500, 0)], "b": [(0, 0), (300, 1)],
               "c": [(0, 0), (600, "x")], "d": [(0, 0), (700, 1)], "e": [(0, 0), (100, 1), (400, 0)]},
         # t0: all five inputs, z = 0 (fall), num_in = 5 -> min of the falls = 11.
         # t100: e -> 1: z rises via c 18 -> 118.  t200, t300: a1, a2 -> 1: z stays 1.
         # t400: e -> 0: z = a1 & a2 = 1, no event.  t500: a1 -> 0: z falls via a1 11 -> 511.
         # t600: b1 -> X with b2 = 0: z stays 0.  t700: b2 -> 1: z = X, 0 -> X is a
         # rise, via b2 16 -> 716.
         expect={"y": [(0, X), (11, 0), (118, 3), (511, 0), (716, X)]}),
    # AO222 (6 distinct inputs) with edge-specific IOPATHs on c1, events on
    # two inputs at once, and a pre-empted candidate.
    dict(name="ao222_six_inputs",
         insts=["GEN_AO222_D1 U0 (.a1(a), .a2(b), .b1(c), .b2(d), .c1(e), .c2(f), .z(y));"],
         sdf={"U0": ["(IOPATH a1 z (10:10:10) (20:20:20))", "(IOPATH a2 z (11:11:11) (21:21:21))",
                     "(IOPATH b1 z (12:12:12) (22:22:22))", "(IOPATH b2 z (13:13:13) (23:23:23))",
                     "(IOPATH (posedge c1) z (5:5:5) (6:6:6))",
                     "(IOPATH (negedge c1) z (7:7:7) (8:8:8))",
                     "(IOPATH c2 z (14:14:14) (24:24:24))"]},
         stim={"a": [(0, 0), (400, 1)], "b": [(0, 0), (400, 1), (405, 0)],
               "c": [(0, 0), (500, 1), (515, 0)], "d": [(0, 0), (510, 1)],
               "e": [(0, 0), (100, 1), (200, 0), (300, 1), (601, 0)],
               "f": [(0, 0), (100, 1), (302, 0), (600, 1)]},
         # t0: all six inputs, z = 0 (fall); c1 X -> 0 is a negedge: fall 8;
         # others 20..24 -> 8.  t100: c1, c2 -> 1: rise, min(posedge c1 5, c2 14)
         # -> 105.  t200: c1 -> 0: fall, negedge c1 8 -> 208.  t300: c1 -> 1:
         # rise 5 -> 305.  t302: c2 -> 0: fall 24 -> 326.  t400: a1, a2 -> 1:
         # rise min(10, 11) -> 410.  t405: a2 -> 0: fall 21 -> 426.  t510: b2
         # -> 1 (b1 = 1): rise 13 -> 523.  t515: b1 -> 0: fall 22 -> 537.
         # t600: c2 -> 1 (c1 = 1): rise 14 -> 614; t601: c1 -> 0: fall 8 ->
         # 609 pre-empts 614, and 0 repeats the value from 537: no change.
         expect={"y": [(0, X), (8, 0), (105, 3), (208, 0), (305, 3), (326, 0), (410, 3), (426, 0),
                       (523, 3), (537, 0)]}),
    # `assign p = m` (src/PreProcess.cpp:780-863): an instance input on the
    # alias reads the waveform of the net it names, here a gate output, so
    # this is two_levels with the alias between the cells.
    dict(name="assign_gate_output",
         insts=["GEN_INV_D1 U0 (.i(a), .zn(m));", "GEN_BUF_D1 U1 (.i(p), .z(y));"],
         assigns=[("p", "m")],
         sdf={"U0": ["(IOPATH i zn (10:10:10) (20:20:20))"], "U1": ["(IOPATH i z (5:5:5) (6:6:6))"]},
         stim={"a": [(0, 0), (100, 1)]},
         expect={"m": [(0, X), (10, 3), (120, 0)], "y": [(0, X), (15, 3), (126, 0)]}),
    # alias of a stimulus net: buf_pass with the alias on the cell input
    dict(name="assign_stimulus",
         insts=["GEN_BUF_D1 U0 (.i(p), .z(y));"],
         assigns=[("p", "a")],
         sdf={"U0": ["(IOPATH i z (5:5:5) (50:50:50))"]},
         stim={"a": [(0, 0), (100, 1), (120, 0)]},
         expect={"y": [(0, X), (50, 0), (105, 3), (170, 0)]}),
    # `assign k = 1'b1`: a constant net (value 1 from time 0), as const_input
    dict(name="assign_constant",
         insts=["GEN_AND2_D1 U0 (.a1(a), .a2(k), .z(y));"],
         assigns=[("k", "1'b1")],
         sdf={"U0": ["(IOPATH a1 z (10:10:10) (20:20:20))", "(IOPATH a2 z (1:1:1) (2:2:2))"]},
         stim={"a": [(0, 0), (100, 1)]},
         expect={"y": [(0, X), (2, 0), (110, 3)]}),
    # the stimulus drives the alias net itself: the VCD waveform wins over
    # the alias (precedence VCD > gate output > constant > assign,
    # src/simulate.cpp:1617-1664)
    dict(name="assign_overridden_by_stimulus",
         insts=["GEN_INV_D1 U0 (.i(a), .zn(m));", "GEN_BUF_D1 U1 (.i(p), .z(y));"],
         assigns=[("p", "m")],
         sdf={"U0": ["(IOPATH i zn (10:10:10) (20:20:20))"], "U1": ["(IOPATH i z (5:5:5) (6:6:6))"]},
         stim={"a": [(0, 0), (100, 1)], "p": [(0, 1), (150, 0)]},
         expect={"m": [(0, X), (10, 3), (120, 0)], "y": [(0, X), (5, 3), (156, 0)]}),
]


def vcd_id(i):
    s = ""
    while True:
        s += chr(33 + i % 94)
        i //= 94
        if i == 0:
            return s


def texts(case):
    """Return (vlib, gv, sdf, vcd) texts of a case."""
    nets_in = list(case["stim"].keys())
    wires = set()
    for inst in case["insts"]:
        body = inst[inst.index("(") + 1:]
        for part in body.split(".")[1:]:
            net = part[part.index("(") + 1:part.index(")")]
            if net not in nets_in and not net.startswith("1'b"):
                wires.add(net)
    for left, right in case.get("assigns", []):
        for net in (left, right):
            if net not in nets_in and not net.startswith("1'b"):
                wires.add(net)
    gv = "module top ( );\n"
    gv += "  input " + ", ".join(nets_in) + ";\n"
    if wires:
        gv += "  wire " + ", ".join(sorted(wires)) + ";\n"
    for left, right in case.get("assigns", []):
        gv += "  assign %s = %s;\n" % (left, right)
    for inst in case["insts"]:
        gv += "  " + inst + "\n"
    gv += "endmodule\n"
    sdf = '(DELAYFILE\n  (SDFVERSION "3.0")\n  (TIMESCALE 1ps)\n'
    for iname, lines in case["sdf"].items():
        sdf += f"  (CELL\n    (CELLTYPE \"X\")\n    (INSTANCE {iname})\n    (DELAY\n      (ABSOLUTE\n"
        for l in lines:
            sdf += "        " + l + "\n"
        sdf += "      )\n    )\n  )\n"
    sdf += ")\n"
    vcd = "$timescale 1ps $end\n$scope module top $end\n"
    ids = {}
    for i, n in enumerate(nets_in):
        ids[n] = vcd_id(i)
        vcd += f"$var wire 1 {ids[n]} {n} $end\n"
    vcd += "$upscope $end\n$enddefinitions $end\n"
    events = []
    for n, w in case["stim"].items():
        for t, v in w:
            events.append((t, n, str(v)))
    events.sort(key=lambda e: (e[0], nets_in.index(e[1])))
    last = None
    for t, n, v in events:
        if t != last:
            vcd += f"#{t}\n"
            last = t
        vcd += f"{v}{ids[n]}\n"
    return VLIB, gv, sdf, vcd


def write_case(case, d):
    import os
    vlib, gv, sdf, vcd = texts(case)
    paths = [os.path.join(d, n) for n in ("lib.vlib", "top.gv", "top.sdf", "stim.vcd")]
    for p, t in zip(paths, (vlib, gv, sdf, vcd)):
        with open(p, "w") as f:
            f.write(t)
    return paths
