"""Closed-form restatement of one gate's event rules, in Python, used to pin
the oracle on small random cases.  It differs from the oracle on purpose:
functions come from the Verilog primitive truth tables (not calc()), and the
output queue is computed as "strict suffix-min filter of the candidates, then
adjacent dedup" instead of the reference's pop/push loop
(src/simulate.cpp:990-1009).  The two agree iff the preemption rule is
implemented as the reference states it.

Codes: 0 -> 0, 1 -> X, 2 -> Z, 3 -> 1.
"""
V0, VX, VZ, V1 = 0, 1, 2, 3


def _kleene(v):  # Z behaves as X on gate inputs
    return VX if v == VZ else v


def f_and(*a):
    a = [_kleene(x) for x in a]
    if V0 in a:
        return V0
    return VX if VX in a else V1


def f_or(*a):
    a = [_kleene(x) for x in a]
    if V1 in a:
        return V1
    return VX if VX in a else V0


def f_not(a):
    a = _kleene(a)
    return {V0: V1, V1: V0}.get(a, VX)


def f_xor(*a):
    if len(a) == 2 and a[0] == VZ and a[1] == VZ:
        return V0  # reference quirk, src/simulate.cpp:545-551
    a = [_kleene(x) for x in a]
    if VX in a:
        return VX
    return V1 if sum(1 for x in a if x == V1) % 2 else V0


def f_mux2(i0, i1, s):
    if s == V0:
        return i0
    if s == V1:
        return i1
    if s == VX:
        return i0 if i0 == i1 else VX
    return VX


CELLS = {
    "INV": (1, lambda a: f_not(a)),
    "BUF": (1, lambda a: _kleene(a)),
    "AND2": (2, f_and),
    "OR2": (2, f_or),
    "NAND2": (2, lambda a, b: f_not(f_and(a, b))),
    "NOR2": (2, lambda a, b: f_not(f_or(a, b))),
    "XOR2": (2, f_xor),
    "XNOR2": (2, lambda a, b: f_not(f_xor(a, b))),
    "MUX2": (3, f_mux2),
    "AOI21": (3, lambda a1, a2, b: f_not(f_or(f_and(a1, a2), b))),
}


def edge_class(pre, cur):
    """1 rise, 2 fall, 3 X->Z, 4 Z->X, 0 none (src/simulate.cpp:1178-1201)."""
    if pre == cur:
        return 0
    if cur == V1 or pre == V0:
        return 1
    if cur == V0 or pre == V1:
        return 2
    return 3 if (pre, cur) == (VX, VZ) else 4


def sel(rise, fall, s):
    return {1: rise, 2: fall, 3: max(rise, fall), 4: min(rise, fall)}[s]


def gate_wave(cell, inputs, delays):
    """inputs: list of waveforms [(t, v)] (distinct nets, strictly increasing t).
    delays: list of (input_index, edge, rise, fall) with edge 0 pos, 1 neg, 2 both.
    Returns the output waveform."""
    fn = CELLS[cell][1]
    times = sorted({t for w in inputs for t, _ in w})

    def at(w, t):  # (value at t, value before t, has entry at t)
        cur = prev = VX
        has = False
        for tt, v in w:
            if tt < t:
                cur = prev = v
            elif tt == t:
                cur, has = v, True
            else:
                break
        return cur, prev, has

    cands = []  # (o, value) in event order
    f_prev = VX
    for t in times:
        cur, prev, has = zip(*[at(w, t) for w in inputs])
        f = fn(*cur)
        s = edge_class(f_prev, f)
        f_prev = f
        if s == 0:
            continue
        active = [j for j in range(len(inputs)) if has[j]]
        rel = [d for d in delays if d[0] in active]

        def in_edge(j):
            e = edge_class(prev[j], cur[j])
            return 1 if e == 2 else 0

        if not rel:
            d = 1
        elif len(active) == 1:
            if len(rel) == 1:
                d = sel(rel[0][2], rel[0][3], s)
            else:
                e = in_edge(rel[0][0])
                m = [x for x in rel if x[1] == e]
                d = sel(m[0][2], m[0][3], s) if m else 1
        else:
            vals = [sel(x[2], x[3], s) for x in rel if x[1] == 2 or x[1] == in_edge(x[0])]
            d = min(vals) if vals else 1
        cands.append((t + d, f))
    cands = [(0, VX)] + cands
    # strict suffix-min filter
    kept, best = [], None
    for o, v in reversed(cands):
        if best is None or o < best:
            kept.append((o, v))
            best = o if best is None else min(best, o)
    kept.reverse()
    out = []
    for o, v in kept:
        if not out or out[-1][1] != v:
            out.append((o, v))
    return out


def saif(wave, start, end):
    """Closed form of the SAIF durations: every entry holds its value until
    the next entry; X before the first entry; clipped to [start, end)."""
    acc = {V0: 0, V1: 0, VX: 0}
    pts = [(0, VX)] + list(wave) if not wave or wave[0][0] > 0 else list(wave)
    for i, (t, v) in enumerate(pts):
        nxt = pts[i + 1][0] if i + 1 < len(pts) else max(end, t)
        lo, hi = max(t, start), min(nxt, end)
        if hi > lo:
            acc[VX if v == VZ else v] += hi - lo
    return acc[V0], acc[V1], acc[VX]
