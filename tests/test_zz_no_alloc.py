"""The allocation-free execution (gls_run_execute_static) against the
ordinary execution (gls_run_execute): same SAIF text and gate events, no
device allocation, two calls in a row (the second on a smaller input), and
a time-chunked run whose chunks shrink to fit a fixed pool."""
import sys

import pytest

import gls
import golden_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = gls.Context(0)
    yield c
    c.close()


def compare(ctx, name, **opts):
    d, s = golden_cases.make(name)
    on, off = golden_cases.CASES[name]["window"]
    ref = ctx.simulate(d, s, on, off, **opts)
    want = (ref.saif_text(), ref.stats()["n_events"])
    ref.close()
    r = ctx.prepare(d, s, on, off, **opts)
    r.execute_static()
    st = r.stats()
    say("static", name, opts, st["chunks"], st["device_allocs"])
    assert st["device_allocs"] == 0
    assert (r.saif_text(), st["n_events"]) == want
    return r, want


def test_static_execution_matches_execute(ctx):
    # two calls in a row on one GPU's array, the second with a smaller input
    r1, _ = compare(ctx, "c1_small")
    r2, _ = compare(ctx, "c0_small")
    # the first run again, after the array held the second one's data
    d, s = golden_cases.make("c1_small")
    on, off = golden_cases.CASES["c1_small"]["window"]
    again = ctx.simulate(d, s, on, off).saif_text()
    r1.execute_static()
    assert r1.saif_text() == again and r1.stats()["device_allocs"] == 0


def test_static_execution_in_time_chunks(ctx):
    # a fixed 3 MiB pool: the run is cut into chunks that fit it, with no
    # allocation; and fixed chunk lengths
    compare(ctx, "c1_small", pool_bytes=3 << 20)
    compare(ctx, "c0_small", chunk_ps=7777)


def test_static_execution_refuses_what_needs_allocation(ctx):
    d, s = golden_cases.make("c0_small")
    r = ctx.prepare(d, s, 0, 200001, keep_waves=True)
    with pytest.raises(gls.GlsError) as e:
        r.execute_static()
    assert e.value.status == "GLS_E_UNSUPPORTED"
    r = ctx.prepare(d, s, 0, 200001, pool_bytes=4096, chunk_ps=gls.NO_CHUNKS)
    with pytest.raises(gls.GlsError) as e:
        r.execute_static()
    assert e.value.status == "GLS_E_OOM"
    # the ordinary execution still works after the refusals
    ok = ctx.simulate(d, s, 0, 200001)
    assert ok.stats()["n_events"] > 0
