"""GPU: the 600M-row variant of the multi-way join configuration, streamed.

Customer 15M, orders 150M (resident in HBM), part 20M, partsupp 80M, supplier
1M, and a fact table of about 600M rows (1..7 per order) held in pinned host
memory. Q9 and Q3 stream the fact table through the engine's double-buffered
chunk path (reference runQ9BenchmarkSF100 src/q9.cpp:245-557,
runQ3BenchmarkSF100 src/q3.cpp:177-538, chunkedStreamLoop
src/infra.h:816-869); both results must equal the oracle over the same rows.
"""
import os
import time

import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import Engine, Timing, generate

pytestmark = pytest.mark.gpu

NC, NO, NP, NS = 15_000_000, 150_000_000, 20_000_000, 1_000_000
SEED = 11
THREADS = max(1, min(len(os.sched_getaffinity(0)), 16))


def say(msg):
    print(f"[sf100] {msg}", flush=True)


@pytest.fixture(scope="module")
def big():
    eng = Engine(0)
    try:
        t0 = time.time()
        fact = generate.lineitem_joined(SEED, NO, NP, NS, empty=eng.pinned_empty)
        say(f"fact {len(fact['l_orderkey'])} rows in pinned memory ({time.time() - t0:.1f} s)")
        js = dict(customer=generate.customer(SEED, NC), orders=generate.orders(SEED, NO, NC),
                  part=generate.part(SEED, NP), partsupp=generate.partsupp(SEED, NP, NS),
                  supplier=generate.supplier(SEED, NS), lineitem=fact)
        dev = {k: eng.table(v) for k, v in js.items() if k != "lineitem"}
        host = eng.host_table(fact)
        say(f"tables ready ({time.time() - t0:.1f} s)")
        yield eng, js, dev, host
    finally:
        eng.close()


def test_q9_600m_streamed_equals_oracle(big):
    eng, js, dev, host = big
    n = len(js["lineitem"]["l_orderkey"])
    assert n > 590_000_000
    tm = Timing()
    say("q9 streamed: start")
    got = eng.q9(dev["part"], dev["supplier"], dev["partsupp"], dev["orders"], host, "green",
                 timing=tm)
    say(f"q9 streamed: {tm.chunks} chunks, {tm.gpu_ms:.1f} ms")
    assert tm.chunks > 1 and tm.h2d_bytes == 28 * n
    exp = O.q9(js["part"], js["supplier"], js["partsupp"], js["orders"], js["lineitem"], "green",
               nthreads=THREADS)
    say("q9 oracle done")
    assert len(got) == 25 * 7 and got == exp


def test_q3_600m_streamed_equals_oracle(big):
    eng, js, dev, host = big
    n = len(js["lineitem"]["l_orderkey"])
    tm = Timing()
    say("q3 streamed: start")
    got = eng.q3(dev["customer"], dev["orders"], host, 1, 1169, 10, timing=tm)
    say(f"q3 streamed: {tm.chunks} chunks, {tm.gpu_ms:.1f} ms")
    assert tm.chunks > 1 and tm.h2d_bytes == 20 * n
    exp = O.q3(js["customer"], js["orders"], js["lineitem"], 1, 1169, 10, nthreads=THREADS)
    say("q3 oracle done")
    assert got == exp and len(got["rows"]) == 10
