"""GPU: host-resident tables streamed in chunks (the reference's SF100
chunkedStreamLoop path, src/infra.h:816-869) give bit-identical results to the
same tables resident in HBM and to the oracle, for every pipeline; misuse is
rejected with an error code."""
import numpy as np
import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import GdbError, Timing, generate
from gpudbmetalbenchmark_amd.abi import GDB_E_INVALID, GDB_E_UNSUPPORTED

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mid():
    return generate.join_set(21, 150_000, 1_500_000, 200_000, 10_000)


def test_q6_q1_streamed_equals_resident(engine):
    li = generate.lineitem_flat(8, 5_000_003)
    res = engine.table(li)
    host = engine.host_table(li)
    engine.set_chunk_rows(1 << 20)
    try:
        tm = Timing()
        assert engine.q6(host, timing=tm) == engine.q6(res) == O.q6(li, nthreads=8)
        assert tm.chunks == 5 and tm.h2d_bytes == 20 * 5_000_003
        assert engine.q1(host, timing=tm) == engine.q1(res)
        assert tm.chunks == 5
    finally:
        engine.set_chunk_rows(0)


def test_q1_streamed_pinned(engine):
    n = 3_000_001
    li = generate.lineitem_flat(4, n, empty=engine.pinned_empty)
    host = engine.host_table(li)
    engine.set_chunk_rows(777 * 1024)
    try:
        got = engine.q1(host)
    finally:
        engine.set_chunk_rows(0)
    exp = O.q1(li, nthreads=8)
    assert [{k: r[k] for k in r if not k.startswith("avg")} for r in got] == \
        [{k: r[k] for k in r if not k.startswith("avg")} for r in exp]


def test_q3_q9_streamed_fact(engine, mid):
    t = {k: engine.table(v) for k, v in mid.items()}
    hli = engine.host_table(mid["lineitem"])
    engine.set_chunk_rows(1 << 20)
    try:
        for seg, date in ((1, 1169), (4, 2000)):
            a = engine.q3(t["customer"], t["orders"], hli, seg, date, 20)
            b = engine.q3(t["customer"], t["orders"], t["lineitem"], seg, date, 20)
            assert a == b
            assert b == O.q3(mid["customer"], mid["orders"], mid["lineitem"], seg, date, 20,
                             nthreads=8)
        a = engine.q9(t["part"], t["supplier"], t["partsupp"], t["orders"], hli, "green")
        b = engine.q9(t["part"], t["supplier"], t["partsupp"], t["orders"], t["lineitem"], "green")
        assert a == b
    finally:
        engine.set_chunk_rows(0)
    assert b == O.q9(mid["part"], mid["supplier"], mid["partsupp"], mid["orders"],
                     mid["lineitem"], "green", nthreads=8)


def test_q13_streamed_orders(engine, mid):
    c = engine.table(mid["customer"])
    o = engine.table(mid["orders"])
    ho = engine.host_table({k: mid["orders"][k] for k in ("o_custkey", "o_comment")})
    engine.set_chunk_rows(300 * 1024)
    try:
        a = engine.q13(c, ho)
    finally:
        engine.set_chunk_rows(0)
    assert a == engine.q13(c, o) == O.q13(mid["customer"], mid["orders"], nthreads=8)


def test_host_tables_rejected_where_residency_is_required(engine, mid):
    hc = engine.host_table(mid["customer"])
    o = engine.table(mid["orders"])
    li = engine.table(mid["lineitem"])
    with pytest.raises(GdbError) as e:
        engine.q3(hc, o, li)
    assert e.value.code == GDB_E_UNSUPPORTED
    hli = engine.host_table({"l_partkey": mid["lineitem"]["l_partkey"]})
    for fn in (lambda: engine.sum(hli, "l_partkey"), lambda: engine.select_lt(hli, "l_partkey", 5)):
        with pytest.raises(GdbError) as e:
            fn()
        assert e.value.code == GDB_E_UNSUPPORTED
    with pytest.raises(GdbError) as e:
        engine.set_chunk_rows(10)
    assert e.value.code == GDB_E_INVALID
