"""GPU: device allocation failures leave nothing behind.

Every device allocation of the library goes through one accounted allocator,
which a test hook can make fail at the k-th call. For table creation and for
each pipeline, the failure is injected at every allocation the call makes in
turn: the call must return GDB_E_NOMEM, the context must hold no more HBM than
its tables and its scratch arenas (released here after each failure, back to
the bytes held before the call), and the same call without the failure must
then give the oracle's result. The reference has no counterpart: its Metal
buffer allocations are not checked (src/q6.cpp:42-47)."""
import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import Engine, GdbError, generate, lib
from gpudbmetalbenchmark_amd.abi import GDB_E_NOMEM

pytestmark = pytest.mark.gpu

MAX_ALLOCS = 48  # more than any pipeline here makes


def fail_after(e, skip, count=1):
    lib().gdb_debug_fail_allocs_after(e._ctx, skip, count)


@pytest.fixture(scope="module")
def ctx():
    with Engine(0) as e:
        e.release_scratch()
        yield e


@pytest.fixture(scope="module")
def data():
    li = generate.lineitem_flat(31, 300_001)
    js = generate.join_set(29, 15_000, 150_000, 20_000, 1_000)
    return li, js


def test_table_create_failure_at_every_column(ctx, data):
    li, _ = data
    base = ctx.hbm_used()
    for k in range(len(li)):
        fail_after(ctx, k)
        try:
            with pytest.raises(GdbError) as ei:
                ctx.table(li)
        finally:
            fail_after(ctx, 0, 0)
        assert ei.value.code == GDB_E_NOMEM, k
        assert ctx.hbm_used() == base, k
    t = ctx.table(li)
    assert ctx.hbm_used() > base
    t.free()
    assert ctx.hbm_used() == base


def exact(rows):
    """Q1 rows without the averages (floating point, checked elsewhere)."""
    return [{k: v for k, v in r.items() if not k.startswith("avg")} for r in rows]


def _pipelines(ctx, li, js, tabs, host):
    c, o, l, p, s, ps = (tabs[k] for k in ("customer", "orders", "lineitem", "part", "supplier",
                                           "partsupp"))
    return [
        ("q6", lambda: ctx.q6(tabs["flat"]), lambda: O.q6(li)),
        ("q1", lambda: exact(ctx.q1(tabs["flat"])), lambda: exact(O.q1(li))),
        ("q3", lambda: ctx.q3(c, o, l, 1, 1169, 10),
         lambda: O.q3(js["customer"], js["orders"], js["lineitem"], 1, 1169, 10)),
        ("q9", lambda: ctx.q9(p, s, ps, o, l, "green"),
         lambda: O.q9(js["part"], js["supplier"], js["partsupp"], js["orders"], js["lineitem"],
                      "green")),
        ("q13", lambda: ctx.q13(c, o), lambda: O.q13(js["customer"], js["orders"])),
        ("join", lambda: ctx.join_count(o, "o_orderkey", l, "l_orderkey"),
         lambda: O.join_count(js["orders"]["o_orderkey"], js["lineitem"]["l_orderkey"])),
        ("q6_streamed", lambda: ctx.q6(host), lambda: O.q6(li)),
    ]


def test_pipeline_failure_at_every_allocation(ctx, data):
    li, js = data
    tabs = {k: ctx.table(v) for k, v in js.items()}
    tabs["flat"] = ctx.table(li)
    host = ctx.host_table(li)
    ctx.set_chunk_rows(64 * 1024)
    try:
        for name, run, expect in _pipelines(ctx, li, js, tabs, host):
            want = expect()
            ctx.release_scratch()
            base = ctx.hbm_used()
            failures = 0
            for k in range(MAX_ALLOCS):
                fail_after(ctx, k)
                try:
                    got = run()
                except GdbError as ex:
                    assert ex.code == GDB_E_NOMEM, (name, k, str(ex))
                    failures += 1
                    ctx.release_scratch()
                    assert ctx.hbm_used() == base, (name, k)
                    continue
                finally:
                    fail_after(ctx, 0, 0)
                # k allocations sufficed: the call made exactly k of them
                assert got == want, (name, k)
                break
            else:
                pytest.fail(f"{name}: still failing after {MAX_ALLOCS} injected failures")
            assert failures >= 1, name  # every pipeline allocates scratch on a fresh context
            print(f"[alloc-fail] {name}: {failures} allocation points", flush=True)
            # the arenas the successful call grew are kept and reused
            held = ctx.hbm_used()
            assert run() == want and ctx.hbm_used() == held, name
            ctx.release_scratch()
            assert ctx.hbm_used() == base, name
    finally:
        ctx.set_chunk_rows(0)
        for t in tabs.values():
            t.free()
    ctx.release_scratch()
