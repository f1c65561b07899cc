"""GPU: the pipelines against the oracle on random tables drawn by hypothesis.

Each example builds tables with randtables.py from a drawn seed and shape
(empty and one-row tables, sizes that end mid-tile, sparse / negative /
extreme / permuted / whole-range keys, orphan foreign keys, every row tied,
group codes outside the TPC-H sets, text with bytes after its NUL) and
parameters, uploads them, and requires the engine's answer to equal the
oracle's bit for bit. Inputs outside a pipeline's documented domain must be
refused with GDB_E_UNSUPPORTED instead (Q9 and Q13 take non-negative keys).
The oracle is pinned on the same kind of tables against sqlite3 by
test_property_oracle.py. Examples are derandomized, so every run checks the
same cases.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

import pyoracle as O
import randtables as R
from gpudbmetalbenchmark_amd import GdbError
from gpudbmetalbenchmark_amd.abi import GDB_E_UNSUPPORTED

pytestmark = pytest.mark.gpu

SETTINGS = settings(max_examples=60, deadline=None, derandomize=True, database=None,
                    suppress_health_check=list(HealthCheck))
seeds = st.integers(0, 2**32 - 1)
sizes = st.one_of(st.integers(0, 70), st.integers(0, 300_000))


@pytest.fixture(scope="module", autouse=True)
def _release(engine):
    yield
    engine.release_scratch()  # extreme key spans grow large direct maps


def upload(engine, tables):
    return {k: engine.table(v) for k, v in tables.items()}


def free(tables):
    for t in tables.values():
        t.free()


def exact(rows):
    return [{k: v for k, v in r.items() if not k.startswith("avg")} for r in rows]


@SETTINGS
@given(seed=seeds, n=sizes, constant=st.booleans(), wide=st.booleans(),
       date_lo=st.integers(-100, 2700), span=st.integers(0, 800), disc_lo=st.integers(-1, 11),
       disc_w=st.integers(0, 4), qty_lt=st.integers(0, 130), cutoff=st.integers(-100, 2800))
def test_q6_q1(engine, seed, n, constant, wide, date_lo, span, disc_lo, disc_w, qty_lt, cutoff):
    li = R.flat(seed, n, constant, wide)
    p = dict(date_lo=date_lo, date_hi=date_lo + span, disc_lo=disc_lo, disc_hi=disc_lo + disc_w,
             qty_lt=qty_lt)
    t = engine.table(li)
    try:
        assert engine.q6(t, **p) == O.q6(li, nthreads=8, **p), (n, p)
        got = engine.q1(t, cutoff)
        assert exact(got) == exact(O.q1(li, cutoff, nthreads=8)), (n, cutoff)
        for g in got:  # the averages are the exact sums over the count
            assert g["avg_qty"] == g["sum_qty"] / g["count"]
    finally:
        t.free()


@SETTINGS
@given(seed=seeds, nc=st.one_of(st.integers(0, 5), st.integers(0, 20_000)),
       no=st.one_of(st.integers(0, 50), st.integers(0, 60_000)),
       cmode=st.sampled_from(R.KEY_MODES), omode=st.sampled_from(R.KEY_MODES),
       segment=st.integers(0, 5), date=st.integers(-10, 2600), k=st.integers(1, 100))
def test_q3(engine, seed, nc, no, cmode, omode, segment, date, k):
    js = R.joined(seed, nc, no, 5, 4, cmode, omode, text=False)
    tabs = {n: js[n] for n in ("customer", "orders", "lineitem")}
    t = upload(engine, tabs)
    try:
        got = engine.q3(t["customer"], t["orders"], t["lineitem"], segment, date, k)
        assert got == O.q3(tabs["customer"], tabs["orders"], tabs["lineitem"], segment, date, k,
                           nthreads=8), (nc, no, cmode, omode)
    finally:
        free(t)


@SETTINGS
@given(seed=seeds, no=st.one_of(st.integers(0, 40), st.integers(0, 40_000)),
       nparts=st.one_of(st.integers(0, 10), st.integers(0, 20_000)), nsupp=st.integers(1, 300),
       omode=st.sampled_from(R.KEY_MODES),
       pattern=st.sampled_from(["green", "gr", "een", "n g", "lime", "zzz", "e", "ivory",
                                "greenish", "forest lime"]))
def test_q9(engine, seed, no, nparts, nsupp, omode, pattern):
    js = R.joined(seed, 10, no, nparts, nsupp, "dense", omode)
    names = ("part", "supplier", "partsupp", "orders", "lineitem")
    t = upload(engine, {n: js[n] for n in names})
    try:
        args = [t[n] for n in names]
        if no and js["orders"]["o_orderkey"].min() < 0:
            with pytest.raises(GdbError) as ei:
                engine.q9(*args, pattern)
            assert ei.value.code == GDB_E_UNSUPPORTED
            return
        got = engine.q9(*args, pattern)
        assert got == O.q9(*[js[n] for n in names], pattern, nthreads=8), (no, nparts, omode)
    finally:
        free(t)


@SETTINGS
@given(seed=seeds, nc=st.one_of(st.integers(0, 5), st.integers(0, 30_000)),
       no=st.one_of(st.integers(0, 50), st.integers(0, 60_000)),
       cmode=st.sampled_from(R.KEY_MODES),
       words=st.sampled_from([("special", "requests"), ("spec", "ial"), ("s", "s"),
                              ("requests", "special"), ("re", "quests"), ("pending", "sp"),
                              ("furiously", "deposits"), ("special requests", "pending")]))
def test_q13(engine, seed, nc, no, cmode, words):
    js = R.joined(seed, nc, no, 0, 1, cmode, "dense")
    cust = dict(c_custkey=js["customer"]["c_custkey"])
    od = dict(o_custkey=js["orders"]["o_custkey"], o_comment=js["orders"]["o_comment"])
    c, o = engine.table(cust), engine.table(od)
    try:
        if nc and cust["c_custkey"].min() < 0:
            with pytest.raises(GdbError) as ei:
                engine.q13(c, o, *words)
            assert ei.value.code == GDB_E_UNSUPPORTED
            return
        assert engine.q13(c, o, *words) == O.q13(cust, od, *words, nthreads=8), (nc, no, cmode)
    finally:
        c.free()
        o.free()


@SETTINGS
@given(seed=seeds, nb=st.one_of(st.integers(0, 50), st.integers(0, 200_000)),
       npr=st.one_of(st.integers(0, 50), st.integers(0, 400_000)),
       mode=st.sampled_from(R.KEY_MODES), dup=st.integers(1, 4))
def test_join_count(engine, seed, nb, npr, mode, dup):
    rng = np.random.default_rng(seed)
    keys = R._keys(rng, nb, mode)
    build = np.repeat(keys, dup)[:max(nb, 0)] if dup > 1 else keys
    probe = rng.choice(np.concatenate([build, rng.integers(-2**31, 2**31, 64).astype(np.int32)]),
                       npr).astype(np.int32) if npr else np.zeros(0, np.int32)
    b, p = engine.table(dict(k=build)), engine.table(dict(k=probe))
    try:
        assert engine.join_count(b, "k", p, "k") == O.join_count(build, probe), (nb, npr, mode)
    finally:
        b.free()
        p.free()
