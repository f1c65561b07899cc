"""CPU: the oracle against sqlite3 on random tables drawn by hypothesis.

The golden vectors pin the oracle on the seeded generator's tables; these
properties pin it on tables that generator never makes (randtables.py):
empty and one-row tables, sparse / negative / extreme / permuted keys, orphan
foreign keys, group codes outside the TPC-H sets, every row tied, text with
bytes after its NUL. sqlite3 runs the reference's SQL (sql/q1.sql, q3.sql,
q6.sql, q9.sql, q13.sql, restated in tests/golden/make_golden.py) in exact
integer arithmetic, so every result must be equal.
"""
import sqlite3

from hypothesis import HealthCheck, given, settings, strategies as st

import make_golden as G
import pyoracle as O
import randtables as R

SETTINGS = settings(max_examples=40, deadline=None, derandomize=True, database=None,
                    suppress_health_check=[HealthCheck.too_slow])
seeds = st.integers(0, 2**32 - 1)


def db_of(tables):
    db = sqlite3.connect(":memory:")
    for name, cols in tables.items():
        G.load(db, name, cols)
    return db


@SETTINGS
@given(seed=seeds, n=st.integers(0, 300), constant=st.booleans(), wide=st.booleans(),
       date_lo=st.integers(-100, 2700), span=st.integers(0, 800), disc_lo=st.integers(-1, 11),
       disc_w=st.integers(0, 4), qty_lt=st.integers(0, 130))
def test_q6_equals_sql(seed, n, constant, wide, date_lo, span, disc_lo, disc_w, qty_lt):
    li = R.flat(seed, n, constant, wide)
    p = dict(date_lo=date_lo, date_hi=date_lo + span, disc_lo=disc_lo, disc_hi=disc_lo + disc_w,
             qty_lt=qty_lt)
    exp = G.q6(db_of(dict(lineitem=li)), p)
    assert O.q6(li, **p) == dict(revenue=exp["revenue"], rows=exp["rows"])


@SETTINGS
@given(seed=seeds, n=st.integers(0, 300), constant=st.booleans(), wide=st.booleans(),
       cutoff=st.integers(-100, 2800))
def test_q1_equals_sql(seed, n, constant, wide, cutoff):
    li = R.flat(seed, n, constant, wide)
    exp = G.q1(db_of(dict(lineitem=li)), cutoff)["rows"]
    got = O.q1(li, cutoff)
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        for k in ("returnflag", "linestatus", "sum_qty", "sum_base_price", "sum_disc_price",
                  "sum_disc", "count"):
            assert g[k] == e[k], k
        assert g["sum_charge"] == int(e["sum_charge"])


@SETTINGS
@given(seed=seeds, nc=st.integers(0, 40), no=st.integers(0, 150),
       cmode=st.sampled_from(R.KEY_MODES), omode=st.sampled_from(R.KEY_MODES),
       segment=st.integers(0, 5), date=st.integers(-10, 2600), k=st.integers(1, 100))
def test_q3_equals_sql(seed, nc, no, cmode, omode, segment, date, k):
    js = R.joined(seed, nc, no, 5, 4, cmode, omode)
    t = {n: js[n] for n in ("customer", "orders", "lineitem")}
    exp = G.q3(db_of(t), segment, date, k)
    got = O.q3(t["customer"], t["orders"], t["lineitem"], segment, date, k)
    assert got == dict(groups=exp["groups"], rows=exp["rows"])


@SETTINGS
@given(seed=seeds, no=st.integers(0, 120), nparts=st.integers(0, 30), nsupp=st.integers(1, 9),
       omode=st.sampled_from(R.KEY_MODES),
       pattern=st.sampled_from(["green", "gr", "een", "n g", "lime", "zzz", "e", "ivory"]))
def test_q9_equals_sql(seed, no, nparts, nsupp, omode, pattern):
    js = R.joined(seed, 10, no, nparts, nsupp, "dense", omode)
    t = {n: js[n] for n in ("part", "supplier", "partsupp", "orders", "lineitem")}
    db = db_of(t)
    db.execute("CREATE TABLE nation (n_nationkey, n_name)")
    db.executemany("INSERT INTO nation VALUES (?, ?)", list(enumerate(G.NATIONS)))
    exp = G.q9(db, pattern)["rows"]
    assert O.q9(t["part"], t["supplier"], t["partsupp"], t["orders"], t["lineitem"],
                pattern) == exp


@SETTINGS
@given(seed=seeds, nc=st.integers(0, 60), no=st.integers(0, 200),
       cmode=st.sampled_from(R.KEY_MODES),
       words=st.sampled_from([("special", "requests"), ("spec", "ial"), ("s", "s"),
                              ("requests", "special"), ("re", "quests"), ("pending", "sp"),
                              ("furiously", "deposits")]))
def test_q13_equals_sql(seed, nc, no, cmode, words):
    js = R.joined(seed, nc, no, 0, 1, cmode, "dense")
    t = dict(customer=dict(c_custkey=js["customer"]["c_custkey"]),
             orders=dict(o_orderkey=js["orders"]["o_orderkey"],
                         o_custkey=js["orders"]["o_custkey"], o_comment=js["orders"]["o_comment"]))
    exp = G.q13(db_of(t), *words)["rows"]
    assert O.q13(t["customer"], t["orders"], *words) == exp
