"""CPU: the TPC-H oracle (oracle/tpch_oracle.py) against sqlite3's answers
(tests/golden/tpch_golden.json, written by tests/golden/make_tpch_golden.py),
plus generator invariants and closed forms the specification fixes."""
import datetime
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import tpch_oracle as T  # noqa: E402
from gpudbmetalbenchmark_amd import tpch  # noqa: E402

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "tpch_golden.json")))


@pytest.fixture(scope="module", params=range(len(GOLD)), ids=lambda i: f"case{i}")
def case(request):
    g = GOLD[request.param]
    return tpch.generate(g["sf"], g["seed"]), g["results"]


def L(rows):
    return [list(r) if isinstance(r, tuple) else r for r in rows]


def test_oracle_matches_sqlite(case):
    db, gold = case
    brass = tpch.codes_like(tpch.P_TYPE, "%BRASS")
    mp = tpch.codes_like(tpch.P_TYPE, "MEDIUM POLISHED%")
    assert L(T.q2(db, 15, brass, 3)) == gold["q2"]
    assert L(T.q4(db)) == gold["q4"]
    assert L(T.q5(db)) == gold["q5"]
    assert L(T.q7(db)) == gold["q7"]
    assert L(T.q8(db)) == gold["q8"]
    rows, groups = T.q10(db)
    assert L(rows) == gold["q10"] and groups == gold["q10_groups"][0][0]
    rows, total = T.q11(db)
    assert L(rows) == gold["q11"] and total == gold["q11_total"][0][0]
    assert L(T.q12(db)) == gold["q12"]
    assert list(T.q14(db)) == gold["q14"][0]
    assert L(T.q15(db)) == gold["q15"]
    assert L(T.q16(db, 45, mp)) == gold["q16"]
    assert T.q17(db) == gold["q17"][0][0]
    assert L(T.q18(db)) == gold["q18"]
    assert L(T.q18(db, 250)) == gold["q18_250"]
    assert T.q19(db) == gold["q19"][0][0]
    assert T.q20(db) == gold["q20"]
    assert L(T.q21(db)) == gold["q21"]
    assert L(T.q22(db)) == gold["q22"]


def test_goldens_are_not_trivial():
    # every query has rows in at least one case, so a vacuous oracle fails
    for name in ("q2", "q4", "q5", "q7", "q8", "q10", "q11", "q12", "q15", "q16", "q18_250",
                 "q20", "q21", "q22"):
        assert any(g["results"][name] for g in GOLD), name
    for name in ("q14", "q17", "q19"):
        assert any(g["results"][name][0][0] for g in GOLD), name


def test_calendar_years():
    days = np.arange(0, 2600)
    exp = [(datetime.date(1992, 1, 1) + datetime.timedelta(days=int(d))).year for d in days]
    assert T.year_of(days).tolist() == exp


def test_generator_invariants():
    db = tpch.generate(0.01, 7)
    li, o, ps, p = db["lineitem"], db["orders"], db["partsupp"], db["part"]
    # lines are clustered by order, 1..7 per order, in order-key order
    ok = li["l_orderkey"]
    assert np.all(np.diff(ok) >= 0)
    _, cnt = np.unique(ok, return_counts=True)
    assert cnt.min() >= 1 and cnt.max() <= 7 and cnt.size == o["o_orderkey"].size
    # every (partkey, suppkey) of a line is a partsupp row
    pairs = set(zip(ps["ps_partkey"].tolist(), ps["ps_suppkey"].tolist()))
    assert all(x in pairs for x in zip(li["l_partkey"][:5000].tolist(),
                                       li["l_suppkey"][:5000].tolist()))
    # four distinct suppliers per part
    assert all(len(set(ps["ps_suppkey"][4 * i:4 * i + 4])) == 4 for i in range(p["p_partkey"].size))
    # no customer key divisible by 3 places an order
    assert not np.any(o["o_custkey"] % 3 == 0)
    # extendedprice = quantity * retailprice; date orderings of the spec
    rp = tpch.retail_price(li["l_partkey"])
    assert np.array_equal(li["l_extendedprice"], li["l_quantity"] * rp)
    assert np.all(li["l_receiptdate"] > li["l_shipdate"])
    # status 'F' iff all lines shipped by the current date
    f = o["o_orderstatus"] == ord("F")
    late = np.zeros(int(ok.max()) + 1, bool)
    late[ok[li["l_linestatus"] == ord("O")]] = True
    assert np.array_equal(f, ~late[o["o_orderkey"]] & (o["o_orderstatus"] != ord("P")))
    # planted supplier comments
    com = [tpch.text(c) for c in db["supplier"]["s_comment"]]
    assert sum("Complaints" in c and c.index("Customer") < c.index("Complaints")
               for c in com if "Customer" in c) >= 2


def test_closed_forms():
    db = tpch.generate(0.01, 7)
    li = db["lineitem"]
    rev = li["l_extendedprice"] * (100 - li["l_discount"].astype(np.int64))
    # Q14 over all dates and every type: promo == total == sum of all revenue
    promo, total = T.q14(db, -1, 10**6, list(range(150)))
    assert promo == total == int(rev.sum())
    # Q14 with no type: promo 0; adjacent windows partition the total
    assert T.q14(db, 0, 5000, [])[0] == 0
    a, b = T.q14(db, 0, 1000)[1], T.q14(db, 1000, 5000)[1]
    assert a + b == int(rev.sum())
    # Q4 over all dates counts every order with a late line
    late_orders = np.unique(li["l_orderkey"][li["l_commitdate"] < li["l_receiptdate"]]).size
    assert sum(c for _, c in T.q4(db, -1, 10**6)) == late_orders
    # Q15 over all dates: the top supplier's revenue is the largest per-supplier sum
    acc = np.zeros(int(li["l_suppkey"].max()) + 1, np.int64)
    np.add.at(acc, li["l_suppkey"], rev)
    assert T.q15(db, -1, 10**6)[0][1] == int(acc.max())
    # Q18 with threshold 0 returns the k largest totalprices
    tp = np.sort(db["orders"]["o_totalprice"])[::-1][:100]
    assert [r[3] for r in T.q18(db, 0)] == tp.tolist()
    # Q11 with fraction 0 keeps every German part; their values sum to the total
    rows, total = T.q11(db, 7, 0, 1)
    assert sum(v for _, v in rows) == total
