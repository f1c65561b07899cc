"""CPU: pins the oracle (oracle/oracle.c) and the table generator.

The expected values in golden.json come from sqlite3 running the reference's
SQL on tables built by an independent numpy restatement of the generator
(tests/golden/make_golden.py); the digests pin libgdb.so's generator to it.
"""
import datetime
import hashlib
import re

import numpy as np
import pytest

import pyoracle as O


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def np_cols(d):
    out = {}
    for k, v in d.items():
        if k in ("l_extendedprice", "ps_supplycost"):
            out[k] = np.array(v, np.int64)
        elif k in ("l_returnflag", "l_linestatus", "c_mktsegment", "o_comment", "p_name"):
            out[k] = np.array(v, np.uint8)
        else:
            out[k] = np.array(v, np.int32)
    return out


# ------------------------------------------------------------ generator pins
def test_generator_flat_digests(golden, flat):
    for k, v in flat.items():
        assert digest(v) == golden["flat"]["digests"][k], k


def test_generator_join_digests(golden, joinset):
    for t in joinset.values():
        for k, v in t.items():
            assert digest(v) == golden["join"]["digests"][k], k
    assert len(joinset["lineitem"]["l_orderkey"]) == golden["join"]["lineitem_rows"]


def test_generator_other_seed_matches_numpy():
    import npgen
    from gpudbmetalbenchmark_amd import generate
    a = generate.lineitem_flat(1234, 7777, row0=99, receipt=True)
    b = npgen.lineitem_flat(1234, 7777, row0=99)
    for k in b:
        assert np.array_equal(a[k], b[k]), k
    a = generate.lineitem_joined(99, 333, 50, 8)
    b = npgen.lineitem_joined(99, 333, 50, 8)
    for k in b:
        assert np.array_equal(a[k], b[k]), k


def test_generator_distributions(flat, joinset):
    s = flat["l_shipdate"]
    assert s.min() == 0 and s.max() == 2525
    assert flat["l_quantity"].min() == 1 and flat["l_quantity"].max() == 50
    assert flat["l_discount"].min() == 0 and flat["l_discount"].max() == 10
    retail = flat["l_extendedprice"] // flat["l_quantity"]
    assert np.all(flat["l_extendedprice"] % flat["l_quantity"] == 0)
    assert retail.min() >= 90000 and retail.max() <= 209900
    # linestatus / returnflag rules of the flat fact table
    ls, rf, rc = flat["l_linestatus"], flat["l_returnflag"], flat["l_receiptdate"]
    assert np.all((ls == ord("F")) == (s < 1263))
    assert np.all((rf == ord("N")) == (rc >= 1263))
    assert np.all((rc - s >= 1) & (rc - s <= 30))
    # joined fact: shipdate = orderdate + 1..121, 1..7 lines per order
    o, li = joinset["orders"], joinset["lineitem"]
    date_of = dict(zip(o["o_orderkey"].tolist(), o["o_orderdate"].tolist()))
    od = np.array([date_of[k] for k in li["l_orderkey"].tolist()])
    d = li["l_shipdate"] - od
    assert d.min() >= 1 and d.max() <= 121
    _, per = np.unique(li["l_orderkey"], return_counts=True)
    assert per.min() >= 1 and per.max() <= 7
    # partsupp: (partkey, suppkey) unique; fact rows use one of the part's suppliers
    ps = joinset["partsupp"]
    pairs = set(zip(ps["ps_partkey"].tolist(), ps["ps_suppkey"].tolist()))
    assert len(pairs) == len(ps["ps_partkey"])
    assert all(p in pairs for p in zip(li["l_partkey"].tolist(), li["l_suppkey"].tolist()))


def test_part_name_green_rate():
    from gpudbmetalbenchmark_amd import generate
    p = generate.part(3, 20000)
    names = [bytes(r).split(b"\0")[0].decode() for r in p["p_name"]]
    assert all(len(n.split()) == 5 and len(set(n.split())) == 5 for n in names)
    rate = sum("green" in n for n in names) / len(names)
    assert abs(rate - 5 / 92) < 0.006  # 5 distinct words of 92


def test_comment_match_rate():
    from gpudbmetalbenchmark_amd import generate
    o = generate.orders(5, 20000, 100, comments=True)
    texts = [bytes(r).split(b"\0")[0].decode() for r in o["o_comment"]]
    rate = sum(bool(re.search("special.*requests", t)) for t in texts) / len(texts)
    assert 0.004 < rate < 0.03
    lens = [len(t) for t in texts]
    assert min(lens) >= 20 and max(lens) == 80


# --------------------------------------------------------------- oracle pins
def test_oracle_q6_golden(golden, flat):
    for case in golden["q6"]:
        r = O.q6(flat, **case["params"], nthreads=4)
        assert r == dict(revenue=case["revenue"], rows=case["rows"]), case["params"]


def _q1_cmp(got, exp):
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        for k in ("returnflag", "linestatus", "sum_qty", "sum_base_price", "sum_disc_price",
                  "sum_disc", "count"):
            assert g[k] == e[k], k
        assert g["sum_charge"] == int(e["sum_charge"])
        for k in ("avg_qty", "avg_price", "avg_disc"):
            assert g[k] == pytest.approx(e[k], rel=1e-6, abs=0), k


def test_oracle_q1_golden(golden, flat):
    for case in golden["q1"]:
        _q1_cmp(O.q1(flat, case["cutoff"], nthreads=3), case["rows"])


def test_oracle_q3_golden(golden, joinset):
    j = joinset
    for case in golden["q3"]:
        r = O.q3(j["customer"], j["orders"], j["lineitem"], case["segment"], case["date"],
                 case["top_k"], nthreads=4)
        assert r["groups"] == case["groups"]
        assert r["rows"] == case["rows"]


def test_oracle_q9_golden(golden, joinset):
    j = joinset
    for case in golden["q9"]:
        r = O.q9(j["part"], j["supplier"], j["partsupp"], j["orders"], j["lineitem"],
                 case["pattern"], nthreads=4)
        assert r == case["rows"], case["pattern"]


def test_oracle_q13_golden(golden, joinset):
    j = joinset
    for case in golden["q13"]:
        r = O.q13(j["customer"], j["orders"], case["word1"], case["word2"], nthreads=4)
        assert r == case["rows"]


def test_oracle_primitives_golden(golden, joinset):
    li, od = joinset["lineitem"], joinset["orders"]
    pr = golden["primitives"]
    for c in pr["select"]:
        assert O.select_lt(li[c["column"]], c["value"])["count"] == c["count"]
    for c in pr["sum"]:
        assert O.sum_col(li[c["column"]]) == c["sum"]
    for c in pr["join"]:
        assert O.join_count(od[c["build"]], li[c["probe"]]) == c["matches"]


def test_oracle_edge_cases(golden):
    e = golden["edge"]
    t = np_cols(e["q6"]["table"])
    for case in e["q6"]["results"]:
        assert O.q6(t, **case["params"]) == dict(revenue=case["revenue"], rows=case["rows"])
    t = np_cols(e["q1"]["table"])
    for case in e["q1"]["results"]:
        _q1_cmp(O.q1(t, case["cutoff"]), case["rows"])
    c, o, li = (np_cols(e["q3"][k]) for k in ("customer", "orders", "lineitem"))
    for case in e["q3"]["results"]:
        r = O.q3(c, o, li, case["segment"], case["date"], case["top_k"])
        assert r["rows"] == case["rows"] and r["groups"] == case["groups"]
    c, o = np_cols(e["q13"]["customer"]), np_cols(e["q13"]["orders"])
    for case in e["q13"]["results"]:
        assert O.q13(c, o, case["word1"], case["word2"]) == case["rows"]


# ------------------------------------------------ closed forms and invariants
def test_q6_closed_forms_and_partition(flat):
    everything = O.q6(flat, 0, 2526, 0, 10, 51)
    assert everything["rows"] == len(flat["l_shipdate"])
    assert everything["revenue"] == int(np.sum(flat["l_extendedprice"] * flat["l_discount"]))
    assert O.q6(flat, 900, 900, 0, 10, 51) == dict(revenue=0, rows=0)
    a, b, c = O.q6(flat, 0, 800), O.q6(flat, 800, 1700), O.q6(flat, 0, 1700)
    assert a["revenue"] + b["revenue"] == c["revenue"] and a["rows"] + b["rows"] == c["rows"]


def test_q1_invariants(flat):
    rows = O.q1(flat, 2435)
    assert sum(r["count"] for r in rows) == int(np.sum(flat["l_shipdate"] <= 2435))
    groups = {(r["returnflag"], r["linestatus"]) for r in rows}
    assert ("A", "O") not in groups and ("R", "O") not in groups
    for r in rows:  # SUM(charge) between 100x and 108x SUM(disc_price) for tax 0..8
        assert 100 * r["sum_disc_price"] <= r["sum_charge"] <= 108 * r["sum_disc_price"]


def test_oracle_thread_invariance(flat, joinset):
    assert O.q6(flat, nthreads=1) == O.q6(flat, nthreads=7)
    assert O.q1(flat, nthreads=1) == O.q1(flat, nthreads=5)
    j = joinset
    assert (O.q3(j["customer"], j["orders"], j["lineitem"], nthreads=1)
            == O.q3(j["customer"], j["orders"], j["lineitem"], nthreads=6))


def test_year_of_day_matches_datetime():
    d0 = datetime.date(1992, 1, 1)
    for day in range(-400, 4000, 7):
        assert O.year_of_day(day) == (d0 + datetime.timedelta(days=day)).year, day


def test_substring_helpers_against_python():
    rng = np.random.default_rng(0)
    alphabet = np.frombuffer(b"sperciaqlut ", dtype=np.uint8)
    for _ in range(3000):
        n = int(rng.integers(0, 30))
        s = bytes(rng.choice(alphabet, n))
        field = np.zeros(30, np.uint8)
        field[:n] = np.frombuffer(s, np.uint8) if n else field[:0]
        t = s.decode()
        assert O.contains(field, "spec") == ("spec" in t)
        assert O.like_two(field, "sp", "la") == bool(re.search("sp.*la", t))
        assert O.like_two(field, "aa", "aa") == bool(re.search("aa.*aa", t))
        # bytes after the terminating NUL are not part of the string
        if n < 29:
            junk = field.copy()
            junk[n + 1:] = np.frombuffer(b"splaaspecla" * 3, np.uint8)[:29 - n]
            assert O.contains(junk, "spec") == ("spec" in t)
            assert O.like_two(junk, "sp", "la") == bool(re.search("sp.*la", t))


def extreme_q1_table(n, seed=3):
    """Rows at the edge of Q1's exact domain: extendedprice up to 2^32-1,
    tax up to 100, discount 0..100, quantity up to 127, every group."""
    rng = np.random.default_rng(seed)
    return dict(
        l_shipdate=rng.integers(0, 2000, n).astype(np.int32),
        l_quantity=rng.integers(120, 128, n).astype(np.int32),
        l_discount=rng.choice(np.array([0, 0, 0, 100, 37], np.int32), n),
        l_tax=np.where(rng.random(n) < 0.5, 100, rng.integers(0, 101, n)).astype(np.int32),
        l_extendedprice=(2**32 - 1 - rng.integers(0, 1000, n)).astype(np.int64),
        l_returnflag=rng.choice(np.frombuffer(b"ANR", np.uint8), n),
        l_linestatus=rng.choice(np.frombuffer(b"FO", np.uint8), n))


def q1_bigint(li, cutoff):
    """Q1 sums with Python integers (no overflow possible)."""
    out = {}
    for i in np.nonzero(li["l_shipdate"] <= cutoff)[0]:
        key = (chr(li["l_returnflag"][i]), chr(li["l_linestatus"][i]))
        p, d, t = int(li["l_extendedprice"][i]), int(li["l_discount"][i]), int(li["l_tax"][i])
        g = out.setdefault(key, [0, 0, 0, 0, 0, 0])
        g[0] += int(li["l_quantity"][i])
        g[1] += p
        g[2] += p * (100 - d)
        g[3] += p * (100 - d) * (100 + t)
        g[4] += d
        g[5] += 1
    return out


def test_oracle_q1_extreme_values_exact_or_refused():
    """Near the int64 limit the oracle's sums must still equal Python's
    integer arithmetic; past its bound (rows x max|price| x (100 + max|disc|)
    >= 2^63) it must refuse rather than wrap."""
    li = extreme_q1_table(40_000)
    exp = q1_bigint(li, 1500)
    got = O.q1(li, 1500, nthreads=4)
    assert len(got) == len(exp)
    for r in got:
        e = exp[(r["returnflag"], r["linestatus"])]
        assert [r["sum_qty"], r["sum_base_price"], r["sum_disc_price"], r["sum_charge"],
                r["sum_disc"], r["count"]] == e
    big = dict(li)
    big["l_extendedprice"] = np.full(len(li["l_shipdate"]), 2**55, np.int64)
    with pytest.raises(OverflowError):
        O.q1({k: v[:3] for k, v in big.items()}, 1500)
