"""GPU parity of the join pipelines (Q3, Q9, Q13) and the join primitive
against the sqlite3 goldens, hand-built edge cases, and the oracle at the
benchmark configurations' full sizes. Integer results must be bit-exact."""
import numpy as np
import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import Timing, generate

pytestmark = pytest.mark.gpu

FULL = dict(seed=11, customers=1_500_000, orders=15_000_000, parts=2_000_000, suppliers=100_000)


def np_cols(d):
    out = {}
    for k, v in d.items():
        if k in ("l_extendedprice", "ps_supplycost"):
            out[k] = np.array(v, np.int64)
        elif k in ("c_mktsegment", "o_comment", "p_name"):
            out[k] = np.array(v, np.uint8)
        else:
            out[k] = np.array(v, np.int32)
    return out


def upload(engine, js):
    return {name: engine.table(cols) for name, cols in js.items()}


@pytest.fixture(scope="module")
def small(engine, joinset):
    t = upload(engine, joinset)
    yield t
    for x in t.values():
        x.free()


@pytest.fixture(scope="module")
def full(engine):
    js = generate.join_set(FULL["seed"], FULL["customers"], FULL["orders"], FULL["parts"],
                           FULL["suppliers"])
    t = upload(engine, js)
    yield js, t
    for x in t.values():
        x.free()


# ------------------------------------------------------------------------ Q3
def test_q3_golden(engine, golden, small):
    for case in golden["q3"]:
        r = engine.q3(small["customer"], small["orders"], small["lineitem"], case["segment"],
                      case["date"], case["top_k"])
        assert r["groups"] == case["groups"]
        assert r["rows"] == case["rows"], (case["segment"], case["date"])


def test_q3_edge_ties(engine, golden):
    e = golden["edge"]["q3"]
    c, o, li = (engine.table(np_cols(e[k])) for k in ("customer", "orders", "lineitem"))
    for case in e["results"]:
        r = engine.q3(c, o, li, case["segment"], case["date"], case["top_k"])
        assert r["rows"] == case["rows"] and r["groups"] == case["groups"]


def test_q3_full_config(engine, full):
    js, t = full
    tm = Timing()
    for seg, date, k in ((1, 1169, 10), (3, 400, 100), (0, 2300, 7)):
        got = engine.q3(t["customer"], t["orders"], t["lineitem"], seg, date, k, timing=tm)
        exp = O.q3(js["customer"], js["orders"], js["lineitem"], seg, date, k, nthreads=16)
        assert got == exp, (seg, date, k)
    again = engine.q3(t["customer"], t["orders"], t["lineitem"], 3, 400, 100)
    assert again == engine.q3(t["customer"], t["orders"], t["lineitem"], 3, 400, 100)
    assert tm.main_kernel_ms > 0


# ------------------------------------------------------------------------ Q9
def test_q9_golden(engine, golden, small):
    s = small
    for case in golden["q9"]:
        r = engine.q9(s["part"], s["supplier"], s["partsupp"], s["orders"], s["lineitem"],
                      case["pattern"])
        assert r == case["rows"], case["pattern"]


def test_q9_full_config(engine, full):
    js, t = full
    for pat in ("green", "ivory"):
        got = engine.q9(t["part"], t["supplier"], t["partsupp"], t["orders"], t["lineitem"], pat)
        exp = O.q9(js["part"], js["supplier"], js["partsupp"], js["orders"], js["lineitem"], pat,
                   nthreads=16)
        assert got == exp, pat
        assert len(got) == 25 * 7


# ----------------------------------------------------------------------- Q13
def test_q13_golden(engine, golden, small):
    for case in golden["q13"]:
        r = engine.q13(small["customer"], small["orders"], case["word1"], case["word2"])
        assert r == case["rows"]


def test_q13_edge_patterns(engine, golden):
    e = golden["edge"]["q13"]
    c, o = engine.table(np_cols(e["customer"])), engine.table(np_cols(e["orders"]))
    for case in e["results"]:
        assert engine.q13(c, o, case["word1"], case["word2"]) == case["rows"]


def test_q13_full_config(engine, full):
    js, t = full
    got = engine.q13(t["customer"], t["orders"])
    assert got == O.q13(js["customer"], js["orders"], nthreads=16)
    assert sum(r["custdist"] for r in got) == FULL["customers"]


def test_q13_random_fields_against_oracle(engine):
    rng = np.random.default_rng(5)
    n = 20000
    alphabet = np.frombuffer(b"specialrqut ", np.uint8)
    com = rng.choice(alphabet, (n, 80)).astype(np.uint8)
    cut = rng.integers(0, 81, n)
    com[np.arange(80)[None, :] >= cut[:, None]] = 0
    cust = dict(c_custkey=np.arange(1, 301, dtype=np.int32))
    od = dict(o_custkey=rng.integers(1, 320, n).astype(np.int32), o_comment=com)
    c, o = engine.table(cust), engine.table(od)
    for w1, w2 in (("spe", "qu"), ("a", "a"), ("l", "t"), ("ecia", "ue")):
        assert engine.q13(c, o, w1, w2) == O.q13(cust, od, w1, w2), (w1, w2)


def test_q13_words_bytes_after_nul_and_long_patterns(engine):
    """Comments made of words, with non-NUL bytes left after the terminating
    NUL (they must be ignored), against patterns of 1 to 31 bytes, overlapping
    pairs and patterns that run past the field."""
    rng = np.random.default_rng(9)
    n = 30000
    words = [b"special", b"requests", b"spec", b"ial", b"re", b"quests", b"pending", b"s", b"sp"]
    com = np.zeros((n, 80), np.uint8)
    for i in range(n):
        s = b" ".join(words[j] for j in rng.integers(0, len(words), rng.integers(0, 14)))[:80]
        com[i, :len(s)] = np.frombuffer(s, np.uint8)
        if i % 3 == 0 and len(s) < 79:  # bytes after the NUL that would match
            tail = b"specialrequests"[:79 - len(s)]
            com[i, len(s) + 1:len(s) + 1 + len(tail)] = np.frombuffer(tail, np.uint8)
    cust = dict(c_custkey=np.arange(1, 501, dtype=np.int32))
    od = dict(o_custkey=rng.integers(1, 520, n).astype(np.int32), o_comment=com)
    c, o = engine.table(cust), engine.table(od)
    for w1, w2 in (("special", "requests"), ("spec", "ial"), ("sp", "sp"), ("s", "s"),
                   ("special requests", "s"), ("requests special", "pending"),
                   ("ial re", "ests spe"),
                   ("special requests special reques", "s"), ("s", "special requests special reques"),
                   ("quests", "re"), ("pending", "special"), ("spec", "ests"), ("reque", "special"),
                   ("special", "special"), ("sts s", "requests"), ("requests", "spec")):
        assert engine.q13(c, o, w1, w2) == O.q13(cust, od, w1, w2), (w1, w2)


@pytest.mark.parametrize("layout", ["dense", "dense_from_0", "permuted", "sparse", "gaps"])
def test_q13_customer_layouts(engine, layout):
    """Customer keys dense and ascending (counts turned into the histogram per
    bucket), starting at 0, in random row order, sparse, or with gaps (counts
    per custkey, then a histogram over customer rows); several custkey buckets,
    orders whose custkey has no customer, and customers with more than 1024
    counted orders (the overflow list)."""
    rng = np.random.default_rng(17)
    nc = 60_000
    keys = {"dense": np.arange(1, nc + 1), "dense_from_0": np.arange(nc),
            "permuted": rng.permutation(np.arange(1, nc + 1)),
            "sparse": np.arange(1, nc + 1) * 3 + 2,
            "gaps": np.sort(rng.choice(np.arange(1, 2 * nc), nc, replace=False))}[layout]
    keys = keys.astype(np.int32)
    n = 300_000
    ock = rng.choice(keys, n)
    ock[:5000] = keys[7]          # a customer with 5000 orders
    ock[5000:6500] = keys[-1]     # and one with 1500
    ock[6500:7000] = keys.max() + 11  # orders of no customer
    com = np.zeros((n, 80), np.uint8)
    pick = rng.random(n)
    for i, t in enumerate((b"very special requests indeed", b"special pending", b"requests special")):
        m = (pick >= i / 3) & (pick < (i + 1) / 3)
        com[m, :len(t)] = np.frombuffer(t, np.uint8)
    cust = dict(c_custkey=keys)
    od = dict(o_custkey=ock.astype(np.int32), o_comment=com)
    c, o = engine.table(cust), engine.table(od)
    got = engine.q13(c, o)
    assert got == O.q13(cust, od)
    assert sum(r["custdist"] for r in got) == nc
    assert max(r["c_count"] for r in got) > 1024


# -------------------------------------------------------------- join primitive
def test_join_count_golden(engine, golden, small):
    for c in golden["primitives"]["join"]:
        assert engine.join_count(small["orders"], c["build"], small["lineitem"],
                                 c["probe"]) == c["matches"]


def test_join_count_duplicates(engine):
    rng = np.random.default_rng(9)
    b = rng.integers(-5000, 5000, 30001).astype(np.int32)
    p = rng.integers(-6000, 6000, 1_000_003).astype(np.int32)
    tb, tp = engine.table({"k": b}), engine.table({"k": p})
    assert engine.join_count(tb, "k", tp, "k") == O.join_count(b, p)


# radix bits: -1 one global table, 0 derived (LDS broadcast for small builds),
# 1 and 4 single-pass partitions far larger than an LDS table (built and probed
# in pieces), 9 and 14 two partition passes
STRATEGIES = (-2, -1, 0, 1, 4, 9, 14)


@pytest.fixture()
def join_strategy(engine):
    saved = engine.get_tuning("join")
    yield lambda bits: engine.set_tuning("join", 256, 0, 1, bits)
    engine.set_tuning("join", saved["block"], saved["blocks_per_cu"], saved["unroll"],
                      saved["radix_bits"])


@pytest.mark.parametrize("bits", STRATEGIES)
def test_join_strategies_against_oracle(engine, join_strategy, bits):
    join_strategy(bits)
    rng = np.random.default_rng(100 + bits)
    cases = [
        (rng.integers(-5000, 5000, 30001), rng.integers(-6000, 6000, 1_000_003)),  # duplicates
        (np.arange(1, 2_000_002, 4), rng.integers(0, 8_000_000, 3_000_001)),  # sparse unique keys
        (np.full(20_000, 7), np.concatenate([np.full(50_000, 7), np.arange(100_000)])),  # skew
        (np.array([3]), np.array([3, 3, 4])),  # tiny
        (np.array([], np.int64), np.arange(10)),  # empty build
        (np.arange(10), np.array([], np.int64)),  # empty probe
        (rng.integers(-2**31 + 1, 2**31 - 1, 300_007), rng.integers(-2**31 + 1, 2**31 - 1, 5)),
    ]
    for b, p in cases:
        b, p = b.astype(np.int32), p.astype(np.int32)
        tb, tp = engine.table({"k": b}), engine.table({"k": p})
        assert engine.join_count(tb, "k", tp, "k") == O.join_count(b, p), (bits, len(b), len(p))


def test_join_full_config_all_strategies(engine, join_strategy, full):
    js, t = full
    exp = O.join_count(js["orders"]["o_orderkey"], js["lineitem"]["l_orderkey"])
    for bits in (-2, -1, 0, 1, 9, 12, 14):  # 1 and 9 are raised to the fewest bits allowed
        join_strategy(bits)
        assert engine.join_count(t["orders"], "o_orderkey", t["lineitem"], "l_orderkey") == exp, bits
    kb = engine.table({"k": js["orders"]["o_orderkey"]})
    kp = engine.table({"k": js["lineitem"]["l_orderkey"]})
    got = engine.autotune("join", [kb, kp], reps=1)
    assert -2 <= got["radix_bits"] <= 14
    assert engine.get_tuning("join")["radix_bits"] == got["radix_bits"]
    assert engine.join_count(kb, "k", kp, "k") == exp


def test_q9_parts_with_more_than_four_suppliers(engine):
    """Parts with 1..7 suppliers: the direct partsupp table holds four per part,
    so a fifth sends the query through the hash table; both must equal the
    oracle (as must parts with fewer than four)."""
    rng = np.random.default_rng(3)
    np_, ns = 400, 12
    name = np.zeros((np_, 55), np.uint8)
    for i in range(np_):
        t = (b"forest green" if i % 3 else b"lime")
        name[i, :len(t)] = np.frombuffer(t, np.uint8)
    pa = dict(p_partkey=np.arange(1, np_ + 1, dtype=np.int32), p_name=name)
    su = dict(s_suppkey=np.arange(1, ns + 1, dtype=np.int32),
              s_nationkey=rng.integers(0, 25, ns).astype(np.int32))
    for most in (4, 7):
        pk, sk = [], []
        for p in range(1, np_ + 1):
            k = 1 + (p * 5) % most
            for s_ in rng.choice(np.arange(1, ns + 1), k, replace=False):
                pk.append(p)
                sk.append(int(s_))
        ps = dict(ps_partkey=np.array(pk, np.int32), ps_suppkey=np.array(sk, np.int32),
                  ps_supplycost=rng.integers(100, 100_000, len(pk)).astype(np.int64))
        no = 3000
        od = dict(o_orderkey=(np.arange(no) * 4 + 1).astype(np.int32),
                  o_orderdate=rng.integers(0, 2406, no).astype(np.int32))
        row = rng.integers(0, len(pk), 20000)
        li = dict(l_partkey=ps["ps_partkey"][row], l_suppkey=ps["ps_suppkey"][row],
                  l_orderkey=od["o_orderkey"][rng.integers(0, no, 20000)],
                  l_quantity=rng.integers(1, 51, 20000).astype(np.int32),
                  l_discount=rng.integers(0, 11, 20000).astype(np.int32),
                  l_extendedprice=rng.integers(90_000, 10_500_000, 20000).astype(np.int64))
        tabs = [engine.table(t) for t in (pa, su, ps, od, li)]
        try:
            for pat in ("green", "lime", "e"):
                assert engine.q9(*tabs, pat) == O.q9(pa, su, ps, od, li, pat), (most, pat)
        finally:
            for t in tabs:
                t.free()


def test_q9_order_year_spans(engine):
    """Q9's year map takes 4-bit codes for orders spanning at most 15 years with
    strictly ascending orderkeys, else a byte per key value: 15 years (codes),
    16 years and 1970-2033 (bytes; suppliers in 6 nations, so the 384 groups
    fit the 400-row result), 7 years with permuted keys (bytes) and 7 years
    with dense keys 1..n (codes; every 4096-order slice of the map builder
    then shares its first word with the slice before), each with both end
    years present, must equal the oracle."""
    import datetime as dt

    def day(y):
        return (dt.date(y, 1, 1) - dt.date(1992, 1, 1)).days

    rng = np.random.default_rng(33)
    np_, ns, no = 2000, 30, 40_000
    part = generate.part(7, np_)
    ps = generate.partsupp(7, np_, ns)
    su = generate.supplier(7, ns)
    su6 = dict(su, s_nationkey=(np.asarray(su["s_nationkey"]) % 6).astype(np.int32))
    sorted_keys = (np.arange(no, dtype=np.int64) // 8 * 32 + np.arange(no) % 8 + 1).astype(np.int32)
    cases = {"15 years": (1992, 2007, sorted_keys), "16 years": (1992, 2008, sorted_keys),
             "1970-2033": (1970, 2034, sorted_keys),
             "7 years, permuted keys": (1992, 1999, rng.permutation(sorted_keys)),
             "7 years, dense keys": (1992, 1999, np.arange(1, no + 1, dtype=np.int32))}
    pt, st_, st6, pst = (engine.table(t) for t in (part, su, su6, ps))
    try:
        for name, (y0, y1, ok) in cases.items():
            sup, stab = (su6, st6) if y0 == 1970 else (su, st_)
            odate = rng.integers(day(y0), day(y1), no).astype(np.int32)
            odate[0], odate[-1] = day(y0), day(y1) - 1
            od = dict(o_orderkey=ok, o_orderdate=odate)
            owner = rng.integers(0, no, 5 * no)
            m = len(owner)
            row = rng.integers(0, len(ps["ps_partkey"]), m)
            li = dict(l_partkey=np.asarray(ps["ps_partkey"])[row].astype(np.int32),
                      l_suppkey=np.asarray(ps["ps_suppkey"])[row].astype(np.int32),
                      l_orderkey=ok[owner].astype(np.int32),
                      l_quantity=rng.integers(1, 51, m).astype(np.int32),
                      l_discount=rng.integers(0, 11, m).astype(np.int32),
                      l_extendedprice=rng.integers(90_000, 10_500_000, m).astype(np.int64))
            ot, lt = engine.table(od), engine.table(li)
            try:
                for pat in ("green", "e"):
                    want = O.q9(part, sup, ps, od, li, pat)
                    assert {r["year"] for r in want} >= ({y0, y1 - 1} if pat == "e" else set())
                    assert engine.q9(pt, stab, pst, ot, lt, pat) == want, (name, pat)
            finally:
                lt.free()
                ot.free()
    finally:
        for t in (pt, st_, st6, pst):
            t.free()


def test_q9_lineitem_layouts(engine):
    """Q9 over clustered and shuffled lineitem, and with orphan orderkeys
    (negative, past the largest orderkey, in the gaps between sparse orders)
    placed at the start of 256-row groups and at random rows: each must equal
    the oracle (an orphan joins no order)."""
    rng = np.random.default_rng(21)
    np_, ns, no = 3000, 40, 60_000
    part = generate.part(5, np_)
    ps = generate.partsupp(5, np_, ns)
    su = generate.supplier(5, ns)
    ok = np.sort(rng.choice(np.arange(1, 8 * no), no, replace=False)).astype(np.int32)
    od = dict(o_orderkey=ok, o_orderdate=rng.integers(0, 2406, no).astype(np.int32))
    lines = rng.integers(1, 8, no)
    owner = np.repeat(np.arange(no), lines)
    m = len(owner)
    row = rng.integers(0, len(ps["ps_partkey"]), m)
    base = dict(l_partkey=np.asarray(ps["ps_partkey"])[row].astype(np.int32),
                l_suppkey=np.asarray(ps["ps_suppkey"])[row].astype(np.int32),
                l_orderkey=ok[owner].astype(np.int32),
                l_quantity=rng.integers(1, 51, m).astype(np.int32),
                l_discount=rng.integers(0, 11, m).astype(np.int32),
                l_extendedprice=rng.integers(90_000, 10_500_000, m).astype(np.int64))
    orphan = base["l_orderkey"].copy()
    starts = np.arange(0, m, 256)
    orphan[starts[::3]] = np.int32(-2**31)          # negative key first in a group
    orphan[starts[1::3]] = np.int32(int(ok[-1]) + 3)  # past the last order
    gaps = np.setdiff1d(np.arange(1, int(ok[-1])), ok)[:2000]
    orphan[rng.integers(0, m, len(gaps))] = gaps    # between orders
    perm = rng.permutation(m)
    layouts = {"clustered": base,
               "shuffled": {k: v[perm] for k, v in base.items()},
               "orphans": dict(base, l_orderkey=orphan)}
    pt, st_, pst, ot = (engine.table(t) for t in (part, su, ps, od))
    try:
        for name, li in layouts.items():
            lt = engine.table(li)
            try:
                for pat in ("green", "e"):
                    got = engine.q9(pt, st_, pst, ot, lt, pat)
                    assert got == O.q9(part, su, ps, od, li, pat), (name, pat)
            finally:
                lt.free()
    finally:
        for t in (pt, st_, pst, ot):
            t.free()
