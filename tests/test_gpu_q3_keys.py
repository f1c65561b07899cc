"""GPU: Q3 over key distributions the benchmark tables do not have.

The order side of Q3 is a structure over the orderkey span (two bits per key
value), where the reference switches from its direct map to a compact hash
table once the map would exceed 30 % of the working set (src/q3.cpp:243-246).
These tests check every key layout that path must cover: sparse keys (64x the
order count), keys up to 2^31-1, negative keys, sparse customer keys, rows in
random order, and more tied groups than the candidate list holds. Results are
checked against the oracle (whose key lookups take any int32 layout) and, in
addition, through an order-preserving key remapping: the top-k of the
remapped tables must be the remapped top-k of the original ones.
"""
import numpy as np
import pytest

import pyoracle as O
from gpudbmetalbenchmark_amd import generate

pytestmark = pytest.mark.gpu

CASES = ((1, 1169, 10), (4, 2000, 100), (0, 300, 5))


@pytest.fixture(scope="module")
def mid():
    js = generate.join_set(23, 150_000, 1_500_000, 200_000, 10_000, comments=False)
    return {k: js[k] for k in ("customer", "orders", "lineitem")}


def run(engine, js, seg, date, k):
    t = {n: engine.table(v) for n, v in js.items()}
    try:
        return engine.q3(t["customer"], t["orders"], t["lineitem"], seg, date, k)
    finally:
        for x in t.values():
            x.free()


def remap(js, okey=None, ckey=None):
    out = {n: dict(v) for n, v in js.items()}
    if okey:
        out["orders"]["o_orderkey"] = okey(js["orders"]["o_orderkey"].astype(np.int64)).astype(np.int32)
        out["lineitem"]["l_orderkey"] = okey(js["lineitem"]["l_orderkey"].astype(np.int64)).astype(np.int32)
    if ckey:
        out["customer"]["c_custkey"] = ckey(js["customer"]["c_custkey"].astype(np.int64)).astype(np.int32)
        out["orders"]["o_custkey"] = ckey(js["orders"]["o_custkey"].astype(np.int64)).astype(np.int32)
    return out


def test_base_equals_oracle(engine, mid):
    for seg, date, k in CASES:
        got = run(engine, mid, seg, date, k)
        assert got == O.q3(mid["customer"], mid["orders"], mid["lineitem"], seg, date, k,
                           nthreads=8), (seg, date, k)


def test_sparse_orderkeys_64x_with_noise(engine, mid):
    noise = lambda k: (k * 2654435761) % 64  # noqa: E731  fixed per key
    f = lambda k: k * 64 + noise(k)  # noqa: E731  order-preserving, unique
    js = remap(mid, okey=f)
    assert js["orders"]["o_orderkey"].max() > 64 * len(js["orders"]["o_orderkey"])
    for seg, date, k in CASES:
        base = run(engine, mid, seg, date, k)
        got = run(engine, js, seg, date, k)
        assert got["groups"] == base["groups"]
        assert [r["orderkey"] for r in got["rows"]] == [int(f(np.int64(r["orderkey"])))
                                                       for r in base["rows"]]
        assert [r["revenue"] for r in got["rows"]] == [r["revenue"] for r in base["rows"]]
    # and directly against the oracle
    seg, date, k = CASES[0]
    assert run(engine, js, seg, date, k) == O.q3(js["customer"], js["orders"], js["lineitem"],
                                                 seg, date, k, nthreads=8)


def test_orderkeys_share_slice_words(engine, mid):
    """Orderkeys remapped by k -> k + R[k // 32], R a random non-decreasing
    sequence (strictly increasing on every key): the order-key words are built
    per 4096-order slice, and with the TPC-H layout (8 of every 32 key
    values) every slice boundary falls on a word boundary; shifted this way
    many slices share their first 32-key word with the slice before, whose
    last orders must then land in the later slice's word. Checked against the
    oracle and through the remapping."""
    top = int(max(mid["orders"]["o_orderkey"].max(), mid["lineitem"]["l_orderkey"].max())) // 32 + 2
    shift = np.cumsum(np.random.default_rng(7).integers(0, 32, top)).astype(np.int64)
    f = lambda k: k + shift[k // 32]  # noqa: E731  order-preserving, unique
    js = remap(mid, okey=f)
    ok = js["orders"]["o_orderkey"].astype(np.int64)
    assert np.all(np.diff(ok) > 0)
    n = len(ok)
    shared = sum(1 for b in range(4096, n, 4096) if (ok[b] - ok[0]) >> 5 == (ok[b - 1] - ok[0]) >> 5)
    assert shared >= 5, shared
    for seg, date, k in CASES:
        base = run(engine, mid, seg, date, k)
        got = run(engine, js, seg, date, k)
        assert got["groups"] == base["groups"]
        assert [r["orderkey"] for r in got["rows"]] == [int(f(np.int64(r["orderkey"])))
                                                       for r in base["rows"]]
        assert got == O.q3(js["customer"], js["orders"], js["lineitem"], seg, date, k,
                           nthreads=8), (seg, date, k)


@pytest.mark.parametrize("name,f", [
    ("near_int32_max", lambda k, m: 2147483647 - 61 * (m - k)),
    ("negative", lambda k, m: k - (1 << 30) - 7 * m),
    ("whole_int32_range", lambda k, m: -2147483648 + k * ((2 ** 32 - 2) // m)),
])
def test_extreme_orderkeys(engine, mid, name, f):
    m = int(mid["orders"]["o_orderkey"].max())
    fk = lambda k: f(k, m)  # noqa: E731
    js = remap(mid, okey=fk)
    ok = js["orders"]["o_orderkey"].astype(np.int64)
    assert len(np.unique(ok)) == len(ok)
    assert np.all(np.diff(ok) > 0)  # order-preserving on the (ascending) keys
    for seg, date, k in CASES:
        base = run(engine, mid, seg, date, k)
        got = run(engine, js, seg, date, k)
        assert got["groups"] == base["groups"], (name, seg, date)
        want = [int(fk(np.int64(r["orderkey"]))) for r in base["rows"]]
        assert [r["orderkey"] for r in got["rows"]] == want, (name, seg, date)
        assert [(r["revenue"], r["orderdate"], r["shippriority"]) for r in got["rows"]] == \
            [(r["revenue"], r["orderdate"], r["shippriority"]) for r in base["rows"]]
        assert got == O.q3(js["customer"], js["orders"], js["lineitem"], seg, date, k,
                           nthreads=8), (name, seg, date)


def test_sparse_negative_custkeys(engine, mid):
    js = remap(mid, ckey=lambda c: c * 1000 - 500_000_000)
    for seg, date, k in CASES:
        got = run(engine, js, seg, date, k)
        assert got == run(engine, mid, seg, date, k)
        assert got == O.q3(js["customer"], js["orders"], js["lineitem"], seg, date, k, nthreads=8)


def test_rows_in_random_order(engine, mid):
    rng = np.random.default_rng(5)
    js = {n: dict(v) for n, v in mid.items()}
    for n in ("orders", "lineitem", "customer"):
        perm = rng.permutation(len(next(iter(js[n].values()))))
        js[n] = {c: v[perm] for c, v in js[n].items()}
    for seg, date, k in CASES:
        assert run(engine, js, seg, date, k) == O.q3(js["customer"], js["orders"], js["lineitem"],
                                                     seg, date, k, nthreads=8), (seg, date, k)


def test_more_ties_than_the_candidate_list(engine):
    """150,000 groups with one revenue: every group lands in the threshold
    bucket, more than the 65,536-entry candidate list, so the ranking takes
    the per-block path; ties break on orderdate, then orderkey."""
    n = 150_000
    i = np.arange(n, dtype=np.int64)
    js = dict(
        customer=dict(c_custkey=np.array([1, 2], np.int32), c_mktsegment=np.array([0, 1], np.uint8)),
        orders=dict(o_orderkey=(i * 3 + 1).astype(np.int32), o_custkey=np.ones(n, np.int32),
                    o_orderdate=(i % 1000).astype(np.int32), o_shippriority=(i % 7).astype(np.int32)),
        lineitem=dict(l_orderkey=(i * 3 + 1).astype(np.int32), l_shipdate=np.full(n, 2000, np.int32),
                      l_discount=np.full(n, 5, np.int32),
                      l_extendedprice=np.full(n, 123456, np.int64)))
    for k in (10, 100):
        got = run(engine, js, 0, 1500, k)
        exp = O.q3(js["customer"], js["orders"], js["lineitem"], 0, 1500, k)
        assert got == exp and got["groups"] == n and len(got["rows"]) == k
