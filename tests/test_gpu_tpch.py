"""GPU: the 17 TPC-H pipelines of include/gdb_tpch.h against the CPU oracle
(oracle/tpch_oracle.py, itself pinned to sqlite3 by tests/golden), bit-exact,
on the golden scale factors and at SF1 (6M lineitem rows), with the reference
SQL's constants and with other parameters (dates, nations, regions, patterns,
thresholds) that move the selectivities, including empty results."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import tpch_oracle as T  # noqa: E402
from gpudbmetalbenchmark_amd import Engine, tpch  # noqa: E402
from gpudbmetalbenchmark_amd.tpchdb import TpchDb  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [(0.01, 7), (0.02, 3), (1.0, 11)]


@pytest.fixture(scope="module")
def engine():
    with Engine(0) as e:
        yield e


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"sf{c[0]}-seed{c[1]}")
def db(request, engine):
    sf, seed = request.param
    print(f"\ncase sf={sf} seed={seed}", flush=True)
    tabs = tpch.generate(sf, seed)
    d = TpchDb(engine, tabs)
    yield tabs, d
    d.free()


def L(rows):
    return [tuple(r) for r in rows]


BRASS = tpch.codes_like(tpch.P_TYPE, "%BRASS")
MEDPOL = tpch.codes_like(tpch.P_TYPE, "MEDIUM POLISHED%")


def test_q2(db):
    tabs, d = db
    assert L(d.q2()) == L(T.q2(tabs, 15, BRASS, 3))
    tin = tpch.codes_like(tpch.P_TYPE, "%TIN")
    assert L(d.q2(size=7, type_like="%TIN", region="ASIA", top_k=17)) == \
        L(T.q2(tabs, 7, tin, 2, 17))
    assert d.q2(size=99) == []


def test_q4(db):
    tabs, d = db
    assert d.q4() == T.q4(tabs)
    assert d.q4(0, 2500) == T.q4(tabs, 0, 2500)
    assert d.q4(5000, 6000) == []


def test_q5(db):
    tabs, d = db
    assert d.q5() == T.q5(tabs)
    assert d.q5("EUROPE", 1000, 1800) == T.q5(tabs, 3, 1000, 1800)
    assert d.q5("AFRICA", 0, 0) == []


def test_q7(db):
    tabs, d = db
    assert d.q7() == T.q7(tabs)
    assert d.q7("JAPAN", "CHINA", 0, 2921) == T.q7(tabs, 12, 18, 0, 2921)
    assert d.q7("PERU", "PERU", 500, 2000) == T.q7(tabs, 17, 17, 500, 2000)


def test_q8(db):
    tabs, d = db
    assert d.q8() == T.q8(tabs)
    code = tpch.P_TYPE.index("PROMO PLATED TIN")
    assert d.q8("JAPAN", "ASIA", "PROMO PLATED TIN", 0, 2921) == \
        T.q8(tabs, 12, 2, code, 0, 2921)


def test_q10(db):
    tabs, d = db
    rows, groups = d.q10()
    exp_rows, exp_groups = T.q10(tabs)
    assert (L(rows), groups) == (L(exp_rows), exp_groups)
    rows, groups = d.q10(0, 1500, top_k=100)
    exp_rows, exp_groups = T.q10(tabs, 0, 1500, 100)
    assert (L(rows), groups) == (L(exp_rows), exp_groups)
    assert d.q10(3000, 3100) == ([], 0)


def test_q11(db):
    tabs, d = db
    rows, total = d.q11()
    exp_rows, exp_total = T.q11(tabs)
    assert (L(rows), total) == (L(exp_rows), exp_total)
    rows, total = d.q11("PERU", 1, 1000)
    assert (L(rows), total) == tuple(map(lambda x: L(x) if isinstance(x, list) else x,
                                         T.q11(tabs, 17, 1, 1000)))
    rows, total = d.q11("CHINA", 0, 1, cap=50)
    exp_rows, exp_total = T.q11(tabs, 18, 0, 1)
    assert (L(rows), total) == (L(exp_rows[:50]), exp_total)


def test_q12(db):
    tabs, d = db
    assert d.q12() == T.q12(tabs)
    assert d.q12("AIR", "TRUCK", 0, 2600, ("3-MEDIUM",)) == T.q12(tabs, 1, 4, 0, 2600, (2,))


def test_q14(db):
    tabs, d = db
    assert d.q14() == T.q14(tabs)
    std = tpch.codes_like(tpch.P_TYPE, "STANDARD%")
    assert d.q14(0, 2600, "STANDARD%") == T.q14(tabs, 0, 2600, std)
    assert d.q14(9000, 9100) == (0, 0)


def test_q15(db):
    tabs, d = db
    assert d.q15() == T.q15(tabs)
    assert d.q15(0, 2600) == T.q15(tabs, 0, 2600)
    assert d.q15(9000, 9100) == []


def test_q16(db):
    tabs, d = db
    assert L(d.q16().tolist()) == L(T.q16(tabs, 45, MEDPOL))
    sm = tpch.codes_like(tpch.P_TYPE, "SMALL%")
    assert L(d.q16("Brand#11", "SMALL%", (1, 2, 3), ("Customer", "Recommends")).tolist()) == \
        L(T.q16(tabs, 11, sm, (1, 2, 3), ("Customer", "Recommends")))


def test_q17(db):
    tabs, d = db
    assert d.q17() == T.q17(tabs)
    cont = tpch.P_CONTAINER.index("SM PACK")
    assert d.q17("Brand#11", "SM PACK", 1, 1) == T.q17(tabs, 11, cont, 1, 1)


def test_q18(db):
    tabs, d = db
    assert L(d.q18()) == L(T.q18(tabs))
    assert L(d.q18(250)) == L(T.q18(tabs, 250))
    assert L(d.q18(0, 37)) == L(T.q18(tabs, 0, 37))


def test_q19(db):
    tabs, d = db
    assert d.q19() == T.q19(tabs)
    blocks = (("Brand#11", ("SM CASE", "LG BOX", "JUMBO DRUM"), 1, 50, 1, 50),)
    assert d.q19(blocks, ("MAIL",), "NONE") == T.q19(tabs, blocks, ("MAIL",), "NONE")


def test_q20(db):
    tabs, d = db
    assert d.q20() == T.q20(tabs)
    assert d.q20("green", 0, 2600, "JAPAN", 1, 10) == T.q20(tabs, b"green", 0, 2600, 12, 1, 10)


def test_q21(db):
    tabs, d = db
    assert L(d.q21()) == L(T.q21(tabs))
    assert L(d.q21("KENYA", "O", 7)) == L(T.q21(tabs, 14, ord("O"), 7))


def test_q22(db):
    tabs, d = db
    assert d.q22() == T.q22(tabs)
    assert d.q22((10, 11, 34)) == T.q22(tabs, (10, 11, 34))


def test_unclustered_lineitem_takes_the_general_paths(engine):
    """Q4, Q18 and Q21 with lineitem rows shuffled (l_orderkey no longer
    non-decreasing, so the per-order owner kernels are not used): same answers."""
    import numpy as np
    tabs = tpch.generate(0.05, 5)
    perm = np.random.default_rng(1).permutation(len(tabs["lineitem"]["l_orderkey"]))
    shuffled = dict(tabs, lineitem={k: v[perm] for k, v in tabs["lineitem"].items()})
    d = TpchDb(engine, shuffled)
    try:
        assert L(d.q18(250)) == L(T.q18(tabs, 250))
        assert L(d.q21()) == L(T.q21(tabs))
        assert L(d.q21("KENYA", "O", 7)) == L(T.q21(tabs, 14, ord("O"), 7))
        assert d.q4(0, 2500) == T.q4(tabs, 0, 2500)
    finally:
        d.free()


def test_unsorted_orders_table(engine):
    """Orders rows shuffled (o_orderkey no longer ascending, so Q18 cannot look
    its candidates up by binary search and sums per orderkey instead), lineitem
    still clustered: same answers for every query that reads orders."""
    import numpy as np
    tabs = tpch.generate(0.02, 8)
    perm = np.random.default_rng(3).permutation(len(tabs["orders"]["o_orderkey"]))
    shuffled = dict(tabs, orders={k: v[perm] for k, v in tabs["orders"].items()})
    d = TpchDb(engine, shuffled)
    try:
        assert L(d.q18()) == L(T.q18(tabs))
        assert L(d.q18(250)) == L(T.q18(tabs, 250))
        assert L(d.q18(0, 37)) == L(T.q18(tabs, 0, 37))
        assert L(d.q21()) == L(T.q21(tabs))
        assert d.q4() == T.q4(tabs)
        assert d.q5() == T.q5(tabs)
        rows, groups = d.q10()
        assert (L(rows), groups) == (L(T.q10(tabs)[0]), T.q10(tabs)[1])
    finally:
        d.free()


def test_long_orders_run_past_the_owner_window(engine):
    """Q18 and Q21 over a lineitem whose l_orderkey stays non-decreasing but
    has runs of 13 to 60 rows (longer than the 12-row window an owner lane holds
    in registers), placed at every offset within a wave's 256 rows."""
    import numpy as np
    tabs = tpch.generate(0.02, 4)
    ok = tabs["lineitem"]["l_orderkey"].copy()
    rng = np.random.default_rng(2)
    b = 0
    while b < len(ok) - 64:
        run = int(rng.integers(13, 61))
        ok[b:b + run] = ok[b]
        b += run + int(rng.integers(1, 300))
    assert np.all(np.diff(ok) >= 0)
    tabs = dict(tabs, lineitem=dict(tabs["lineitem"], l_orderkey=ok))
    d = TpchDb(engine, tabs)
    try:
        assert L(d.q18()) == L(T.q18(tabs))
        assert L(d.q18(250)) == L(T.q18(tabs, 250))
        assert L(d.q21()) == L(T.q21(tabs))
        assert L(d.q21("KENYA", "O", 7)) == L(T.q21(tabs, 14, ord("O"), 7))
    finally:
        d.free()


def test_top_k_ties_past_the_candidate_list(engine):
    """Q18 with every o_totalprice equal and quantity 0: all 75K orders tie in
    one value bucket, more than the 8192 candidates the top-k collects, so its
    ranking workgroup scans every candidate; the ties resolve by orderdate, then
    orderkey. Both the clustered (candidate list) and shuffled-lineitem paths."""
    import numpy as np
    tabs = tpch.generate(0.05, 6)
    tp = np.full_like(tabs["orders"]["o_totalprice"], 12345678)
    tabs = dict(tabs, orders=dict(tabs["orders"], o_totalprice=tp))
    perm = np.random.default_rng(4).permutation(len(tabs["lineitem"]["l_orderkey"]))
    shuffled = dict(tabs, lineitem={k: v[perm] for k, v in tabs["lineitem"].items()})
    for t in (tabs, shuffled):
        d = TpchDb(engine, t)
        try:
            assert L(d.q18(0, 100)) == L(T.q18(tabs, 0, 100))
            assert L(d.q18(0, 1)) == L(T.q18(tabs, 0, 1))
        finally:
            d.free()


def test_streamed_lineitem_gives_the_same_results(engine):
    """Every pipeline with lineitem in host memory, streamed through the GPU in
    64K-row chunks (two passes for Q17 and Q21, the general paths for Q18 and
    Q21): the same answers as the oracle."""
    tabs = tpch.generate(0.05, 12)
    engine.set_chunk_rows(65536)
    d = TpchDb(engine, tabs, stream_lineitem=True)
    try:
        from gpudbmetalbenchmark_amd import Timing
        tm = Timing()
        assert d.q14(timing=tm) == T.q14(tabs) and tm.chunks > 1
        assert L(d.q2()) == L(T.q2(tabs, 15, BRASS, 3))
        assert d.q4() == T.q4(tabs)
        assert d.q5() == T.q5(tabs)
        assert d.q7() == T.q7(tabs)
        assert d.q8() == T.q8(tabs)
        rows, groups = d.q10()
        assert (L(rows), groups) == (L(T.q10(tabs)[0]), T.q10(tabs)[1])
        assert d.q12() == T.q12(tabs)
        assert d.q15() == T.q15(tabs)
        assert d.q17() == T.q17(tabs)
        assert d.q17("Brand#11", "SM PACK", 1, 1) == T.q17(tabs, 11, tpch.P_CONTAINER.index("SM PACK"), 1, 1)
        assert L(d.q18(250)) == L(T.q18(tabs, 250))
        assert d.q19() == T.q19(tabs)
        assert d.q20("green", 0, 2600, "JAPAN", 1, 10) == T.q20(tabs, b"green", 0, 2600, 12, 1, 10)
        assert L(d.q21()) == L(T.q21(tabs))
        assert L(d.q21("KENYA", "O", 7)) == L(T.q21(tabs, 14, ord("O"), 7))
        assert d.q22() == T.q22(tabs)
    finally:
        d.free()
        engine.set_chunk_rows(0)


def test_q20_pair_table_grows_past_its_estimate():
    """Q20's pair table is sized from the last run's pair count (1/32 of
    partsupp on a fresh context). Prefixes that select more parts than that
    make the device flag the overflow and the query rerun with a table sized
    from the count: the answers stay exact, in either order of prefixes."""
    tabs = tpch.generate(0.02, 9)
    with Engine(0) as e:
        d = TpchDb(e, tabs)
        try:
            for pre in ("a", "forest", "b", "a", "lemon"):
                assert d.q20(pre, 0, 2600, "JAPAN", 1, 10) == \
                    T.q20(tabs, pre.encode(), 0, 2600, 12, 1, 10), pre
        finally:
            d.free()
