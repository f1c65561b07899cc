"""CPU: the .tbl ingest (gdb_tbl_count_rows / gdb_tbl_parse) and the binary
column format. Values are checked against Python's own decimal and calendar
arithmetic, and generated tables must survive a dump -> parse round trip
unchanged."""
import datetime
import os

import numpy as np
import pytest

from gpudbmetalbenchmark_amd import GdbError, generate, tbl

DAY0 = datetime.date(1992, 1, 1)


def day(s):
    return (datetime.date.fromisoformat(s) - DAY0).days


def test_tpch_spec_line(tmp_path):
    # the first lineitem row dbgen writes for the TPC-H scale-1 database
    p = tmp_path / "lineitem.tbl"
    p.write_text("1|155190|7706|1|17|21168.23|0.04|0.02|N|O|1996-03-13|1996-02-12|1996-03-22|"
                 "DELIVER IN PERSON|TRUCK|egular courts above the|\n")
    c = tbl.load(str(p), "lineitem")
    assert c["l_orderkey"].tolist() == [1] and c["l_partkey"].tolist() == [155190]
    assert c["l_suppkey"].tolist() == [7706] and c["l_quantity"].tolist() == [17]
    assert c["l_extendedprice"].tolist() == [2116823]
    assert c["l_discount"].tolist() == [4] and c["l_tax"].tolist() == [2]
    assert bytes(c["l_returnflag"]) == b"N" and bytes(c["l_linestatus"]) == b"O"
    assert c["l_shipdate"].tolist() == [day("1996-03-13")]
    assert c["l_receiptdate"].tolist() == [day("1996-03-22")]


def test_values_against_python(tmp_path):
    rng = np.random.default_rng(3)
    dates = [DAY0 + datetime.timedelta(days=int(d)) for d in rng.integers(-9000, 20000, 500)]
    dates += [datetime.date(1996, 2, 29), datetime.date(2000, 2, 29), datetime.date(1900, 3, 1),
              datetime.date(1992, 1, 1), datetime.date(1991, 12, 31)]
    cents = rng.integers(-10**12, 10**12, len(dates))
    lines = []
    for i, (d, c) in enumerate(zip(dates, cents)):
        money = f"{'-' if c < 0 else ''}{abs(int(c)) // 100}.{abs(int(c)) % 100:02d}"
        if i % 7 == 0 and c % 100 == 0:
            money = f"{'-' if c < 0 else ''}{abs(int(c)) // 100}"       # no fraction
        if i % 11 == 0 and c % 10 == 0:
            money = f"{'-' if c < 0 else ''}{abs(int(c)) // 100}.{abs(int(c)) % 100 // 10}"  # one digit
        lines.append(f"{i}|x|x|{i % 25}|x|{money}|x|\n")
    p = tmp_path / "supplier.tbl"
    p.write_text("".join(lines))
    # parse field 5 (s_acctbal) as money via the generic entry point
    from gpudbmetalbenchmark_amd import abi
    lib = abi.load()
    n = len(lines)
    money = np.empty(n, np.int64)
    keys = np.empty(n, np.int32)
    f = (abi.TblField * 2)(abi.TblField(5, abi.TBL_CENTS, 0, money.ctypes.data),
                           abi.TblField(0, abi.TBL_INT32, 0, keys.ctypes.data))
    abi.check(lib.gdb_tbl_parse(os.fsencode(str(p)), f, 2, n, 3))
    assert money.tolist() == [int(c) for c in cents]
    assert keys.tolist() == list(range(n))
    # dates, through a one-field file
    q = tmp_path / "dates.tbl"
    q.write_text("".join(f"{d.isoformat()}|\n" for d in dates))
    out = np.empty(n, np.int32)
    g = (abi.TblField * 1)(abi.TblField(0, abi.TBL_DATE, 0, out.ctypes.data))
    abi.check(lib.gdb_tbl_parse(os.fsencode(str(q)), g, 1, n, 2))
    assert out.tolist() == [(d - DAY0).days for d in dates]


def test_text_fields_and_segments(tmp_path):
    p = tmp_path / "customer.tbl"
    names = tbl.segment_names()
    rows = [f"{i}|Customer#{i}|addr|3|11-111|1.00|{names[i % 5]}|c|\n" for i in range(1, 11)]
    rows.append("11|C|a|3|1|0.00|UNKNOWN|c|\n")
    p.write_text("".join(rows))
    c = tbl.load(str(p), "customer")
    assert c["c_custkey"].tolist() == list(range(1, 12))
    assert c["c_mktsegment"].tolist() == [i % 5 for i in range(1, 11)] + [255]
    q = tmp_path / "part.tbl"
    long = "x" * 70
    q.write_text(f"1|green lime|m|b|t|1|c|1.00|c|\n2|{long}|m|b|t|1|c|1.00|c|\n3||m|b|t|1|c|1.00|c|\n")
    pt = tbl.load(str(q), "part")
    assert bytes(pt["p_name"][0]).rstrip(b"\0") == b"green lime"
    assert bytes(pt["p_name"][1]) == b"x" * 55          # truncated to the field width
    assert bytes(pt["p_name"][2]) == b"\0" * 55


@pytest.mark.parametrize("text,what", [
    ("1|2|\n3|\n", "too few fields"),
    ("1|2|4|12.345|c|\n", "row 1 field 3: cannot parse"),
    ("1|2|4|7.00|c|\n1|2|4|1e3|c|\n", "row 2 field 3: cannot parse"),
])
def test_malformed_rows_are_rejected(tmp_path, text, what):
    p = tmp_path / "partsupp.tbl"
    p.write_text(text)
    with pytest.raises(GdbError, match=what):
        tbl.load(str(p), "partsupp")


def test_bad_dates_are_rejected(tmp_path):
    p = tmp_path / "orders.tbl"
    p.write_text("1|2|O|1.00|1995-13-01|x|x|0|c|\n")
    with pytest.raises(GdbError, match="row 1 field 4"):
        tbl.load(str(p), "orders")


def test_missing_file():
    with pytest.raises(GdbError, match="cannot open"):
        tbl.load("/nonexistent/lineitem.tbl", "lineitem")


def test_round_trip_generated_tables(tmp_path):
    seed = 5
    tables = {
        "customer": generate.customer(seed, 300),
        "orders": generate.orders(seed, 1000, 300),
        "lineitem": generate.lineitem_joined(seed, 1000, 200, 50),
        "part": generate.part(seed, 200),
        "partsupp": generate.partsupp(seed, 200, 50),
        "supplier": generate.supplier(seed, 50),
    }
    for name, cols in tables.items():
        p = str(tmp_path / f"{name}.tbl")
        tbl.dump(p, name, cols)
        back = tbl.load(p, name, columns=list(cols), nthreads=4)
        assert set(back) == set(cols), name
        for k in cols:
            assert back[k].dtype == cols[k].dtype, (name, k)
            assert np.array_equal(back[k], cols[k]), (name, k)
    flat = generate.lineitem_flat(seed, 5000, receipt=True)
    p = str(tmp_path / "flat.tbl")
    tbl.dump(p, "lineitem", flat)
    back = tbl.load(p, "lineitem", columns=list(flat))
    for k in flat:
        assert np.array_equal(back[k], flat[k]), k


def test_binary_columns_round_trip(tmp_path):
    cols = generate.orders(7, 500, 100)
    d = str(tmp_path / "orders.cols")
    tbl.save_columns(d, cols)
    for mm in (True, False):
        back = tbl.load_columns(d, mmap=mm)
        assert set(back) == set(cols)
        for k in cols:
            assert back[k].dtype == cols[k].dtype and np.array_equal(back[k], cols[k]), k


def test_tpch_database_round_trip(tmp_path):
    """A whole TPC-H-shaped database written in dbgen's field layout (every
    field, dictionary text, Brand#MN, Manufacturer#M, names, phones) parses
    back into exactly the same engine columns."""
    from gpudbmetalbenchmark_amd import tpch
    db = tpch.generate(0.005, 9)
    tbl.dump_database(str(tmp_path), db)
    back = tbl.load_database(str(tmp_path))
    for t, cols in db.items():
        assert set(back[t]) == set(cols), t
        for k, v in cols.items():
            assert np.array_equal(np.asarray(v), back[t][k]), (t, k)
    line = open(tmp_path / "part.tbl").readline().split("|")
    assert line[2].startswith("Manufacturer#") and line[3].startswith("Brand#")
    assert line[4] in tpch.P_TYPE and line[6] in tpch.P_CONTAINER
