"""Writes tests/golden/golden.json: seeded tables, their column digests, and the
results sqlite3 computes for the reference's SQL (sql/q1.sql, q3.sql, q6.sql,
q9.sql, q13.sql) on them, plus hand-built edge-case tables.

Neither the engine nor the oracle is used here: the tables come from
npgen.py (an independent numpy restatement of the generator) and the answers
from stdlib sqlite3 in exact integer arithmetic (money in cents, discount and
tax in hundredths). SQLite's LIKE is ASCII case-insensitive, so the substring
filters use instr() / GLOB, which are case-sensitive like the reference's.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sqlite3
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import npgen  # noqa: E402

NATIONS = ["ALGERIA", "ARGENTINA", "BRAZIL", "CANADA", "EGYPT", "ETHIOPIA", "FRANCE",
           "GERMANY", "INDIA", "INDONESIA", "IRAN", "IRAQ", "JAPAN", "JORDAN", "KENYA",
           "MOROCCO", "MOZAMBIQUE", "PERU", "CHINA", "ROMANIA", "SAUDI ARABIA", "VIETNAM",
           "RUSSIA", "UNITED KINGDOM", "UNITED STATES"]

FLAT = dict(seed=42, n=50000)
JOIN = dict(seed=7, customers=500, orders=5000, parts=800, suppliers=40)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def text(field):
    b = bytes(field)
    return b.split(b"\0", 1)[0].decode("ascii")


def load(db, name, cols):
    keys = list(cols)
    n = len(cols[keys[0]])
    rows = []
    for i in range(n):
        row = []
        for k in keys:
            v = cols[k][i]
            if isinstance(v, np.ndarray):
                row.append(text(v))
            elif k in ("l_returnflag", "l_linestatus"):
                row.append(chr(int(v)))
            else:
                row.append(int(v))
        rows.append(row)
    db.execute(f"CREATE TABLE {name} ({', '.join(keys)})")
    db.executemany(f"INSERT INTO {name} VALUES ({', '.join('?' * len(keys))})", rows)


# ---------------------------------------------------------------- queries
def q6(db, p):
    r = db.execute(
        "SELECT COALESCE(SUM(l_extendedprice * l_discount), 0), COUNT(*) FROM lineitem "
        "WHERE l_shipdate >= ? AND l_shipdate < ? AND l_discount BETWEEN ? AND ? "
        "AND l_quantity < ?",
        (p["date_lo"], p["date_hi"], p["disc_lo"], p["disc_hi"], p["qty_lt"])).fetchone()
    return dict(params=p, revenue=r[0], rows=r[1])


def q1(db, cutoff):
    rows = db.execute(
        "SELECT l_returnflag, l_linestatus, SUM(l_quantity), SUM(l_extendedprice), "
        "SUM(l_extendedprice * (100 - l_discount)), "
        "SUM(l_extendedprice * (100 - l_discount) * (100 + l_tax)), "
        "SUM(l_discount), COUNT(*), AVG(l_quantity), AVG(l_extendedprice) / 100.0, "
        "AVG(l_discount) / 100.0 FROM lineitem WHERE l_shipdate <= ? "
        "AND l_returnflag IN ('A', 'N', 'R') AND l_linestatus IN ('F', 'O') "
        "GROUP BY l_returnflag, l_linestatus ORDER BY l_returnflag, l_linestatus",
        (cutoff,)).fetchall()
    return dict(cutoff=cutoff, rows=[dict(returnflag=r[0], linestatus=r[1], sum_qty=r[2],
                                          sum_base_price=r[3], sum_disc_price=r[4],
                                          sum_charge=str(r[5]), sum_disc=r[6], count=r[7],
                                          avg_qty=r[8], avg_price=r[9], avg_disc=r[10])
                                     for r in rows])


def q3(db, segment, date, k=10):
    base = ("FROM customer, orders, lineitem WHERE c_mktsegment = ? AND c_custkey = o_custkey "
            "AND l_orderkey = o_orderkey AND o_orderdate < ? AND l_shipdate > ? ")
    rows = db.execute(
        "SELECT l_orderkey, SUM(l_extendedprice * (100 - l_discount)) AS revenue, o_orderdate, "
        "o_shippriority " + base + "GROUP BY l_orderkey, o_orderdate, o_shippriority "
        "ORDER BY revenue DESC, o_orderdate, l_orderkey LIMIT ?",
        (segment, date, date, k)).fetchall()
    groups = db.execute("SELECT COUNT(DISTINCT l_orderkey) " + base,
                        (segment, date, date)).fetchone()[0]
    return dict(segment=segment, date=date, top_k=k, groups=groups,
                rows=[dict(orderkey=r[0], revenue=r[1], orderdate=r[2], shippriority=r[3])
                      for r in rows])


def q9(db, pattern):
    rows = db.execute(
        "SELECT n_nationkey, o_year, SUM(amount) FROM (SELECT n_name, n_nationkey, "
        "CAST(strftime('%Y', date('1992-01-01', '+' || o_orderdate || ' days')) AS INTEGER) "
        "AS o_year, l_extendedprice * (100 - l_discount) - ps_supplycost * l_quantity * 100 "
        "AS amount FROM part, supplier, lineitem, partsupp, orders, nation "
        "WHERE s_suppkey = l_suppkey AND ps_suppkey = l_suppkey AND ps_partkey = l_partkey "
        "AND p_partkey = l_partkey AND o_orderkey = l_orderkey AND s_nationkey = n_nationkey "
        "AND instr(p_name, ?) > 0) GROUP BY n_name, o_year ORDER BY n_name, o_year DESC",
        (pattern,)).fetchall()
    return dict(pattern=pattern, rows=[dict(nationkey=r[0], year=r[1], profit=r[2]) for r in rows])


def q13(db, w1, w2):
    rows = db.execute(
        "SELECT c_count, COUNT(*) AS custdist FROM (SELECT c_custkey, COUNT(o_orderkey) AS "
        "c_count FROM customer LEFT OUTER JOIN orders ON c_custkey = o_custkey AND "
        f"o_comment NOT GLOB '*{w1}*{w2}*' GROUP BY c_custkey) GROUP BY c_count "
        "ORDER BY custdist DESC, c_count DESC").fetchall()
    return dict(word1=w1, word2=w2, rows=[dict(c_count=r[0], custdist=r[1]) for r in rows])


def primitives(db):
    out = dict(select=[], sum=[], join=[])
    for v in (1000, 10000, 50000, 0, 10 ** 9):
        out["select"].append(dict(column="l_partkey", value=v, count=db.execute(
            "SELECT COUNT(*) FROM lineitem WHERE l_partkey < ?", (v,)).fetchone()[0]))
    for c in ("l_quantity", "l_extendedprice"):
        out["sum"].append(dict(column=c, sum=db.execute(f"SELECT SUM({c}) FROM lineitem").fetchone()[0]))
    out["join"].append(dict(build="o_orderkey", probe="l_orderkey", matches=db.execute(
        "SELECT COUNT(*) FROM orders, lineitem WHERE o_orderkey = l_orderkey").fetchone()[0]))
    return out


# --------------------------------------------------------- edge-case tables
def edge_q6():
    # Boundaries of every predicate around the window [731, 1096), band [5, 7], qty < 24.
    rows = [  # ship, disc, qty, price
        (731, 6, 10, 1000), (730, 6, 10, 2000), (1095, 6, 10, 4000), (1096, 6, 10, 8000),
        (800, 4, 10, 16000), (800, 5, 10, 32000), (800, 7, 10, 64000), (800, 8, 10, 128000),
        (800, 6, 23, 256000), (800, 6, 24, 512000), (800, 6, 0, 1024000),
        (800, 7, 1, 2147483647)]
    cols = dict(l_shipdate=[r[0] for r in rows], l_discount=[r[1] for r in rows],
                l_quantity=[r[2] for r in rows], l_extendedprice=[r[3] for r in rows])
    return cols


def edge_q1():
    rows = [  # ship, qty, disc, tax, price, rf, ls
        (2435, 1, 0, 0, 100, "A", "F"), (2436, 2, 0, 0, 200, "A", "F"),
        (0, 3, 10, 8, 209900 * 50, "R", "F"), (5, 50, 100, 0, 7, "N", "O"),
        (10, 4, 3, 2, 999, "X", "F"), (10, 4, 3, 2, 999, "A", "Z"),
        (-5, 7, 1, 100, 2147483647, "N", "F"), (2435, 9, 2, 3, 12345, "R", "O")]
    keys = ["l_shipdate", "l_quantity", "l_discount", "l_tax", "l_extendedprice",
            "l_returnflag", "l_linestatus"]
    return {k: [r[i] if not isinstance(r[i], str) else ord(r[i]) for r in rows]
            for i, k in enumerate(keys)}


def edge_q3():
    customer = dict(c_custkey=[1, 2, 3], c_mktsegment=[1, 1, 0])
    orders = dict(o_orderkey=[10, 11, 12, 13, 14], o_custkey=[1, 2, 3, 1, 2],
                  o_orderdate=[100, 100, 100, 99, 1169], o_shippriority=[0, 0, 0, 0, 0])
    # Orders 10, 11 and 13 tie on revenue; 11 and 10 also tie on date.
    lineitem = dict(l_orderkey=[10, 11, 13, 12, 14, 10, 11, 13, 99],
                    l_shipdate=[1170, 1170, 1170, 1170, 1170, 1169, 1169, 1169, 1170],
                    l_discount=[0, 0, 0, 0, 0, 0, 0, 0, 0],
                    l_extendedprice=[500, 500, 500, 900, 900, 7, 7, 7, 100])
    return customer, orders, lineitem


def comment_field(s):
    b = s.encode()
    assert len(b) <= 80
    return list(b) + [0] * (80 - len(b))


def edge_q13():
    comments = [
        "special only here",
        "requests before special",
        "special requests",
        "specialrequests",
        "x" * 64 + "special" + "requests"[:1] + "",  # 72 bytes, cut word
        "y" * 65 + "special" + "requests",            # exactly 80 bytes, no NUL
        "y" * 66 + "special" + "request",             # 80 bytes, word cut
        "",
        "speciaL requests",
        "special special requests requests",
    ]
    custkeys = [1, 1, 2, 2, 3, 3, 4, 4, 5, 7]  # 7 is not a customer
    customer = dict(c_custkey=[1, 2, 3, 4, 5, 6])
    orders = dict(o_orderkey=list(range(1, len(comments) + 1)), o_custkey=custkeys,
                  o_comment=[comment_field(c) for c in comments])
    return customer, orders


def to_np(cols):
    out = {}
    for k, v in cols.items():
        if k in ("o_comment",):
            out[k] = np.array(v, dtype=np.uint8)
        elif k in ("l_extendedprice", "ps_supplycost"):
            out[k] = np.array(v, dtype=np.int64)
        elif k in ("l_returnflag", "l_linestatus", "c_mktsegment"):
            out[k] = np.array(v, dtype=np.uint8)
        else:
            out[k] = np.array(v, dtype=np.int32)
    return out


def main():
    g = dict(flat=dict(FLAT), join=dict(JOIN))

    # ---- flat fact table: Q6, Q1
    li = npgen.lineitem_flat(FLAT["seed"], FLAT["n"])
    g["flat"]["digests"] = {k: digest(v) for k, v in li.items()}
    db = sqlite3.connect(":memory:")
    load(db, "lineitem", {k: v for k, v in li.items() if k != "l_receiptdate"})
    g["q6"] = [q6(db, p) for p in (
        dict(date_lo=731, date_hi=1096, disc_lo=5, disc_hi=7, qty_lt=24),
        dict(date_lo=0, date_hi=2526, disc_lo=0, disc_hi=10, qty_lt=51),
        dict(date_lo=2000, date_hi=2000, disc_lo=5, disc_hi=7, qty_lt=24),
        dict(date_lo=1500, date_hi=1865, disc_lo=1, disc_hi=3, qty_lt=25),
        dict(date_lo=100, date_hi=465, disc_lo=7, disc_hi=5, qty_lt=24))]
    g["q1"] = [q1(db, c) for c in (2435, 2525, 1262, -1)]
    db.close()

    # ---- joined tables: Q3, Q9, Q13, primitives
    s = JOIN["seed"]
    cu = npgen.customer(s, JOIN["customers"])
    od = npgen.orders(s, JOIN["orders"], JOIN["customers"], comments=True)
    lj = npgen.lineitem_joined(s, JOIN["orders"], JOIN["parts"], JOIN["suppliers"])
    pa = npgen.part(s, JOIN["parts"])
    ps = npgen.partsupp(s, JOIN["parts"], JOIN["suppliers"])
    su = npgen.supplier(s, JOIN["suppliers"])
    g["join"]["lineitem_rows"] = int(len(lj["l_orderkey"]))
    g["join"]["digests"] = {k: digest(v) for t in (cu, od, lj, pa, ps, su) for k, v in t.items()}
    db = sqlite3.connect(":memory:")
    load(db, "customer", cu)
    load(db, "orders", od)
    load(db, "lineitem", lj)
    load(db, "part", pa)
    load(db, "partsupp", ps)
    load(db, "supplier", su)
    db.execute("CREATE TABLE nation (n_nationkey, n_name)")
    db.executemany("INSERT INTO nation VALUES (?, ?)", list(enumerate(NATIONS)))
    g["q3"] = [q3(db, 1, 1169), q3(db, 3, 500), q3(db, 0, 2000, 100), q3(db, 4, 0)]
    g["q9"] = [q9(db, "green"), q9(db, "blue"), q9(db, "zzz")]
    g["q13"] = [q13(db, "special", "requests"), q13(db, "requests", "special")]
    g["primitives"] = primitives(db)
    db.close()

    # ---- edge cases
    e = {}
    t = edge_q6()
    db = sqlite3.connect(":memory:")
    load(db, "lineitem", to_np(t))
    e["q6"] = dict(table=t, results=[q6(db, dict(date_lo=731, date_hi=1096, disc_lo=5, disc_hi=7,
                                                 qty_lt=24))])
    db.close()
    t = edge_q1()
    db = sqlite3.connect(":memory:")
    load(db, "lineitem", to_np(t))
    e["q1"] = dict(table=t, results=[q1(db, 2435)])
    db.close()
    c, o, li3 = edge_q3()
    db = sqlite3.connect(":memory:")
    load(db, "customer", to_np(c))
    load(db, "orders", to_np(o))
    load(db, "lineitem", to_np(li3))
    e["q3"] = dict(customer=c, orders=o, lineitem=li3, results=[q3(db, 1, 1169, 3), q3(db, 1, 1169)])
    db.close()
    c, o = edge_q13()
    db = sqlite3.connect(":memory:")
    load(db, "customer", to_np(c))
    load(db, "orders", to_np(o))
    e["q13"] = dict(customer=c, orders=o, results=[q13(db, "special", "requests")])
    db.close()
    g["edge"] = e

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
