"""Writes tests/golden/tpch_golden.json: expected outputs of the TPC-H queries
Q2, Q4, Q5, Q7, Q8, Q10, Q11, Q12, Q14-Q22 over the seeded TPC-H-shaped tables
(gpudbmetalbenchmark_amd.tpch.generate), computed by stdlib sqlite3 running
each query's SQL (reference sql/qN.sql) over the same rows with every
dictionary code decoded back to its TPC-H text.

The SQL keeps the reference's correlated subqueries (EXISTS, NOT EXISTS,
scalar MIN / AVG / SUM subqueries), so sqlite's evaluation is independent of
the oracle's decorrelated numpy. Dialect changes only: money is integer cents
and discount integer hundredths, so l_extendedprice * (1 - l_discount)
becomes l_extendedprice * (100 - l_discount) (1e-4 dollars); dates are ISO
text with interval arithmetic resolved to literals; EXTRACT(YEAR ...) is
strftime('%Y'); LIKE is replaced by GLOB or instr where case matters
(sqlite's LIKE ignores ASCII case); decimal factors (0.2*AVG, 0.5*SUM,
0.0001*SUM) become integer cross-multiplications; ties the SQL leaves open are
broken by key columns, as the oracle does.

    python tests/golden/make_tpch_golden.py
"""
import datetime
import json
import os
import sqlite3
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpudbmetalbenchmark_amd import tpch  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tpch_golden.json")
DAY0 = datetime.date(1992, 1, 1)
CASES = [dict(sf=0.01, seed=7), dict(sf=0.02, seed=3)]


def iso(day):
    return (DAY0 + datetime.timedelta(days=int(day))).isoformat()


def load(db):
    con = sqlite3.connect(":memory:")
    cur = con.cursor()
    cur.execute("CREATE TABLE nation(n_nationkey INT, n_name TEXT, n_regionkey INT)")
    cur.executemany("INSERT INTO nation VALUES (?,?,?)",
                    [(i, tpch.N_NAME[i], tpch.N_REGION[i]) for i in range(25)])
    cur.execute("CREATE TABLE region(r_regionkey INT, r_name TEXT)")
    cur.executemany("INSERT INTO region VALUES (?,?)", list(enumerate(tpch.R_NAME)))
    p = db["part"]
    cur.execute("CREATE TABLE part(p_partkey INT PRIMARY KEY, p_name TEXT, p_mfgr TEXT, "
                "p_brand TEXT, p_type TEXT, p_size INT, p_container TEXT, p_retailprice INT)")
    cur.executemany("INSERT INTO part VALUES (?,?,?,?,?,?,?,?)", [
        (int(p["p_partkey"][i]), tpch.text(p["p_name"][i]), f"Manufacturer#{int(p['p_mfgr'][i])}",
         tpch.brand_name(p["p_brand"][i]), tpch.P_TYPE[p["p_type"][i]], int(p["p_size"][i]),
         tpch.P_CONTAINER[p["p_container"][i]], int(p["p_retailprice"][i]))
        for i in range(len(p["p_partkey"]))])
    s = db["supplier"]
    cur.execute("CREATE TABLE supplier(s_suppkey INT PRIMARY KEY, s_name TEXT, s_nationkey INT, "
                "s_acctbal INT, s_comment TEXT)")
    cur.executemany("INSERT INTO supplier VALUES (?,?,?,?,?)", [
        (int(s["s_suppkey"][i]), f"Supplier#{int(s['s_suppkey'][i]):09d}",
         int(s["s_nationkey"][i]), int(s["s_acctbal"][i]), tpch.text(s["s_comment"][i]))
        for i in range(len(s["s_suppkey"]))])
    ps = db["partsupp"]
    cur.execute("CREATE TABLE partsupp(ps_partkey INT, ps_suppkey INT, ps_availqty INT, "
                "ps_supplycost INT)")
    cur.executemany("INSERT INTO partsupp VALUES (?,?,?,?)", zip(
        *(ps[k].tolist() for k in ("ps_partkey", "ps_suppkey", "ps_availqty", "ps_supplycost"))))
    c = db["customer"]
    cur.execute("CREATE TABLE customer(c_custkey INT PRIMARY KEY, c_nationkey INT, "
                "c_phone TEXT, c_acctbal INT, c_mktsegment TEXT)")
    cur.executemany("INSERT INTO customer VALUES (?,?,?,?,?)", [
        (int(c["c_custkey"][i]), int(c["c_nationkey"][i]),
         f"{int(c['c_nationkey'][i]) + 10}-{100 + i % 900}-{200 + i % 700}-{1000 + i % 9000}",
         int(c["c_acctbal"][i]), tpch.C_MKTSEGMENT[c["c_mktsegment"][i]])
        for i in range(len(c["c_custkey"]))])
    o = db["orders"]
    cur.execute("CREATE TABLE orders(o_orderkey INT PRIMARY KEY, o_custkey INT, "
                "o_orderstatus TEXT, o_totalprice INT, o_orderdate TEXT, o_orderpriority TEXT)")
    cur.executemany("INSERT INTO orders VALUES (?,?,?,?,?,?)", [
        (int(o["o_orderkey"][i]), int(o["o_custkey"][i]), chr(o["o_orderstatus"][i]),
         int(o["o_totalprice"][i]), iso(o["o_orderdate"][i]),
         tpch.O_ORDERPRIORITY[o["o_orderpriority"][i]]) for i in range(len(o["o_orderkey"]))])
    li = db["lineitem"]
    cur.execute("CREATE TABLE lineitem(l_orderkey INT, l_partkey INT, l_suppkey INT, "
                "l_quantity INT, l_extendedprice INT, l_discount INT, l_returnflag TEXT, "
                "l_shipdate TEXT, l_commitdate TEXT, l_receiptdate TEXT, l_shipinstruct TEXT, "
                "l_shipmode TEXT)")
    cols = [li[k].tolist() for k in ("l_orderkey", "l_partkey", "l_suppkey", "l_quantity",
                                     "l_extendedprice", "l_discount")]
    rf = [chr(x) for x in li["l_returnflag"]]
    sd = [iso(x) for x in li["l_shipdate"]]
    cd = [iso(x) for x in li["l_commitdate"]]
    rd = [iso(x) for x in li["l_receiptdate"]]
    si = [tpch.L_SHIPINSTRUCT[x] for x in li["l_shipinstruct"]]
    sm = [tpch.L_SHIPMODE[x] for x in li["l_shipmode"]]
    cur.executemany("INSERT INTO lineitem VALUES (?,?,?,?,?,?,?,?,?,?,?,?)",
                    zip(*cols, rf, sd, cd, rd, si, sm))
    for ddl in ("CREATE INDEX li_ok ON lineitem(l_orderkey)",
                "CREATE INDEX li_pk ON lineitem(l_partkey, l_suppkey)",
                "CREATE INDEX ps_pk ON partsupp(ps_partkey)",
                "CREATE INDEX o_ck ON orders(o_custkey)"):
        cur.execute(ddl)
    con.commit()
    return con


SQL = {
    "q2": """
SELECT s_acctbal, s_suppkey, n_nationkey, p_partkey, CAST(substr(p_mfgr, 14) AS INT)
FROM part, supplier, partsupp, nation, region
WHERE p_partkey = ps_partkey AND s_suppkey = ps_suppkey AND p_size = 15
  AND p_type GLOB '*BRASS' AND s_nationkey = n_nationkey AND n_regionkey = r_regionkey
  AND r_name = 'EUROPE'
  AND ps_supplycost = (SELECT MIN(ps_supplycost) FROM partsupp, supplier, nation, region
                       WHERE p_partkey = ps_partkey AND s_suppkey = ps_suppkey
                         AND s_nationkey = n_nationkey AND n_regionkey = r_regionkey
                         AND r_name = 'EUROPE')
ORDER BY s_acctbal DESC, n_name, s_name, p_partkey LIMIT 100""",
    "q4": """
SELECT o_orderpriority, COUNT(*) FROM orders
WHERE o_orderdate >= '1993-07-01' AND o_orderdate < '1993-10-01'
  AND EXISTS (SELECT * FROM lineitem WHERE l_orderkey = o_orderkey
              AND l_commitdate < l_receiptdate)
GROUP BY o_orderpriority ORDER BY o_orderpriority""",
    "q5": """
SELECT n_nationkey, SUM(l_extendedprice * (100 - l_discount)) AS revenue
FROM customer, orders, lineitem, supplier, nation, region
WHERE c_custkey = o_custkey AND l_orderkey = o_orderkey AND l_suppkey = s_suppkey
  AND c_nationkey = s_nationkey AND s_nationkey = n_nationkey
  AND n_regionkey = r_regionkey AND r_name = 'ASIA'
  AND o_orderdate >= '1994-01-01' AND o_orderdate < '1995-01-01'
GROUP BY n_name ORDER BY revenue DESC, n_name""",
    "q7": """
SELECT supp_nation, cust_nation, l_year, SUM(volume) FROM (
  SELECT n1.n_nationkey AS supp_nation, n2.n_nationkey AS cust_nation,
         n1.n_name AS sn, n2.n_name AS cn,
         CAST(strftime('%Y', l_shipdate) AS INT) AS l_year,
         l_extendedprice * (100 - l_discount) AS volume
  FROM supplier, lineitem, orders, customer, nation n1, nation n2
  WHERE s_suppkey = l_suppkey AND o_orderkey = l_orderkey AND c_custkey = o_custkey
    AND s_nationkey = n1.n_nationkey AND c_nationkey = n2.n_nationkey
    AND ((n1.n_name = 'FRANCE' AND n2.n_name = 'GERMANY')
      OR (n1.n_name = 'GERMANY' AND n2.n_name = 'FRANCE'))
    AND l_shipdate BETWEEN '1995-01-01' AND '1996-12-31') AS shipping
GROUP BY sn, cn, l_year ORDER BY sn, cn, l_year""",
    "q8": """
SELECT o_year, SUM(CASE WHEN nation = 'BRAZIL' THEN volume ELSE 0 END), SUM(volume) FROM (
  SELECT CAST(strftime('%Y', o_orderdate) AS INT) AS o_year,
         l_extendedprice * (100 - l_discount) AS volume, n2.n_name AS nation
  FROM part, supplier, lineitem, orders, customer, nation n1, nation n2, region
  WHERE p_partkey = l_partkey AND s_suppkey = l_suppkey AND l_orderkey = o_orderkey
    AND o_custkey = c_custkey AND c_nationkey = n1.n_nationkey
    AND n1.n_regionkey = r_regionkey AND r_name = 'AMERICA'
    AND s_nationkey = n2.n_nationkey
    AND o_orderdate BETWEEN '1995-01-01' AND '1996-12-31'
    AND p_type = 'ECONOMY ANODIZED STEEL') AS all_nations
GROUP BY o_year ORDER BY o_year""",
    "q10": """
SELECT c_custkey, SUM(l_extendedprice * (100 - l_discount)) AS revenue, c_acctbal, n_nationkey
FROM customer, orders, lineitem, nation
WHERE c_custkey = o_custkey AND l_orderkey = o_orderkey
  AND o_orderdate >= '1993-10-01' AND o_orderdate < '1994-01-01'
  AND l_returnflag = 'R' AND c_nationkey = n_nationkey
GROUP BY c_custkey, c_acctbal, n_name ORDER BY revenue DESC, c_custkey LIMIT 20""",
    "q10_groups": """
SELECT COUNT(DISTINCT c_custkey) FROM customer, orders, lineitem
WHERE c_custkey = o_custkey AND l_orderkey = o_orderkey
  AND o_orderdate >= '1993-10-01' AND o_orderdate < '1994-01-01' AND l_returnflag = 'R'""",
    "q11": """
SELECT ps_partkey, SUM(ps_supplycost * ps_availqty) AS value
FROM partsupp, supplier, nation
WHERE ps_suppkey = s_suppkey AND s_nationkey = n_nationkey AND n_name = 'GERMANY'
GROUP BY ps_partkey
HAVING SUM(ps_supplycost * ps_availqty) * 10000 > (
  SELECT SUM(ps_supplycost * ps_availqty) FROM partsupp, supplier, nation
  WHERE ps_suppkey = s_suppkey AND s_nationkey = n_nationkey AND n_name = 'GERMANY')
ORDER BY value DESC, ps_partkey""",
    "q11_total": """
SELECT SUM(ps_supplycost * ps_availqty) FROM partsupp, supplier, nation
WHERE ps_suppkey = s_suppkey AND s_nationkey = n_nationkey AND n_name = 'GERMANY'""",
    "q12": """
SELECT l_shipmode,
  SUM(CASE WHEN o_orderpriority = '1-URGENT' OR o_orderpriority = '2-HIGH' THEN 1 ELSE 0 END),
  SUM(CASE WHEN o_orderpriority <> '1-URGENT' AND o_orderpriority <> '2-HIGH' THEN 1 ELSE 0 END)
FROM orders, lineitem
WHERE o_orderkey = l_orderkey AND l_shipmode IN ('MAIL', 'SHIP')
  AND l_commitdate < l_receiptdate AND l_shipdate < l_commitdate
  AND l_receiptdate >= '1994-01-01' AND l_receiptdate < '1995-01-01'
GROUP BY l_shipmode ORDER BY l_shipmode""",
    "q14": """
SELECT SUM(CASE WHEN p_type GLOB 'PROMO*' THEN l_extendedprice * (100 - l_discount) ELSE 0 END),
       SUM(l_extendedprice * (100 - l_discount))
FROM lineitem, part
WHERE l_partkey = p_partkey AND l_shipdate >= '1995-09-01' AND l_shipdate < '1995-10-01'""",
    "q15": """
WITH revenue0 AS (
  SELECT l_suppkey AS supplier_no, SUM(l_extendedprice * (100 - l_discount)) AS total_revenue
  FROM lineitem WHERE l_shipdate >= '1996-01-01' AND l_shipdate < '1996-04-01'
  GROUP BY l_suppkey)
SELECT s_suppkey, total_revenue FROM supplier, revenue0
WHERE s_suppkey = supplier_no AND total_revenue = (SELECT MAX(total_revenue) FROM revenue0)
ORDER BY s_suppkey""",
    "q16": """
SELECT CAST(substr(p_brand, 7) AS INT), p_type, p_size, COUNT(DISTINCT ps_suppkey) AS supplier_cnt
FROM partsupp, part
WHERE p_partkey = ps_partkey AND p_brand <> 'Brand#45'
  AND p_type NOT GLOB 'MEDIUM POLISHED*'
  AND p_size IN (49, 14, 23, 45, 19, 3, 36, 9)
  AND ps_suppkey NOT IN (SELECT s_suppkey FROM supplier
                         WHERE s_comment GLOB '*Customer*Complaints*')
GROUP BY p_brand, p_type, p_size
ORDER BY supplier_cnt DESC, p_brand, p_type, p_size""",
    "q17": """
SELECT SUM(l_extendedprice) FROM lineitem, part
WHERE p_partkey = l_partkey AND p_brand = 'Brand#23' AND p_container = 'MED BOX'
  AND l_quantity * 5 * (SELECT COUNT(*) FROM lineitem l2 WHERE l2.l_partkey = p_partkey)
      < (SELECT SUM(l_quantity) FROM lineitem l2 WHERE l2.l_partkey = p_partkey)""",
    "q18": """
SELECT c_custkey, o_orderkey,
       CAST(julianday(o_orderdate) - julianday('1992-01-01') AS INT), o_totalprice,
       SUM(l_quantity)
FROM customer, orders, lineitem
WHERE o_orderkey IN (SELECT l_orderkey FROM lineitem GROUP BY l_orderkey
                     HAVING SUM(l_quantity) > 300)
  AND c_custkey = o_custkey AND o_orderkey = l_orderkey
GROUP BY c_custkey, o_orderkey, o_orderdate, o_totalprice
ORDER BY o_totalprice DESC, o_orderdate, o_orderkey LIMIT 100""",
    "q18_250": """
SELECT c_custkey, o_orderkey,
       CAST(julianday(o_orderdate) - julianday('1992-01-01') AS INT), o_totalprice,
       SUM(l_quantity)
FROM customer, orders, lineitem
WHERE o_orderkey IN (SELECT l_orderkey FROM lineitem GROUP BY l_orderkey
                     HAVING SUM(l_quantity) > 250)
  AND c_custkey = o_custkey AND o_orderkey = l_orderkey
GROUP BY c_custkey, o_orderkey, o_orderdate, o_totalprice
ORDER BY o_totalprice DESC, o_orderdate, o_orderkey LIMIT 100""",
    "q19": """
SELECT SUM(l_extendedprice * (100 - l_discount)) FROM lineitem, part
WHERE (p_partkey = l_partkey AND p_brand = 'Brand#12'
       AND p_container IN ('SM CASE', 'SM BOX', 'SM PACK', 'SM PKG')
       AND l_quantity >= 1 AND l_quantity <= 11 AND p_size BETWEEN 1 AND 5
       AND l_shipmode IN ('AIR', 'REG AIR') AND l_shipinstruct = 'DELIVER IN PERSON')
   OR (p_partkey = l_partkey AND p_brand = 'Brand#23'
       AND p_container IN ('MED BAG', 'MED BOX', 'MED PKG', 'MED PACK')
       AND l_quantity >= 10 AND l_quantity <= 20 AND p_size BETWEEN 1 AND 10
       AND l_shipmode IN ('AIR', 'REG AIR') AND l_shipinstruct = 'DELIVER IN PERSON')
   OR (p_partkey = l_partkey AND p_brand = 'Brand#34'
       AND p_container IN ('LG CASE', 'LG BOX', 'LG PACK', 'LG PKG')
       AND l_quantity >= 20 AND l_quantity <= 30 AND p_size BETWEEN 1 AND 15
       AND l_shipmode IN ('AIR', 'REG AIR') AND l_shipinstruct = 'DELIVER IN PERSON')""",
    "q20": """
SELECT s_suppkey FROM supplier, nation
WHERE s_suppkey IN (
    SELECT ps_suppkey FROM partsupp
    WHERE ps_partkey IN (SELECT p_partkey FROM part WHERE p_name GLOB 'forest*')
      AND ps_availqty * 2 > (SELECT SUM(l_quantity) FROM lineitem
                             WHERE l_partkey = ps_partkey AND l_suppkey = ps_suppkey
                               AND l_shipdate >= '1994-01-01' AND l_shipdate < '1995-01-01'))
  AND s_nationkey = n_nationkey AND n_name = 'CANADA'
ORDER BY s_name""",
    "q21": """
SELECT s_suppkey, COUNT(*) AS numwait
FROM supplier, lineitem l1, orders, nation
WHERE s_suppkey = l1.l_suppkey AND o_orderkey = l1.l_orderkey AND o_orderstatus = 'F'
  AND l1.l_receiptdate > l1.l_commitdate
  AND EXISTS (SELECT * FROM lineitem l2 WHERE l2.l_orderkey = l1.l_orderkey
              AND l2.l_suppkey <> l1.l_suppkey)
  AND NOT EXISTS (SELECT * FROM lineitem l3 WHERE l3.l_orderkey = l1.l_orderkey
                  AND l3.l_suppkey <> l1.l_suppkey AND l3.l_receiptdate > l3.l_commitdate)
  AND s_nationkey = n_nationkey AND n_name = 'SAUDI ARABIA'
GROUP BY s_name ORDER BY numwait DESC, s_name LIMIT 100""",
    "q22": """
SELECT cntrycode, COUNT(*), SUM(c_acctbal) FROM (
  SELECT CAST(substr(c_phone, 1, 2) AS INT) AS cntrycode, c_acctbal FROM customer
  WHERE substr(c_phone, 1, 2) IN ('13', '31', '23', '29', '30', '18', '17')
    AND c_acctbal * (SELECT COUNT(*) FROM customer WHERE c_acctbal > 0
                     AND substr(c_phone, 1, 2) IN ('13', '31', '23', '29', '30', '18', '17'))
        > (SELECT SUM(c_acctbal) FROM customer WHERE c_acctbal > 0
           AND substr(c_phone, 1, 2) IN ('13', '31', '23', '29', '30', '18', '17'))
    AND NOT EXISTS (SELECT * FROM orders WHERE o_custkey = c_custkey)) AS custsale
GROUP BY cntrycode ORDER BY cntrycode""",
}


def run_case(sf, seed):
    db = tpch.generate(sf, seed)
    con = load(db)
    out = {}
    for name, sql in SQL.items():
        rows = con.execute(sql).fetchall()
        if name == "q4":
            rows = [(tpch.O_ORDERPRIORITY.index(a), b) for a, b in rows]
        elif name == "q12":
            rows = [(tpch.L_SHIPMODE.index(a), b, c) for a, b, c in rows]
        elif name == "q16":
            rows = [(a, tpch.P_TYPE.index(b), c, d) for a, b, c, d in rows]
        elif name in ("q14", "q17", "q19", "q10_groups", "q11_total"):
            rows = [tuple(0 if v is None else v for v in rows[0])]
        elif name == "q20":
            rows = [r[0] for r in rows]
        out[name] = [list(r) if isinstance(r, tuple) else r for r in rows]
    return out


def main():
    res = []
    for case in CASES:
        res.append(dict(case, results=run_case(**case)))
        print("case", case, {k: len(v) for k, v in res[-1]["results"].items()})
    with open(OUT, "w") as f:
        json.dump(res, f, indent=0, separators=(",", ":"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
