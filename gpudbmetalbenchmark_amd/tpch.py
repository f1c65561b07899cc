"""Seeded TPC-H-shaped database: all eight tables, every column the 22 queries read.

The reference runs its 22 pipelines over dbgen `.tbl` files (reference
src/infra.h:231-468; the loaders of src/q2.cpp ... src/q22.cpp). dbgen is not
available offline, so this module synthesises tables with the TPC-H
specification's cardinalities and value distributions (clause 4.2.3) at any
scale factor, deterministically from a seed. Columns use the engine's exact
layout:

  money (p_retailprice, ps_supplycost, s_acctbal, c_acctbal, l_extendedprice,
  o_totalprice)                    int64 cents
  l_discount, l_tax                int32 hundredths
  dates                            int32 day numbers, 1992-01-01 = day 0
  low-cardinality text             uint8 dictionary codes (tables below):
      p_type (150 values), p_container (40), p_brand (Brand#MN -> 10*M+N),
      p_mfgr (1..5), l_shipmode (7), l_shipinstruct (4), o_orderpriority (5),
      c_mktsegment (5), l_returnflag / l_linestatus / o_orderstatus (ASCII)
  free text                        fixed-width NUL-padded bytes:
      p_name 55 B, o_comment 80 B, s_comment 48 B

A predicate on a dictionary-coded column (LIKE 'PROMO%', IN ('MAIL','SHIP'))
becomes a set of codes: `codes_like(DICT, pattern)` evaluates the SQL pattern
against every dictionary entry once. c_phone's country code is
c_nationkey + 10 (clause 4.2.3), so Q22 reads c_nationkey.

Deviations from dbgen, all deliberate: o_orderkey keeps TPC-H's sparse key
pattern (8 of every 32) but dense 1..N keys everywhere else, o_totalprice is
the sum of each line's charge rounded half-up to cents, and comments are drawn
from small vocabularies (with the Q13 and Q16 patterns planted at fixed rates).
"""
import ctypes as C
import functools
import re

import numpy as np

from . import abi

# ---- dictionaries (TPC-H specification clause 4.2.2.13 and 4.2.3) ----------
TYPE_S1 = ["STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO"]
TYPE_S2 = ["ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED"]
TYPE_S3 = ["TIN", "NICKEL", "BRASS", "STEEL", "COPPER"]
P_TYPE = [f"{a} {b} {c}" for a in TYPE_S1 for b in TYPE_S2 for c in TYPE_S3]  # code = index
CONT_S1 = ["SM", "LG", "MED", "JUMBO", "WRAP"]
CONT_S2 = ["CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM"]
P_CONTAINER = [f"{a} {b}" for a in CONT_S1 for b in CONT_S2]
L_SHIPMODE = ["REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"]
L_SHIPINSTRUCT = ["DELIVER IN PERSON", "COLLECT COD", "NONE", "TAKE BACK RETURN"]
O_ORDERPRIORITY = ["1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW"]
C_MKTSEGMENT = ["AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY"]
R_NAME = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
N_NAME = ["ALGERIA", "ARGENTINA", "BRAZIL", "CANADA", "EGYPT", "ETHIOPIA", "FRANCE",
          "GERMANY", "INDIA", "INDONESIA", "IRAN", "IRAQ", "JAPAN", "JORDAN", "KENYA",
          "MOROCCO", "MOZAMBIQUE", "PERU", "CHINA", "ROMANIA", "SAUDI ARABIA", "VIETNAM",
          "RUSSIA", "UNITED KINGDOM", "UNITED STATES"]
N_REGION = [0, 1, 1, 1, 4, 0, 3, 3, 2, 2, 4, 4, 2, 4, 0, 0, 0, 1, 2, 3, 4, 2, 3, 3, 1]

CURRENT_DATE = 1263  # 1995-06-17, the spec's CURRENTDATE
ORDER_DATE_MAX = 2405  # ENDDATE - 151 days (1998-08-02)
S_COMMENT_WIDTH = 48
_S_WORDS = ["furiously", "regular", "deposits", "carefully", "final", "ideas", "blithely",
            "express", "accounts", "pending", "slyly", "ironic", "packages", "quickly",
            "bold", "requests"]

# table / column stream ids of the generator
_T = {"part": 1, "supplier": 2, "partsupp": 3, "customer": 4, "orders": 5, "lineitem": 6}


def brand_code(name):
    """'Brand#23' -> 23."""
    m = re.fullmatch(r"Brand#([1-5])([1-5])", name)
    if not m:
        raise ValueError(f"not a TPC-H brand: {name!r}")
    return int(m.group(1)) * 10 + int(m.group(2))


def brand_name(code):
    return f"Brand#{int(code)}"


@functools.lru_cache(maxsize=256)
def _like_rx(pattern):
    rx = "".join(".*" if ch == "%" else "." if ch == "_" else re.escape(ch) for ch in pattern)
    return re.compile(rx, flags=re.S)


def like(value, pattern):
    """SQL LIKE (case-sensitive, % and _ wildcards) of one string."""
    return _like_rx(pattern).fullmatch(value) is not None


@functools.lru_cache(maxsize=256)
def _codes_like(dictionary, pattern):
    rx = _like_rx(pattern)
    return tuple(i for i, v in enumerate(dictionary) if rx.fullmatch(v))


def codes_like(dictionary, pattern):
    """Codes of the dictionary entries that match the LIKE pattern (memoised per
    dictionary and pattern)."""
    return list(_codes_like(tuple(dictionary), pattern))


def code_mask(codes, words=3):
    """Bit mask (list of `words` uint64) with bit c set for every code c."""
    m = [0] * words
    for c in codes:
        m[c // 64] |= 1 << (c % 64)
    return m


def sizes(sf):
    """Row counts at scale factor sf (TPC-H clause 4.2.5)."""
    return dict(supplier=max(8, int(round(10_000 * sf))),
                part=max(8, int(round(200_000 * sf))),
                customer=max(8, int(round(150_000 * sf))),
                orders=max(8, int(round(1_500_000 * sf))))


def _rng(seed, table, col):
    return np.random.Generator(np.random.PCG64([seed, _T[table], col]))


def _uni(rng, lo, hi, n, dtype=np.int32):
    return rng.integers(lo, hi + 1, size=n, dtype=np.int64).astype(dtype)


def retail_price(partkey):
    """p_retailprice in cents (TPC-H clause 4.2.3)."""
    pk = np.asarray(partkey, np.int64)
    return 90000 + (pk // 10) % 20001 + 100 * (pk % 1000)


def supplier_of(partkey, i, nsupp):
    """The i-th (0..3) supplier of a part (TPC-H clause 4.2.3, PS_SUPPKEY)."""
    pk = np.asarray(partkey, np.int64)
    return ((pk + i * (nsupp // 4 + (pk - 1) // nsupp)) % nsupp + 1).astype(np.int32)


def order_key(idx):
    """Sparse TPC-H order key of the idx-th order: 8 of every 32 keys used."""
    idx = np.asarray(idx, np.int64)
    return ((idx // 8) * 32 + idx % 8 + 1).astype(np.int32)


def _s_comments(seed, n):
    """Supplier comments (48-byte fields): four vocabulary words, with
    'Customer ... Complaints' planted in about 5 per 10,000 rows and
    'Customer ... Recommends' in as many (TPC-H clause 4.2.3, S_COMMENT), plus
    decoys with the words reversed or alone."""
    rng = _rng(seed, "supplier", 9)
    w = rng.integers(0, len(_S_WORDS), size=(n, 4))
    kind = np.zeros(n, np.int8)
    k = max(2, int(round(n * 5 / 10_000)))
    pick = rng.permutation(n)[:4 * k]
    kind[pick[:k]] = 1           # Customer <w> Complaints
    kind[pick[k:2 * k]] = 2      # Customer <w> Recommends
    kind[pick[2 * k:3 * k]] = 3  # Complaints <w> Customer (decoy)
    kind[pick[3 * k:]] = 4       # Customer <w> <w> (decoy)
    out = np.zeros((n, S_COMMENT_WIDTH), np.uint8)
    for r in range(n):
        a, b, c, d = (_S_WORDS[j] for j in w[r])
        kd = kind[r]
        if kd == 0:
            s = f"{a} {b} {c} {d}"
        elif kd == 1:
            s = f"{a} Customer {b} Complaints {c}"
        elif kd == 2:
            s = f"{a} Customer {b} Recommends {c}"
        elif kd == 3:
            s = f"{a} Complaints {b} Customer {c}"
        else:
            s = f"Customer {a} {b} {c}"
        bs = s.encode()[:S_COMMENT_WIDTH]
        out[r, :len(bs)] = np.frombuffer(bs, np.uint8)
    return out


def generate(sf=0.01, seed=7, comments=True):
    """All eight tables as dicts of numpy columns at scale factor `sf`."""
    n = sizes(sf)
    S, P, Cn, O = n["supplier"], n["part"], n["customer"], n["orders"]
    lib = abi.load()

    region = dict(r_regionkey=np.arange(5, dtype=np.int32))
    nation = dict(n_nationkey=np.arange(25, dtype=np.int32),
                  n_regionkey=np.array(N_REGION, np.int32))

    # part
    pk = np.arange(1, P + 1, dtype=np.int32)
    name = np.empty((P, 55), np.uint8)
    dummy = np.empty(P, np.int32)
    abi.check(lib.gdb_gen_part(seed, P, dummy.ctypes.data_as(C.POINTER(C.c_int32)),
                               name.ctypes.data_as(C.POINTER(C.c_uint8)), 0))
    mfgr = _uni(_rng(seed, "part", 1), 1, 5, P, np.uint8)
    brand = (mfgr.astype(np.int32) * 10 + _uni(_rng(seed, "part", 2), 1, 5, P)).astype(np.uint8)
    part = dict(p_partkey=pk, p_name=name, p_mfgr=mfgr, p_brand=brand,
                p_type=_uni(_rng(seed, "part", 3), 0, len(P_TYPE) - 1, P, np.uint8),
                p_size=_uni(_rng(seed, "part", 4), 1, 50, P),
                p_container=_uni(_rng(seed, "part", 5), 0, len(P_CONTAINER) - 1, P, np.uint8),
                p_retailprice=retail_price(pk).astype(np.int64))

    # supplier
    supplier = dict(s_suppkey=np.arange(1, S + 1, dtype=np.int32),
                    s_nationkey=_uni(_rng(seed, "supplier", 1), 0, 24, S),
                    s_acctbal=_uni(_rng(seed, "supplier", 2), -99999, 999999, S, np.int64),
                    s_comment=_s_comments(seed, S))

    # partsupp: 4 rows per part, partkey-major
    ps_pk = np.repeat(pk, 4)
    ps_i = np.tile(np.arange(4, dtype=np.int64), P)
    partsupp = dict(ps_partkey=ps_pk, ps_suppkey=supplier_of(ps_pk, ps_i, S),
                    ps_availqty=_uni(_rng(seed, "partsupp", 1), 1, 9999, 4 * P),
                    ps_supplycost=_uni(_rng(seed, "partsupp", 2), 100, 100000, 4 * P, np.int64))

    # customer
    customer = dict(c_custkey=np.arange(1, Cn + 1, dtype=np.int32),
                    c_nationkey=_uni(_rng(seed, "customer", 1), 0, 24, Cn),
                    c_acctbal=_uni(_rng(seed, "customer", 2), -99999, 999999, Cn, np.int64),
                    c_mktsegment=_uni(_rng(seed, "customer", 3), 0, 4, Cn, np.uint8))

    # orders: custkeys skip multiples of 3 (a third of customers has no orders)
    oidx = np.arange(O, dtype=np.int64)
    okey = order_key(oidx)
    kk = _uni(_rng(seed, "orders", 1), 0, Cn - Cn // 3 - 1, O, np.int64)
    ocust = ((kk // 2) * 3 + kk % 2 + 1).astype(np.int32)
    odate = _uni(_rng(seed, "orders", 2), 0, ORDER_DATE_MAX, O)
    oprio = _uni(_rng(seed, "orders", 3), 0, 4, O, np.uint8)

    # lineitem: 1..7 lines per order, clustered by order
    nl = _uni(_rng(seed, "lineitem", 1), 1, 7, O, np.int64)
    L = int(nl.sum())
    own = np.repeat(oidx, nl)
    lpk = _uni(_rng(seed, "lineitem", 2), 1, P, L)
    lsk = supplier_of(lpk, _uni(_rng(seed, "lineitem", 3), 0, 3, L, np.int64), S)
    qty = _uni(_rng(seed, "lineitem", 4), 1, 50, L)
    price = qty.astype(np.int64) * retail_price(lpk)
    disc = _uni(_rng(seed, "lineitem", 5), 0, 10, L)
    tax = _uni(_rng(seed, "lineitem", 6), 0, 8, L)
    lod = odate[own]
    ship = lod + _uni(_rng(seed, "lineitem", 7), 1, 121, L)
    commit = lod + _uni(_rng(seed, "lineitem", 8), 30, 90, L)
    receipt = ship + _uni(_rng(seed, "lineitem", 9), 1, 30, L)
    ra = np.where(_uni(_rng(seed, "lineitem", 10), 0, 1, L) == 1, ord("R"), ord("A"))
    rflag = np.where(receipt <= CURRENT_DATE, ra, ord("N")).astype(np.uint8)
    lstat = np.where(ship > CURRENT_DATE, ord("O"), ord("F")).astype(np.uint8)
    lineitem = dict(l_orderkey=okey[own], l_partkey=lpk, l_suppkey=lsk, l_quantity=qty,
                    l_extendedprice=price, l_discount=disc, l_tax=tax,
                    l_returnflag=rflag, l_linestatus=lstat, l_shipdate=ship,
                    l_commitdate=commit, l_receiptdate=receipt,
                    l_shipinstruct=_uni(_rng(seed, "lineitem", 11), 0, 3, L, np.uint8),
                    l_shipmode=_uni(_rng(seed, "lineitem", 12), 0, 6, L, np.uint8))

    # order status and total price follow from the lines
    starts = np.concatenate([[0], np.cumsum(nl)[:-1]])
    nf = np.add.reduceat((lstat == ord("F")).astype(np.int64), starts)
    ostat = np.where(nf == nl, ord("F"), np.where(nf == 0, ord("O"), ord("P"))).astype(np.uint8)
    charge = (price * (100 + tax) * (100 - disc) + 5000) // 10000
    total = np.add.reduceat(charge, starts).astype(np.int64)
    orders = dict(o_orderkey=okey, o_custkey=ocust, o_orderstatus=ostat, o_totalprice=total,
                  o_orderdate=odate, o_orderpriority=oprio,
                  o_shippriority=np.zeros(O, np.int32))
    if comments:
        com = np.empty((O, 80), np.uint8)
        tmp = [np.empty(O, np.int32) for _ in range(4)]
        abi.check(lib.gdb_gen_orders(seed, 0, O, Cn, *(t.ctypes.data_as(C.POINTER(C.c_int32))
                                                        for t in tmp),
                                     com.ctypes.data_as(C.POINTER(C.c_uint8)), 0))
        orders["o_comment"] = com
    return dict(region=region, nation=nation, part=part, supplier=supplier,
                partsupp=partsupp, customer=customer, orders=orders, lineitem=lineitem)


def text(field):
    """A fixed-width NUL-padded byte field as str."""
    return bytes(field).split(b"\0", 1)[0].decode()


if __name__ == "__main__":
    import argparse
    import os
    ap = argparse.ArgumentParser(description="Write a seeded TPC-H-shaped database as dbgen "
                                             "format .tbl files (for gdb_bench --data DIR)")
    ap.add_argument("--dump", required=True, help="output directory")
    ap.add_argument("--sf", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args()
    from . import tbl
    tbl.dump_database(a.dump, generate(a.sf, a.seed))
    print("wrote", sorted(os.listdir(a.dump)))
