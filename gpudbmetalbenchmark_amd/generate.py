"""Seeded synthetic tables as numpy columns, produced by libgdb.so's host-side
generator (gdb_gen_*). The shapes are the benchmark configurations':

  flat fact (Q6, Q1):     lineitem_flat(seed, n)
  joined set (Q3, Q9, Q13): customer, orders (+80-byte comments), lineitem
                          joined to orders (1..7 lines per order), part,
                          partsupp (4 suppliers per part), supplier.
"""
import ctypes as C
import os

import numpy as np

from . import abi

NTHREADS = int(os.environ.get("GDB_GEN_THREADS", "0"))  # 0: all cores


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def lineitem_flat(seed, n, row0=0, receipt=False, empty=np.empty):
    """`empty(n, dtype)` allocates each column (e.g. Engine.pinned_empty)."""
    lib = abi.load()
    cols = dict(l_shipdate=empty(n, np.int32), l_quantity=empty(n, np.int32),
                l_discount=empty(n, np.int32), l_tax=empty(n, np.int32),
                l_extendedprice=empty(n, np.int64), l_returnflag=empty(n, np.uint8),
                l_linestatus=empty(n, np.uint8))
    rc = np.empty(n, np.int32) if receipt else None
    abi.check(lib.gdb_gen_lineitem_flat(
        seed, row0, n, _p(cols["l_shipdate"], C.c_int32), _p(cols["l_quantity"], C.c_int32),
        _p(cols["l_discount"], C.c_int32), _p(cols["l_tax"], C.c_int32),
        _p(cols["l_extendedprice"], C.c_int64), _p(cols["l_returnflag"], C.c_uint8),
        _p(cols["l_linestatus"], C.c_uint8), _p(rc, C.c_int32), NTHREADS))
    if receipt:
        cols["l_receiptdate"] = rc
    return cols


def customer(seed, n):
    lib = abi.load()
    c = dict(c_custkey=np.empty(n, np.int32), c_mktsegment=np.empty(n, np.uint8))
    abi.check(lib.gdb_gen_customer(seed, n, _p(c["c_custkey"], C.c_int32),
                                   _p(c["c_mktsegment"], C.c_uint8), NTHREADS))
    return c


def orders(seed, n, ncustomers, comments=False, order0=0, empty=np.empty):
    """`empty(shape, dtype)` allocates each column (e.g. Engine.pinned_empty)."""
    lib = abi.load()
    o = dict(o_orderkey=empty(n, np.int32), o_custkey=empty(n, np.int32),
             o_orderdate=empty(n, np.int32), o_shippriority=empty(n, np.int32))
    com = empty((n, 80), np.uint8) if comments else None
    abi.check(lib.gdb_gen_orders(seed, order0, n, ncustomers, _p(o["o_orderkey"], C.c_int32),
                                 _p(o["o_custkey"], C.c_int32), _p(o["o_orderdate"], C.c_int32),
                                 _p(o["o_shippriority"], C.c_int32), _p(com, C.c_uint8),
                                 NTHREADS))
    if comments:
        o["o_comment"] = com
    return o


def lineitem_joined(seed, norders, nparts, nsuppliers, order0=0, empty=np.empty):
    """`empty(n, dtype)` allocates each column (e.g. Engine.pinned_empty)."""
    lib = abi.load()
    n = lib.gdb_gen_lineitem_rows(seed, order0, norders)
    li = dict(l_orderkey=empty(n, np.int32), l_partkey=empty(n, np.int32),
              l_suppkey=empty(n, np.int32), l_quantity=empty(n, np.int32),
              l_discount=empty(n, np.int32), l_extendedprice=empty(n, np.int64),
              l_shipdate=empty(n, np.int32))
    abi.check(lib.gdb_gen_lineitem_joined(
        seed, order0, norders, nparts, nsuppliers, _p(li["l_orderkey"], C.c_int32),
        _p(li["l_partkey"], C.c_int32), _p(li["l_suppkey"], C.c_int32),
        _p(li["l_quantity"], C.c_int32), _p(li["l_discount"], C.c_int32),
        _p(li["l_extendedprice"], C.c_int64), _p(li["l_shipdate"], C.c_int32), NTHREADS))
    return li


def part(seed, n):
    lib = abi.load()
    p = dict(p_partkey=np.empty(n, np.int32), p_name=np.empty((n, 55), np.uint8))
    abi.check(lib.gdb_gen_part(seed, n, _p(p["p_partkey"], C.c_int32),
                               _p(p["p_name"], C.c_uint8), NTHREADS))
    return p


def partsupp(seed, nparts, nsuppliers):
    lib = abi.load()
    n = nparts * 4
    ps = dict(ps_partkey=np.empty(n, np.int32), ps_suppkey=np.empty(n, np.int32),
              ps_supplycost=np.empty(n, np.int64))
    abi.check(lib.gdb_gen_partsupp(seed, nparts, nsuppliers, _p(ps["ps_partkey"], C.c_int32),
                                   _p(ps["ps_suppkey"], C.c_int32),
                                   _p(ps["ps_supplycost"], C.c_int64), NTHREADS))
    return ps


def supplier(seed, n):
    lib = abi.load()
    s = dict(s_suppkey=np.empty(n, np.int32), s_nationkey=np.empty(n, np.int32))
    abi.check(lib.gdb_gen_supplier(seed, n, _p(s["s_suppkey"], C.c_int32),
                                   _p(s["s_nationkey"], C.c_int32), NTHREADS))
    return s


def join_set(seed, customers, orders_n, parts, suppliers, comments=True):
    """The customer/orders/lineitem/part/partsupp/supplier tables of one seed."""
    return dict(customer=customer(seed, customers),
                orders=orders(seed, orders_n, customers, comments=comments),
                lineitem=lineitem_joined(seed, orders_n, parts, suppliers),
                part=part(seed, parts), partsupp=partsupp(seed, parts, suppliers),
                supplier=supplier(seed, suppliers))
