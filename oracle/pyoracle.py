"""ctypes wrapper of liboracle.so, the CPU restatement in oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

I32P, I64P, U8P, U64P = (C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint8),
                         C.POINTER(C.c_uint64))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        _lib = C.CDLL(LIB)
        _lib.oracle_q6.restype = None
        _lib.oracle_q1.restype = C.c_int
        _lib.oracle_q3.restype = C.c_int
        _lib.oracle_q9.restype = C.c_int
        _lib.oracle_q13.restype = C.c_int
        _lib.oracle_select_lt.restype = C.c_uint64
        _lib.oracle_sum_i32.restype = C.c_int64
        _lib.oracle_sum_i64.restype = C.c_int64
        _lib.oracle_join_count.restype = C.c_uint64
        _lib.oracle_contains.restype = C.c_int
        _lib.oracle_like_two.restype = C.c_int
        _lib.oracle_year_of_day.restype = C.c_int32
    return _lib


def _p(a, t):
    a = np.ascontiguousarray(a)
    return a.ctypes.data_as(C.POINTER(t))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def q6(li, date_lo=731, date_hi=1096, disc_lo=5, disc_hi=7, qty_lt=24, nthreads=1):
    s, d, q, p = (_c(li["l_shipdate"], np.int32), _c(li["l_discount"], np.int32),
                  _c(li["l_quantity"], np.int32), _c(li["l_extendedprice"], np.int64))
    rev, rows = C.c_int64(), C.c_uint64()
    lib().oracle_q6(C.c_uint64(len(s)), _p(s, C.c_int32), _p(d, C.c_int32), _p(q, C.c_int32),
                    _p(p, C.c_int64), date_lo, date_hi, disc_lo, disc_hi, qty_lt, nthreads,
                    C.byref(rev), C.byref(rows))
    return dict(revenue=rev.value, rows=rows.value)


def q1(li, cutoff=2435, nthreads=1):
    cols = [_c(li[k], np.int32) for k in ("l_shipdate", "l_quantity", "l_discount", "l_tax")]
    price = _c(li["l_extendedprice"], np.int64)
    rf, ls = _c(li["l_returnflag"], np.uint8), _c(li["l_linestatus"], np.uint8)
    out = np.zeros(30, np.int64)
    lo = np.zeros(6, np.uint64)
    hi = np.zeros(6, np.int64)
    rc = lib().oracle_q1(C.c_uint64(len(price)), *[_p(x, C.c_int32) for x in cols],
                         _p(price, C.c_int64), _p(rf, C.c_uint8), _p(ls, C.c_uint8), cutoff,
                         nthreads, _p(out, C.c_int64), _p(lo, C.c_uint64), _p(hi, C.c_int64))
    if rc != 0:
        raise OverflowError("Q1 sums could exceed int64 for this table")
    rows = []
    for g in range(6):
        v = out[g * 5:(g + 1) * 5]
        if v[4] == 0:
            continue
        cnt = int(v[4])
        rows.append(dict(returnflag="ANR"[g // 2], linestatus="FO"[g % 2], sum_qty=int(v[0]),
                         sum_base_price=int(v[1]), sum_disc_price=int(v[2]),
                         sum_charge=(int(hi[g]) << 64) | int(lo[g]), sum_disc=int(v[3]),
                         count=cnt, avg_qty=int(v[0]) / cnt, avg_price=int(v[1]) / 100.0 / cnt,
                         avg_disc=int(v[3]) / 100.0 / cnt))
    return rows


def q3(cu, od, li, segment=1, date=1169, top_k=10, nthreads=1):
    ck, cs = _c(cu["c_custkey"], np.int32), _c(cu["c_mktsegment"], np.uint8)
    ok, oc, odt, op = (_c(od[k], np.int32) for k in ("o_orderkey", "o_custkey", "o_orderdate",
                                                     "o_shippriority"))
    lk, ls, ld = (_c(li[k], np.int32) for k in ("l_orderkey", "l_shipdate", "l_discount"))
    lp = _c(li["l_extendedprice"], np.int64)
    r_ok, r_od, r_sp = (np.zeros(max(top_k, 1), np.int32) for _ in range(3))
    r_rev = np.zeros(max(top_k, 1), np.int64)
    groups = C.c_int64()
    m = lib().oracle_q3(C.c_uint64(len(ck)), _p(ck, C.c_int32), _p(cs, C.c_uint8),
                        C.c_uint64(len(ok)), _p(ok, C.c_int32), _p(oc, C.c_int32),
                        _p(odt, C.c_int32), _p(op, C.c_int32), C.c_uint64(len(lk)),
                        _p(lk, C.c_int32), _p(ls, C.c_int32), _p(ld, C.c_int32),
                        _p(lp, C.c_int64), segment, date, top_k, nthreads,
                        _p(r_ok, C.c_int32), _p(r_od, C.c_int32), _p(r_sp, C.c_int32),
                        _p(r_rev, C.c_int64), C.byref(groups))
    rows = [dict(orderkey=int(r_ok[i]), revenue=int(r_rev[i]), orderdate=int(r_od[i]),
                 shippriority=int(r_sp[i])) for i in range(m)]
    return dict(groups=groups.value, rows=rows)


def q9(pa, su, ps, od, li, pattern="green", nthreads=1):
    pk, nm = _c(pa["p_partkey"], np.int32), _c(pa["p_name"], np.uint8)
    sk, sn = _c(su["s_suppkey"], np.int32), _c(su["s_nationkey"], np.int32)
    ppk, psk, pc = (_c(ps["ps_partkey"], np.int32), _c(ps["ps_suppkey"], np.int32),
                    _c(ps["ps_supplycost"], np.int64))
    ok, odt = _c(od["o_orderkey"], np.int32), _c(od["o_orderdate"], np.int32)
    lpk, lsk, lok, lq, ld = (_c(li[k], np.int32) for k in ("l_partkey", "l_suppkey",
                                                           "l_orderkey", "l_quantity",
                                                           "l_discount"))
    lp = _c(li["l_extendedprice"], np.int64)
    mx = 25 * 64
    rn, ry, rp = np.zeros(mx, np.int32), np.zeros(mx, np.int32), np.zeros(mx, np.int64)
    m = lib().oracle_q9(C.c_uint64(len(pk)), _p(pk, C.c_int32), _p(nm, C.c_uint8), nm.shape[1],
                        C.c_uint64(len(sk)), _p(sk, C.c_int32), _p(sn, C.c_int32),
                        C.c_uint64(len(ppk)), _p(ppk, C.c_int32), _p(psk, C.c_int32),
                        _p(pc, C.c_int64), C.c_uint64(len(ok)), _p(ok, C.c_int32),
                        _p(odt, C.c_int32), C.c_uint64(len(lpk)), _p(lpk, C.c_int32),
                        _p(lsk, C.c_int32), _p(lok, C.c_int32), _p(lq, C.c_int32),
                        _p(ld, C.c_int32), _p(lp, C.c_int64), pattern.encode(), nthreads,
                        _p(rn, C.c_int32), _p(ry, C.c_int32), _p(rp, C.c_int64), mx)
    return [dict(nationkey=int(rn[i]), year=int(ry[i]), profit=int(rp[i])) for i in range(m)]


def q13(cu, od, word1="special", word2="requests", nthreads=1):
    ck = _c(cu["c_custkey"], np.int32)
    oc, com = _c(od["o_custkey"], np.int32), _c(od["o_comment"], np.uint8)
    mx = 4096
    rc, rd = np.zeros(mx, np.int32), np.zeros(mx, np.int64)
    m = lib().oracle_q13(C.c_uint64(len(ck)), _p(ck, C.c_int32), C.c_uint64(len(oc)),
                         _p(oc, C.c_int32), _p(com, C.c_uint8), com.shape[1], word1.encode(),
                         word2.encode(), nthreads, _p(rc, C.c_int32), _p(rd, C.c_int64), mx)
    return [dict(c_count=int(rc[i]), custdist=int(rd[i])) for i in range(m)]


def select_lt(col, value):
    col = _c(col, np.int32)
    bm = np.zeros((len(col) + 63) // 64, np.uint64)
    c = lib().oracle_select_lt(C.c_uint64(len(col)), _p(col, C.c_int32), value,
                               _p(bm, C.c_uint64))
    return dict(count=int(c), bitmap=bm)


def sum_col(col):
    if col.dtype == np.int64:
        return int(lib().oracle_sum_i64(C.c_uint64(len(col)), _p(col, C.c_int64)))
    col = _c(col, np.int32)
    return int(lib().oracle_sum_i32(C.c_uint64(len(col)), _p(col, C.c_int32)))


def join_count(build, probe):
    b, p = _c(build, np.int32), _c(probe, np.int32)
    return int(lib().oracle_join_count(C.c_uint64(len(b)), _p(b, C.c_int32),
                                       C.c_uint64(len(p)), _p(p, C.c_int32)))


def contains(field, pattern):
    f = _c(field, np.uint8)
    return bool(lib().oracle_contains(_p(f, C.c_uint8), len(f), pattern.encode()))


def like_two(field, w1, w2):
    f = _c(field, np.uint8)
    return bool(lib().oracle_like_two(_p(f, C.c_uint8), len(f), w1.encode(), w2.encode()))


def year_of_day(day):
    return int(lib().oracle_year_of_day(day))
