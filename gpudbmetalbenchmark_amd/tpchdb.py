"""Resident TPC-H database and the remaining 17 pipelines over libgdb.so.

`TpchDb(engine, tables)` uploads the tables of `tpch.generate(...)` (or of
`tbl.load`-ed dbgen files with the same columns) into HBM once; each method
runs one pipeline through the C ABI of include/gdb_tpch.h (gdb_q2 ... gdb_q22)
and returns rows in the same form as the CPU oracle, with the reference SQL's
default parameters (reference sql/qN.sql; runQNBenchmark in src/qN.cpp).
Text predicates are evaluated against the dictionaries of `tpch` on the host
(LIKE 'PROMO%' -> the set of p_type codes it selects) and passed as code masks.
Nothing here computes a result on the CPU beyond ordering the few rows a
pipeline returns; a missing library or GPU raises.
"""
import ctypes as C

import numpy as np

from . import abi, tpch
from .abi import check, tref

TABLES = ("lineitem", "orders", "customer", "part", "partsupp", "supplier", "nation")


def _lib():
    return abi.load()


_Q11_DT = np.dtype([("partkey", np.int32), ("pad", np.int32), ("value", np.int64)])


def _mask3(codes):
    return abi.M3(*tpch.code_mask(codes, 3))


class TpchDb:
    """stream_lineitem=True keeps lineitem in host memory (gdb_table_create_host):
    the pipelines then stream it through the GPU in chunks (the reference's SF100
    mode), every other table resident."""

    def __init__(self, engine, tables, stream_lineitem=False):
        self.engine = engine
        self.tables = {}
        for name in TABLES:
            if name in tables:
                make = engine.host_table if (stream_lineitem and name == "lineitem") else engine.table
                self.tables[name] = make(tables[name])
        self.db = abi.TpchDb(*[self.tables[n].handle.value if n in self.tables else None
                               for n in TABLES])

    def free(self):
        for t in self.tables.values():
            t.free()
        self.tables = {}

    def _call(self, name, *args):
        check(getattr(_lib(), name)(self.engine._ctx, C.byref(self.db), *args))

    # ---- pipelines (defaults: the reference SQL's constants) ---------------
    def q2(self, size=15, type_like="%BRASS", region="EUROPE", top_k=100, timing=None):
        p = abi.Q2Params(size, tpch.R_NAME.index(region), top_k, 0,
                         _mask3(tpch.codes_like(tpch.P_TYPE, type_like)))
        r = abi.Q2Result()
        self._call("gdb_q2", C.byref(p), C.byref(r), tref(timing))
        return [(x.s_acctbal, x.suppkey, x.nationkey, x.partkey, x.mfgr) for x in r.rows[:r.nrows]]

    def q4(self, date_lo=547, date_hi=639, timing=None):
        r = abi.Q4Result()
        self._call("gdb_q4", C.byref(abi.Q4Params(date_lo, date_hi)), C.byref(r), tref(timing))
        return [(i, int(c)) for i, c in enumerate(r.counts) if c]

    def q5(self, region="ASIA", date_lo=731, date_hi=1096, timing=None):
        r = abi.Q5Result()
        self._call("gdb_q5", C.byref(abi.Q5Params(tpch.R_NAME.index(region), date_lo, date_hi)),
                   C.byref(r), tref(timing))
        rows = [(n, int(r.revenue[n])) for n in range(25) if (r.has >> n) & 1]
        rows.sort(key=lambda x: (-x[1], tpch.N_NAME[x[0]]))
        return rows

    def q7(self, nation1="FRANCE", nation2="GERMANY", date_lo=1096, date_hi=1826, timing=None):
        n1, n2 = tpch.N_NAME.index(nation1), tpch.N_NAME.index(nation2)
        r = abi.Q7Result()
        self._call("gdb_q7", C.byref(abi.Q7Params(n1, n2, date_lo, date_hi)), C.byref(r),
                   tref(timing))
        rows = []
        for d, (sn, cn) in enumerate(((n1, n2), (n2, n1))):
            if d == 1 and n1 == n2:
                break
            for y in range(8):
                if (r.has[d] >> y) & 1:
                    rows.append((sn, cn, 1992 + y, int(r.revenue[d][y])))
        rows.sort(key=lambda x: (tpch.N_NAME[x[0]], tpch.N_NAME[x[1]], x[2]))
        return rows

    def q8(self, nation="BRAZIL", region="AMERICA", type_name="ECONOMY ANODIZED STEEL",
           date_lo=1096, date_hi=1826, timing=None):
        r = abi.Q8Result()
        p = abi.Q8Params(tpch.N_NAME.index(nation), tpch.R_NAME.index(region),
                         tpch.P_TYPE.index(type_name), date_lo, date_hi)
        self._call("gdb_q8", C.byref(p), C.byref(r), tref(timing))
        return [(1992 + y, int(r.nation_volume[y]), int(r.total_volume[y]))
                for y in range(8) if (r.has >> y) & 1]

    def q10(self, date_lo=639, date_hi=731, top_k=20, returnflag="R", timing=None):
        r = abi.Q10Result()
        p = abi.Q10Params(date_lo, date_hi, top_k, returnflag.encode())
        self._call("gdb_q10", C.byref(p), C.byref(r), tref(timing))
        return [(x.custkey, x.revenue, x.acctbal, x.nationkey) for x in r.rows[:r.nrows]], \
            int(r.groups)

    def q11(self, nation="GERMANY", frac_num=1, frac_den=10000, cap=1 << 20, timing=None):
        rows = np.empty(max(cap, 1), _Q11_DT)
        n, total = C.c_int64(), C.c_int64()
        p = abi.Q11Params(tpch.N_NAME.index(nation), 0, frac_num, frac_den)
        self._call("gdb_q11", C.byref(p), rows.ctypes.data_as(C.POINTER(abi.Q11Row)), cap,
                   C.byref(n), C.byref(total), tref(timing))
        r = rows[:min(n.value, cap)]
        return list(zip(r["partkey"].tolist(), r["value"].tolist())), total.value

    def q12(self, mode1="MAIL", mode2="SHIP", date_lo=731, date_hi=1096,
            high=("1-URGENT", "2-HIGH"), timing=None):
        m = [tpch.L_SHIPMODE.index(mode1), tpch.L_SHIPMODE.index(mode2)]
        mask = sum(1 << tpch.O_ORDERPRIORITY.index(h) for h in high)
        r = abi.Q12Result()
        self._call("gdb_q12", C.byref(abi.Q12Params((C.c_int32 * 2)(*m), date_lo, date_hi, mask)),
                   C.byref(r), tref(timing))
        rows = {}
        for i in range(2):
            if r.high[i] or r.low[i]:
                rows[m[i]] = (m[i], int(r.high[i]), int(r.low[i]))
        return sorted(rows.values(), key=lambda x: tpch.L_SHIPMODE[x[0]])

    def q14(self, date_lo=1339, date_hi=1369, type_like="PROMO%", timing=None):
        r = abi.Q14Result()
        p = abi.Q14Params(date_lo, date_hi, _mask3(tpch.codes_like(tpch.P_TYPE, type_like)))
        self._call("gdb_q14", C.byref(p), C.byref(r), tref(timing))
        return r.promo, r.total

    def q15(self, date_lo=1461, date_hi=1552, timing=None):
        r = abi.Q15Result()
        self._call("gdb_q15", C.byref(abi.Q15Params(date_lo, date_hi)), C.byref(r), tref(timing))
        return [(int(r.suppkey[i]), r.max_revenue) for i in range(r.nrows)]

    def q16(self, brand="Brand#45", type_like="MEDIUM POLISHED%",
            sizes=(49, 14, 23, 45, 19, 3, 36, 9), complaint=("Customer", "Complaints"),
            cap=1 << 20, timing=None):
        rows = np.empty((max(cap, 1), 4), np.int32)  # gdb_q16_row: four int32
        n = C.c_int64()
        p = abi.Q16Params(tpch.brand_code(brand), len(sizes), (C.c_int32 * 8)(*sizes),
                          _mask3(tpch.codes_like(tpch.P_TYPE, type_like)),
                          complaint[0].encode(), complaint[1].encode())
        self._call("gdb_q16", C.byref(p), rows.ctypes.data_as(C.POINTER(abi.Q16Row)), cap,
                   C.byref(n), tref(timing))
        return rows[:min(n.value, cap)]  # (brand, type, size, supplier_cnt) per row

    def q17(self, brand="Brand#23", container="MED BOX", frac_num=1, frac_den=5, timing=None):
        r = abi.Q17Result()
        p = abi.Q17Params(tpch.brand_code(brand), tpch.P_CONTAINER.index(container), frac_num,
                          frac_den)
        self._call("gdb_q17", C.byref(p), C.byref(r), tref(timing))
        return r.sum_price

    def q18(self, quantity=300, top_k=100, timing=None):
        r = abi.Q18Result()
        self._call("gdb_q18", C.byref(abi.Q18Params(quantity, top_k)), C.byref(r), tref(timing))
        return [(x.custkey, x.orderkey, x.orderdate, x.totalprice, x.sum_qty)
                for x in r.rows[:r.nrows]]

    def q19(self, blocks=None, modes=("AIR", "REG AIR"), instruct="DELIVER IN PERSON",
            timing=None):
        if blocks is None:
            blocks = (("Brand#12", ("SM CASE", "SM BOX", "SM PACK", "SM PKG"), 1, 11, 1, 5),
                      ("Brand#23", ("MED BAG", "MED BOX", "MED PKG", "MED PACK"), 10, 20, 1, 10),
                      ("Brand#34", ("LG CASE", "LG BOX", "LG PACK", "LG PKG"), 20, 30, 1, 15))
        p = abi.Q19Params()
        p.nblocks = len(blocks)
        p.mode_mask = sum(1 << tpch.L_SHIPMODE.index(m) for m in modes)
        p.instruct = tpch.L_SHIPINSTRUCT.index(instruct)
        for i, (br, conts, qlo, qhi, slo, shi) in enumerate(blocks):
            cm = sum(1 << tpch.P_CONTAINER.index(c) for c in conts)
            p.blocks[i] = abi.Q19Block(tpch.brand_code(br), qlo, qhi, slo, shi, 0, cm)
        r = abi.Q19Result()
        self._call("gdb_q19", C.byref(p), C.byref(r), tref(timing))
        return r.revenue

    def q20(self, prefix="forest", date_lo=731, date_hi=1096, nation="CANADA", frac_num=1,
            frac_den=2, cap=1 << 20, timing=None):
        out = np.empty(max(cap, 1), np.int32)
        n = C.c_int64()
        p = abi.Q20Params(prefix.encode(), date_lo, date_hi, tpch.N_NAME.index(nation),
                          frac_num, frac_den, 0)
        self._call("gdb_q20", C.byref(p), out.ctypes.data_as(abi.I32P), cap, C.byref(n),
                   tref(timing))
        return out[:min(n.value, cap)].tolist()

    def q21(self, nation="SAUDI ARABIA", status="F", top_k=100, timing=None):
        r = abi.Q21Result()
        p = abi.Q21Params(tpch.N_NAME.index(nation), top_k, status.encode())
        self._call("gdb_q21", C.byref(p), C.byref(r), tref(timing))
        return [(x.suppkey, x.numwait) for x in r.rows[:r.nrows]]

    def q22(self, codes=(13, 31, 23, 29, 30, 18, 17), timing=None):
        r = abi.Q22Result()
        p = abi.Q22Params(len(codes), (C.c_int32 * 7)(*codes))
        self._call("gdb_q22", C.byref(p), C.byref(r), tref(timing))
        return [(x.cntrycode, x.numcust, x.totacctbal) for x in r.rows[:r.nrows]]
