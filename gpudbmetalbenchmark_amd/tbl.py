"""TPC-H `.tbl` files and the engine's binary column format.

`load(path, table)` parses a dbgen `.tbl` file (pipe-delimited, one row per
line) with the library's multi-threaded parser (gdb_tbl_count_rows /
gdb_tbl_parse) into the engine's exact columns: money in int64 cents,
discount and tax in int32 hundredths, dates as day numbers from 1992-01-01,
p_name / o_comment as fixed-width NUL-padded bytes, c_mktsegment as the
segment code 0..4. It replaces the reference's per-query column loaders
(src/infra.h:59-171, e.g. src/q6.cpp:13-18), which parse decimals into
float32.

`dump(path, table, cols)` writes columns back as a `.tbl` file (the inverse,
for exports and tests). `save_columns` / `load_columns` keep parsed columns
as one raw little-endian file per column plus a JSON manifest, so a dataset
is parsed once and then mapped straight into numpy.
"""
import ctypes as C
import datetime
import json
import os

import numpy as np

from . import abi

# (engine column, 0-based .tbl field, kind, width) per TPC-H table
SCHEMAS = {
    "lineitem": [("l_orderkey", 0, abi.TBL_INT32, 0), ("l_partkey", 1, abi.TBL_INT32, 0),
                 ("l_suppkey", 2, abi.TBL_INT32, 0), ("l_quantity", 4, abi.TBL_INT32, 0),
                 ("l_extendedprice", 5, abi.TBL_CENTS, 0),
                 ("l_discount", 6, abi.TBL_HUNDREDTHS, 0), ("l_tax", 7, abi.TBL_HUNDREDTHS, 0),
                 ("l_returnflag", 8, abi.TBL_CHAR, 0), ("l_linestatus", 9, abi.TBL_CHAR, 0),
                 ("l_shipdate", 10, abi.TBL_DATE, 0), ("l_receiptdate", 12, abi.TBL_DATE, 0)],
    "orders": [("o_orderkey", 0, abi.TBL_INT32, 0), ("o_custkey", 1, abi.TBL_INT32, 0),
               ("o_orderdate", 4, abi.TBL_DATE, 0), ("o_shippriority", 7, abi.TBL_INT32, 0),
               ("o_comment", 8, abi.TBL_TEXT, 80)],
    "customer": [("c_custkey", 0, abi.TBL_INT32, 0), ("c_mktsegment", 6, abi.TBL_TEXT, 10)],
    "part": [("p_partkey", 0, abi.TBL_INT32, 0), ("p_name", 1, abi.TBL_TEXT, 55)],
    "partsupp": [("ps_partkey", 0, abi.TBL_INT32, 0), ("ps_suppkey", 1, abi.TBL_INT32, 0),
                 ("ps_supplycost", 3, abi.TBL_CENTS, 0)],
    "supplier": [("s_suppkey", 0, abi.TBL_INT32, 0), ("s_nationkey", 3, abi.TBL_INT32, 0)],
}
# number of '|'-separated fields per line of each dbgen table
FIELDS = {"lineitem": 16, "orders": 9, "customer": 8, "part": 9, "partsupp": 5, "supplier": 7}

_DTYPE = {abi.TBL_INT32: np.int32, abi.TBL_DATE: np.int32, abi.TBL_CENTS: np.int64,
          abi.TBL_HUNDREDTHS: np.int32, abi.TBL_CHAR: np.uint8, abi.TBL_TEXT: np.uint8}
DAY0 = datetime.date(1992, 1, 1)


def segment_names():
    lib = abi.load()
    return [lib.gdb_segment_name(i).decode() for i in range(5)]


def load(path, table, columns=None, nthreads=0, empty=np.empty):
    """Columns of one TPC-H table from its .tbl file, as a dict of numpy
    arrays named and typed as the engine's tables expect."""
    if table not in SCHEMAS:
        raise ValueError(f"unknown table {table!r}; one of {sorted(SCHEMAS)}")
    lib = abi.load()
    spec = [c for c in SCHEMAS[table] if columns is None or c[0] in columns]
    bpath = os.fsencode(path)
    rows = C.c_uint64()
    abi.check(lib.gdb_tbl_count_rows(bpath, nthreads, C.byref(rows)))
    n = rows.value
    out = {}
    fields = (abi.TblField * len(spec))()
    for i, (name, field, kind, width) in enumerate(spec):
        shape = (n, width) if kind == abi.TBL_TEXT else (n,)
        arr = empty(shape, _DTYPE[kind])
        out[name] = arr
        fields[i] = abi.TblField(field, kind, width, arr.ctypes.data if n else None)
    if n:
        abi.check(lib.gdb_tbl_parse(bpath, fields, len(spec), n, nthreads))
    if "c_mktsegment" in out:  # segment names -> codes 0..4 (255: other)
        text = out["c_mktsegment"]
        code = np.full(n, 255, np.uint8)
        for i, s in enumerate(segment_names()):
            pat = np.zeros(10, np.uint8)
            pat[:len(s)] = np.frombuffer(s.encode(), np.uint8)
            code[(text == pat).all(axis=1)] = i
        out["c_mktsegment"] = code
    return out


def _date_text(day):
    return (DAY0 + datetime.timedelta(days=int(day))).isoformat()


def _money_text(v):
    v = int(v)
    sign = "-" if v < 0 else ""
    return f"{sign}{abs(v) // 100}.{abs(v) % 100:02d}"


def _bytes_text(row):
    b = bytes(row)
    return b.split(b"\0", 1)[0].decode("latin-1")


def dump(path, table, cols):
    """Writes engine columns as a .tbl file; fields the engine does not keep
    are written empty."""
    spec = {field: (name, kind) for name, field, kind, _ in SCHEMAS[table]}
    segs = segment_names()
    n = len(next(iter(cols.values())))
    with open(path, "w", encoding="latin-1", newline="\n") as f:
        for r in range(n):
            parts = []
            for field in range(FIELDS[table]):
                if field not in spec or spec[field][0] not in cols:
                    parts.append("")
                    continue
                name, kind = spec[field]
                v = cols[name][r]
                if name == "c_mktsegment":
                    parts.append(segs[int(v)] if int(v) < 5 else "")
                elif kind == abi.TBL_DATE:
                    parts.append(_date_text(v))
                elif kind in (abi.TBL_CENTS, abi.TBL_HUNDREDTHS):
                    parts.append(_money_text(v))
                elif kind == abi.TBL_CHAR:
                    parts.append(chr(int(v)) if int(v) else "")
                elif kind == abi.TBL_TEXT:
                    parts.append(_bytes_text(v))
                else:
                    parts.append(str(int(v)))
            f.write("|".join(parts) + "|\n")


def save_columns(directory, cols):
    """One raw file per column plus manifest.json (name -> dtype, shape)."""
    os.makedirs(directory, exist_ok=True)
    man = {}
    for name, arr in cols.items():
        arr = np.ascontiguousarray(arr)
        arr.tofile(os.path.join(directory, name + ".bin"))
        man[name] = {"dtype": arr.dtype.str, "shape": list(arr.shape)}
    with open(os.path.join(directory, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1, sort_keys=True)


def load_columns(directory, mmap=True):
    """Columns written by save_columns (memory-mapped unless mmap=False)."""
    with open(os.path.join(directory, "manifest.json")) as f:
        man = json.load(f)
    out = {}
    for name, m in man.items():
        p = os.path.join(directory, name + ".bin")
        dt, shape = np.dtype(m["dtype"]), tuple(m["shape"])
        if mmap and int(np.prod(shape)) > 0:
            out[name] = np.memmap(p, dtype=dt, mode="r", shape=shape)
        else:
            out[name] = np.fromfile(p, dtype=dt).reshape(shape)
    return out


# ---- whole TPC-H databases (all eight tables, the layout of tpch.generate) ----
# dbgen's field order per table; None marks a field the engine does not keep.
DB_FIELDS = {
    "region": ["r_regionkey", "r_name", None],
    "nation": ["n_nationkey", "n_name", "n_regionkey", None],
    "part": ["p_partkey", "p_name", "p_mfgr", "p_brand", "p_type", "p_size", "p_container",
             "p_retailprice", None],
    "supplier": ["s_suppkey", "s_name", None, "s_nationkey", "s_phone", "s_acctbal",
                 "s_comment"],
    "partsupp": ["ps_partkey", "ps_suppkey", "ps_availqty", "ps_supplycost", None],
    "customer": ["c_custkey", "c_name", None, "c_nationkey", "c_phone", "c_acctbal",
                 "c_mktsegment", None],
    "orders": ["o_orderkey", "o_custkey", "o_orderstatus", "o_totalprice", "o_orderdate",
               "o_orderpriority", "o_clerk", "o_shippriority", "o_comment"],
    "lineitem": ["l_orderkey", "l_partkey", "l_suppkey", "l_linenumber", "l_quantity",
                 "l_extendedprice", "l_discount", "l_tax", "l_returnflag", "l_linestatus",
                 "l_shipdate", "l_commitdate", "l_receiptdate", "l_shipinstruct", "l_shipmode",
                 None],
}
# engine column -> (parse kind, text width); dictionary columns parse as text first
_DB_KIND = {
    "l_orderkey": abi.TBL_INT32, "l_partkey": abi.TBL_INT32, "l_suppkey": abi.TBL_INT32,
    "l_quantity": abi.TBL_INT32, "l_extendedprice": abi.TBL_CENTS,
    "l_discount": abi.TBL_HUNDREDTHS, "l_tax": abi.TBL_HUNDREDTHS, "l_returnflag": abi.TBL_CHAR,
    "l_linestatus": abi.TBL_CHAR, "l_shipdate": abi.TBL_DATE, "l_commitdate": abi.TBL_DATE,
    "l_receiptdate": abi.TBL_DATE, "l_shipinstruct": (abi.TBL_TEXT, 25),
    "l_shipmode": (abi.TBL_TEXT, 10),
    "o_orderkey": abi.TBL_INT32, "o_custkey": abi.TBL_INT32, "o_orderstatus": abi.TBL_CHAR,
    "o_totalprice": abi.TBL_CENTS, "o_orderdate": abi.TBL_DATE,
    "o_orderpriority": (abi.TBL_TEXT, 15), "o_shippriority": abi.TBL_INT32,
    "o_comment": (abi.TBL_TEXT, 80),
    "c_custkey": abi.TBL_INT32, "c_nationkey": abi.TBL_INT32, "c_acctbal": abi.TBL_CENTS,
    "c_mktsegment": (abi.TBL_TEXT, 10),
    "p_partkey": abi.TBL_INT32, "p_name": (abi.TBL_TEXT, 55), "p_mfgr": (abi.TBL_TEXT, 25),
    "p_brand": (abi.TBL_TEXT, 10), "p_type": (abi.TBL_TEXT, 25), "p_size": abi.TBL_INT32,
    "p_container": (abi.TBL_TEXT, 10), "p_retailprice": abi.TBL_CENTS,
    "ps_partkey": abi.TBL_INT32, "ps_suppkey": abi.TBL_INT32, "ps_availqty": abi.TBL_INT32,
    "ps_supplycost": abi.TBL_CENTS,
    "s_suppkey": abi.TBL_INT32, "s_nationkey": abi.TBL_INT32, "s_acctbal": abi.TBL_CENTS,
    "s_comment": (abi.TBL_TEXT, 48),
    "n_nationkey": abi.TBL_INT32, "n_regionkey": abi.TBL_INT32, "r_regionkey": abi.TBL_INT32,
}


def _dicts():
    from . import tpch
    return {"l_shipinstruct": tpch.L_SHIPINSTRUCT, "l_shipmode": tpch.L_SHIPMODE,
            "o_orderpriority": tpch.O_ORDERPRIORITY, "c_mktsegment": tpch.C_MKTSEGMENT,
            "p_type": tpch.P_TYPE, "p_container": tpch.P_CONTAINER}


def _codes(text, dictionary):
    """Fixed-width text rows -> dictionary codes (255 for values not in it)."""
    out = np.full(len(text), 255, np.uint8)
    w = text.shape[1]
    for i, s in enumerate(dictionary):
        pat = np.zeros(w, np.uint8)
        b = s.encode()[:w]
        pat[:len(b)] = np.frombuffer(b, np.uint8)
        out[(text == pat).all(axis=1)] = i
    return out


def _tail_number(text):
    """'Brand#23' / 'Manufacturer#4' rows -> 23 / 4 (digits after the last '#')."""
    out = np.zeros(len(text), np.uint8)
    for r, row in enumerate(text):
        s = bytes(row).split(b"\0", 1)[0]
        i = s.rfind(b"#")
        out[r] = int(s[i + 1:]) if i >= 0 and s[i + 1:].isdigit() else 0
    return out


def load_database(directory, nthreads=0):
    """The eight .tbl files of a TPC-H database (dbgen output, or dump_database)
    as the engine's columns, in the layout of tpch.generate."""
    lib = abi.load()
    dicts = _dicts()
    db = {}
    for table, names in DB_FIELDS.items():
        path = os.path.join(directory, table + ".tbl")
        bpath = os.fsencode(path)
        rows = C.c_uint64()
        abi.check(lib.gdb_tbl_count_rows(bpath, nthreads, C.byref(rows)))
        n = rows.value
        spec = [(i, nm) for i, nm in enumerate(names) if nm in _DB_KIND]
        fields = (abi.TblField * len(spec))()
        cols = {}
        for k, (i, nm) in enumerate(spec):
            kind = _DB_KIND[nm]
            kind, width = kind if isinstance(kind, tuple) else (kind, 0)
            arr = np.empty((n, width) if kind == abi.TBL_TEXT else (n,), _DTYPE[kind])
            cols[nm] = arr
            fields[k] = abi.TblField(i, kind, width, arr.ctypes.data if n else None)
        if n:
            abi.check(lib.gdb_tbl_parse(bpath, fields, len(spec), n, nthreads))
        for nm in list(cols):
            if nm in dicts:
                cols[nm] = _codes(cols[nm], dicts[nm])
            elif nm in ("p_brand", "p_mfgr"):
                cols[nm] = _tail_number(cols[nm])
        if table == "orders" and "o_shippriority" not in cols:
            cols["o_shippriority"] = np.zeros(n, np.int32)
        db[table] = cols
    return db


def dump_database(directory, db):
    """Writes a tpch.generate database as dbgen-format .tbl files (one per
    table); text the engine does not keep is synthesised from the keys, as
    dbgen formats it (Supplier#000000001, Customer#..., Clerk#...)."""
    from . import tpch
    os.makedirs(directory, exist_ok=True)
    dicts = _dicts()
    for table, names in DB_FIELDS.items():
        cols = db[table]
        n = len(next(iter(cols.values())))
        key = cols[names[0]]
        with open(os.path.join(directory, table + ".tbl"), "w", encoding="latin-1",
                  newline="\n") as f:
            for r in range(n):
                parts = []
                for nm in names:
                    if nm is None:
                        parts.append("")
                    elif nm == "r_name":
                        parts.append(tpch.R_NAME[int(key[r])])
                    elif nm == "n_name":
                        parts.append(tpch.N_NAME[int(key[r])])
                    elif nm == "s_name":
                        parts.append(f"Supplier#{int(key[r]):09d}")
                    elif nm == "c_name":
                        parts.append(f"Customer#{int(key[r]):09d}")
                    elif nm == "o_clerk":
                        parts.append(f"Clerk#{1 + int(key[r]) % 1000:09d}")
                    elif nm in ("s_phone", "c_phone"):
                        nat = int(cols[nm[0] + "_nationkey"][r])
                        k = int(key[r])
                        parts.append(f"{nat + 10}-{100 + k % 900}-{200 + k % 700}-{1000 + k % 9000}")
                    elif nm == "l_linenumber":
                        parts.append("1")
                    elif nm == "p_mfgr":
                        parts.append(f"Manufacturer#{int(cols[nm][r])}")
                    elif nm == "p_brand":
                        parts.append(tpch.brand_name(cols[nm][r]))
                    elif nm in dicts:
                        parts.append(dicts[nm][int(cols[nm][r])])
                    else:
                        kind = _DB_KIND[nm]
                        kind = kind[0] if isinstance(kind, tuple) else kind
                        v = cols[nm][r]
                        if kind == abi.TBL_DATE:
                            parts.append(_date_text(v))
                        elif kind in (abi.TBL_CENTS, abi.TBL_HUNDREDTHS):
                            parts.append(_money_text(v))
                        elif kind == abi.TBL_CHAR:
                            parts.append(chr(int(v)))
                        elif kind == abi.TBL_TEXT:
                            parts.append(_bytes_text(v))
                        else:
                            parts.append(str(int(v)))
                f.write("|".join(parts) + "|\n")
