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
