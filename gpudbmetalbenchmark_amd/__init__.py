"""Python host side of the MI355X query engine: ctypes bindings over libgdb.so.

This mirrors the reference's per-pipeline entry points (runQ6Benchmark,
runQ1Benchmark, ... in reference src/infra.h:874-922, dispatched by name from
src/main.cpp:161-188): one function per pipeline over tables resident in HBM.
All computation happens in libgdb.so's HIP kernels; if the library is missing
or no GPU is visible, calls raise instead of computing anything on the CPU.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import abi
from .abi import (GdbError, Q1Params, Q1Result, Q3Params, Q3Result, Q6Params, Q6Result,
                  Q9Params, Q9Result, Q13Params, Q13Result, Timing)

__all__ = ["Engine", "Table", "Server", "GdbError", "lib", "generate", "request", "LIB_PATH"]

LIB_PATH = abi.LIB_PATH


def lib():
    """The loaded libgdb.so (raises GdbError if it has not been built)."""
    return abi.load()


# numpy views of the integer result rows: one frombuffer + tolist per column
# instead of a ctypes attribute read per field (Q9's 175 rows: 43 -> 29 us).
_Q9_DT = np.dtype({"names": ["nationkey", "year", "profit"], "formats": ["<i4", "<i4", "<i8"],
                   "offsets": [abi.Q9Row.nationkey.offset, abi.Q9Row.year.offset,
                               abi.Q9Row.profit.offset], "itemsize": C.sizeof(abi.Q9Row)})
_Q13_DT = np.dtype({"names": ["c_count", "custdist"], "formats": ["<i4", "<i8"],
                    "offsets": [abi.Q13Row.c_count.offset, abi.Q13Row.custdist.offset],
                    "itemsize": C.sizeof(abi.Q13Row)})


def _columns(res, dt, names):
    a = np.frombuffer(res, dtype=dt, count=res.nrows, offset=type(res).rows.offset)
    return [a[n].tolist() for n in names]


class Table:
    """A columnar table resident in HBM (gdb_table_create), or, with host=True,
    a host-resident table streamed in chunks by the pipelines
    (gdb_table_create_host; the numpy arrays are referenced, not copied)."""

    def __init__(self, engine, columns, nrows=None, host=False):
        self.engine = engine
        self._keep = []
        descs = (abi.ColDesc * len(columns))()
        n = None
        for i, (name, arr) in enumerate(columns.items()):
            arr = np.ascontiguousarray(arr)
            if arr.dtype == np.int32:
                typ, width, rows = abi.GDB_I32, 4, len(arr)
            elif arr.dtype == np.int64:
                typ, width, rows = abi.GDB_I64, 8, len(arr)
            elif arr.dtype == np.uint8 and arr.ndim == 1:
                typ, width, rows = abi.GDB_U8, 1, len(arr)
            elif arr.dtype == np.uint8 and arr.ndim == 2:
                typ, width, rows = abi.GDB_BYTES, arr.shape[1], arr.shape[0]
            else:
                raise GdbError(abi.GDB_E_INVALID, f"column {name}: unsupported dtype {arr.dtype}")
            if n is None:
                n = rows
            elif rows != n:
                raise GdbError(abi.GDB_E_INVALID, f"column {name} has {rows} rows, expected {n}")
            self._keep.append(arr)
            descs[i].name = name.encode()
            descs[i].type = typ
            descs[i].width = width
            descs[i].host_data = arr.ctypes.data_as(C.c_void_p)
        if nrows is not None and columns and nrows != n:
            raise GdbError(abi.GDB_E_INVALID, f"nrows={nrows} but the columns hold {n} rows")
        if n is None:
            n = nrows
        self.nrows = n or 0
        self.names = list(columns)
        self._t = C.c_void_p()
        self.host = host
        create = lib().gdb_table_create_host if host else lib().gdb_table_create
        abi.check(create(engine._ctx, descs, len(columns), self.nrows, C.byref(self._t)))
        engine._tables.add(self)
        if not host:
            self._keep = []

    @property
    def handle(self):
        return self._t

    def free(self):
        """Releases the table; a no-op once freed or once its engine is closed
        (gdb_close releases the context the table belongs to)."""
        if self._t and self.engine._ctx:
            lib().gdb_table_free(self._t)
        self._t = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One GPU context (gdb_open / gdb_close)."""

    def __init__(self, device=0, hbm_budget=0):
        self._ctx = C.c_void_p()
        self._pinned = []
        self._tables = weakref.WeakSet()
        abi.check(lib().gdb_open(device, hbm_budget, C.byref(self._ctx)))

    def close(self):
        if self._ctx:
            for t in list(self._tables):
                t.free()
            for p in self._pinned:
                lib().gdb_host_free(self._ctx, p)
            self._pinned = []
            lib().gdb_close(self._ctx)
            self._ctx = C.c_void_p()

    def pinned_empty(self, shape, dtype):
        """A numpy array in pinned host memory (gdb_host_alloc), freed by close()."""
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) if np.ndim(shape) else int(shape)
        p = C.c_void_p()
        abi.check(lib().gdb_host_alloc(self._ctx, max(1, n * dtype.itemsize), C.byref(p)))
        self._pinned.append(p)
        buf = (C.c_char * max(1, n * dtype.itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape)

    def set_chunk_rows(self, rows):
        abi.check(lib().gdb_set_chunk_rows(self._ctx, rows))

    def host_table(self, columns):
        return Table(self, columns, host=True)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        info = abi.DeviceInfo()
        abi.check(lib().gdb_device_info_get(self._ctx, C.byref(info)))
        return info

    def hbm_used(self):
        return lib().gdb_hbm_used(self._ctx)

    def release_scratch(self):
        """Frees the scratch arenas the pipelines grew (tables stay)."""
        abi.check(lib().gdb_release_scratch(self._ctx))

    def autotune(self, pipeline, tables, reps=3):
        """Times the candidate launch shapes of `pipeline` ('q6', 'q1', 'q3',
        'q9', 'q13', 'join') on `tables` (argument order of the pipeline; join:
        build, probe on their first columns) and keeps the fastest for this
        engine; returns it as a dict."""
        arr = (C.c_void_p * len(tables))(*[t.handle.value for t in tables])
        out = abi.Tuning()
        abi.check(lib().gdb_autotune(self._ctx, abi.PIPE_IDS[pipeline], arr, len(tables), reps,
                                     C.byref(out)))
        return dict(pipeline=pipeline, block=out.block, blocks_per_cu=out.blocks_per_cu,
                    unroll=out.unroll, radix_bits=out.radix_bits, ms=out.ms,
                    candidates=out.candidates)

    def get_tuning(self, pipeline):
        out = abi.Tuning()
        abi.check(lib().gdb_get_tuning(self._ctx, abi.PIPE_IDS[pipeline], C.byref(out)))
        return dict(pipeline=pipeline, block=out.block, blocks_per_cu=out.blocks_per_cu,
                    unroll=out.unroll, radix_bits=out.radix_bits)

    def set_tuning(self, pipeline, block, blocks_per_cu, unroll, radix_bits=0):
        t = abi.Tuning(abi.PIPE_IDS[pipeline], block, blocks_per_cu, unroll, 0.0, 0, radix_bits)
        abi.check(lib().gdb_set_tuning(self._ctx, C.byref(t)))

    def set_kernel_timing(self, on):
        """On (the default): pipelines time their dominant kernel with events
        (Timing.main_kernel_ms). Off: no events between a pipeline's kernels,
        main_kernel_ms = -1."""
        abi.check(lib().gdb_set_kernel_timing(self._ctx, 1 if on else 0))

    def bandwidth_probe(self, nbytes=1 << 30, reps=10):
        """(read GB/s, copy GB/s) of plain streaming kernels over nbytes."""
        r, c = C.c_double(), C.c_double()
        abi.check(lib().gdb_bandwidth_probe(self._ctx, nbytes, reps, C.byref(r), C.byref(c)))
        return r.value, c.value

    def h2d_probe(self, host_array, reps=5):
        """Host-to-device copy GB/s of `host_array` (pinned: pinned_empty) on
        the stream streamed tables are copied on."""
        g = C.c_double()
        abi.check(lib().gdb_h2d_probe(self._ctx, host_array.ctypes.data_as(C.c_void_p),
                                      host_array.nbytes, reps, C.byref(g)))
        return g.value

    def table(self, columns):
        return Table(self, columns)

    # ---- pipelines ---------------------------------------------------------
    def q6(self, lineitem, date_lo=731, date_hi=1096, disc_lo=5, disc_hi=7, qty_lt=24,
           timing=None):
        p = Q6Params(date_lo, date_hi, disc_lo, disc_hi, qty_lt)
        r = Q6Result()
        abi.check(lib().gdb_q6(self._ctx, lineitem.handle, C.byref(p), C.byref(r),
                               abi.tref(timing)))
        return dict(revenue=r.revenue, rows=r.rows)

    def q6_noalloc(self, lineitem, date_lo=731, date_hi=1096, disc_lo=5, disc_hi=7, qty_lt=24,
                   timing=None):
        """gdb_q6_noalloc: Q6 with no allocation in the call (fixed device scratch)."""
        p = Q6Params(date_lo, date_hi, disc_lo, disc_hi, qty_lt)
        r = Q6Result()
        abi.check(lib().gdb_q6_noalloc(self._ctx, lineitem.handle, C.byref(p), C.byref(r),
                                       abi.tref(timing)))
        return dict(revenue=r.revenue, rows=r.rows)

    def q1(self, lineitem, cutoff=2435, timing=None):
        p = Q1Params(cutoff)
        r = Q1Result()
        abi.check(lib().gdb_q1(self._ctx, lineitem.handle, C.byref(p), C.byref(r),
                               abi.tref(timing)))
        return [abi.q1_row_dict(r.rows[i]) for i in range(r.nrows)]

    def q3(self, customer, orders, lineitem, segment=1, date=1169, top_k=10, timing=None):
        p = Q3Params(segment, date, top_k)
        r = Q3Result()
        abi.check(lib().gdb_q3(self._ctx, customer.handle, orders.handle, lineitem.handle,
                               C.byref(p), C.byref(r), abi.tref(timing)))
        rows = [dict(orderkey=x.orderkey, revenue=x.revenue, orderdate=x.orderdate,
                     shippriority=x.shippriority) for x in r.rows[:r.nrows]]
        return dict(groups=r.groups, rows=rows)

    def q9(self, part, supplier, partsupp, orders, lineitem, pattern="green", timing=None):
        p = Q9Params(pattern.encode())
        r = Q9Result()
        abi.check(lib().gdb_q9(self._ctx, part.handle, supplier.handle, partsupp.handle,
                               orders.handle, lineitem.handle, C.byref(p), C.byref(r),
                               abi.tref(timing)))
        nk, yr, pr = _columns(r, _Q9_DT, ("nationkey", "year", "profit"))
        return [{"nationkey": a, "year": b, "profit": c} for a, b, c in zip(nk, yr, pr)]

    def q13(self, customer, orders, word1="special", word2="requests", timing=None):
        p = Q13Params(word1.encode(), word2.encode())
        r = Q13Result()
        abi.check(lib().gdb_q13(self._ctx, customer.handle, orders.handle, C.byref(p),
                                C.byref(r), abi.tref(timing)))
        cc, cd = _columns(r, _Q13_DT, ("c_count", "custdist"))
        return [{"c_count": a, "custdist": b} for a, b in zip(cc, cd)]

    # ---- primitives --------------------------------------------------------
    def select_lt(self, table, column, value, bitmap=True, rowids=True, timing=None):
        n = table.nrows
        bm = np.zeros((n + 63) // 64, dtype=np.uint64) if bitmap else None
        ids = np.zeros(max(n, 1), dtype=np.uint32) if rowids else None
        cnt = C.c_uint64()
        abi.check(lib().gdb_select_lt(
            self._ctx, table.handle, column.encode(), value,
            bm.ctypes.data_as(C.POINTER(C.c_uint64)) if bm is not None else None,
            ids.ctypes.data_as(C.POINTER(C.c_uint32)) if ids is not None else None,
            C.byref(cnt), abi.tref(timing)))
        return dict(count=cnt.value, bitmap=bm,
                    rowids=ids[:cnt.value] if ids is not None else None)

    def sum(self, table, column, timing=None):
        s = C.c_int64()
        abi.check(lib().gdb_sum(self._ctx, table.handle, column.encode(), C.byref(s),
                                abi.tref(timing)))
        return s.value

    def join_count(self, build, build_col, probe, probe_col, timing=None):
        m = C.c_uint64()
        abi.check(lib().gdb_join_count(self._ctx, build.handle, build_col.encode(), probe.handle,
                                       probe_col.encode(), C.byref(m), abi.tref(timing)))
        return m.value


def stats_dict(st):
    return {k: getattr(st, k) for k, _ in abi.ServeStats._fields_}


def response_dict(r):
    out = dict(id=r.id, kind=("q6", "q1", "q3")[r.kind], status=r.status,
               latency_ms=r.latency_ms, batch_size=r.batch_size)
    if r.kind == abi.REQ_Q6:
        out["result"] = dict(revenue=r.q6.revenue, rows=r.q6.rows)
    elif r.kind == abi.REQ_Q1:
        out["result"] = [abi.q1_row_dict(r.q1.rows[i]) for i in range(r.q1.nrows)]
    else:
        out["result"] = dict(groups=r.q3.groups, rows=[
            dict(orderkey=x.orderkey, revenue=x.revenue, orderdate=x.orderdate,
                 shippriority=x.shippriority) for x in r.q3.rows[:r.q3.nrows]])
    return out


def request(kind, **p):
    """A gdb_request: request('q6', date_lo=.., ...), request('q1', cutoff=..),
    request('q3', segment=.., date=.., top_k=..)."""
    r = abi.Request()
    if kind == "q6":
        r.kind = abi.REQ_Q6
        r.q6 = Q6Params(p["date_lo"], p["date_hi"], p["disc_lo"], p["disc_hi"], p["qty_lt"])
    elif kind == "q1":
        r.kind = abi.REQ_Q1
        r.q1 = Q1Params(p["cutoff"])
    else:
        r.kind = abi.REQ_Q3
        r.q3 = Q3Params(p["segment"], p["date"], p.get("top_k", 10))
    return r


def loadgen_request(seed, i, q3_share=0.2):
    """The i-th request of the built-in load generator for `seed`, as a dict."""
    r = abi.Request()
    lib().gdb_loadgen_request(seed, i, q3_share, C.byref(r))
    if r.kind == abi.REQ_Q6:
        return "q6", dict(date_lo=r.q6.date_lo, date_hi=r.q6.date_hi, disc_lo=r.q6.disc_lo,
                          disc_hi=r.q6.disc_hi, qty_lt=r.q6.qty_lt)
    if r.kind == abi.REQ_Q1:
        return "q1", dict(cutoff=r.q1.cutoff)
    return "q3", dict(segment=r.q3.segment, date=r.q3.date, top_k=r.q3.top_k)


class Server:
    """A serving loop on an engine (gdb_serve_start ... gdb_serve_stop)."""

    def __init__(self, engine, lineitem_flat=None, customer=None, orders=None, lineitem=None,
                 mode="batched", max_batch=64, slots=2, hbm_budget=0, graphs=True):
        cfg = abi.ServeCfg(
            lineitem_flat.handle.value if lineitem_flat else None,
            customer.handle.value if customer else None,
            orders.handle.value if orders else None,
            lineitem.handle.value if lineitem else None,
            abi.SERVE_BATCHED if mode == "batched" else abi.SERVE_SERIAL,
            max_batch, slots, 1 if graphs else 0, hbm_budget)
        self._s = C.c_void_p()
        abi.check(lib().gdb_serve_start(engine._ctx, C.byref(cfg), C.byref(self._s)))

    def submit(self, req):
        i = C.c_uint64()
        abi.check(lib().gdb_submit(self._s, C.byref(req), C.byref(i)))
        return i.value

    def poll(self, max_n=256, timeout_ms=100):
        buf = (abi.Response * max_n)()
        n = lib().gdb_poll(self._s, buf, max_n, timeout_ms)
        if n < 0:
            abi.check(n)
        return [response_dict(buf[i]) for i in range(n)]

    def stats(self):
        st = abi.ServeStats()
        abi.check(lib().gdb_serve_stats_get(self._s, C.byref(st)))
        return stats_dict(st)

    def loadgen(self, rate, seconds, seed=1, q3_share=0.2):
        st = abi.ServeStats()
        abi.check(lib().gdb_serve_loadgen(self._s, rate, seconds, seed, q3_share, C.byref(st)))
        return stats_dict(st)

    def stop(self):
        if self._s:
            rc = lib().gdb_serve_stop(self._s)
            self._s = C.c_void_p()
            abi.check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


from . import generate  # noqa: E402  (after Engine: generate imports abi only)
