"""ctypes declarations of include/gdb.h (structs, status codes, prototypes)."""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgdb.so")

GDB_OK, GDB_E_INVALID, GDB_E_NOMEM, GDB_E_HIP, GDB_E_ADMISSION = 0, -1, -2, -3, -4
GDB_E_UNSUPPORTED, GDB_E_NODEVICE, GDB_E_TIMEOUT, GDB_E_CANCELLED, GDB_E_NOTFOUND = (
    -5, -6, -7, -8, -9)
GDB_I32, GDB_I64, GDB_U8, GDB_BYTES = 1, 2, 3, 4


class GdbError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"gdb error {code}: {msg}")
        self.code = code


class ColDesc(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int), ("width", C.c_int),
                ("host_data", C.c_void_p)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("cu_count", C.c_int),
                ("xcd_count", C.c_int), ("hbm_bytes", C.c_uint64), ("l2_bytes", C.c_uint64),
                ("lds_bytes_per_cu", C.c_int), ("wave_size", C.c_int),
                ("hbm_peak_gbps", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("gpu_ms", C.c_double), ("main_kernel_ms", C.c_double), ("post_ms", C.c_double),
                ("bytes", C.c_uint64), ("main_kernel_bytes", C.c_uint64),
                ("h2d_bytes", C.c_uint64), ("chunks", C.c_int32), ("pad", C.c_int32)]


class Q6Params(C.Structure):
    _fields_ = [("date_lo", C.c_int32), ("date_hi", C.c_int32), ("disc_lo", C.c_int32),
                ("disc_hi", C.c_int32), ("qty_lt", C.c_int32)]


class Q6Result(C.Structure):
    _fields_ = [("revenue", C.c_int64), ("rows", C.c_uint64)]


class Q1Params(C.Structure):
    _fields_ = [("cutoff", C.c_int32)]


class Q1Row(C.Structure):
    _fields_ = [("returnflag", C.c_char), ("linestatus", C.c_char), ("sum_qty", C.c_int64),
                ("sum_base_price", C.c_int64), ("sum_disc_price", C.c_int64),
                ("sum_charge_lo", C.c_uint64), ("sum_charge_hi", C.c_int64),
                ("sum_disc", C.c_int64), ("count", C.c_int64), ("avg_qty", C.c_double),
                ("avg_price", C.c_double), ("avg_disc", C.c_double)]


class Q1Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("rows", Q1Row * 6)]


class Q3Params(C.Structure):
    _fields_ = [("segment", C.c_int32), ("date", C.c_int32), ("top_k", C.c_int32)]


class Q3Row(C.Structure):
    _fields_ = [("orderkey", C.c_int32), ("orderdate", C.c_int32), ("shippriority", C.c_int32),
                ("pad", C.c_int32), ("revenue", C.c_int64)]


class Q3Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("groups", C.c_int64), ("rows", Q3Row * 100)]


class Q9Params(C.Structure):
    _fields_ = [("pattern", C.c_char * 32)]


class Q9Row(C.Structure):
    _fields_ = [("nationkey", C.c_int32), ("year", C.c_int32), ("profit", C.c_int64)]


class Q9Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("rows", Q9Row * 400)]


class Q13Params(C.Structure):
    _fields_ = [("word1", C.c_char * 32), ("word2", C.c_char * 32)]


class Q13Row(C.Structure):
    _fields_ = [("c_count", C.c_int32), ("pad", C.c_int32), ("custdist", C.c_int64)]


class Q13Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("rows", Q13Row * 512)]


class Tuning(C.Structure):
    _fields_ = [("pipeline", C.c_int32), ("block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("unroll", C.c_int32), ("ms", C.c_double), ("candidates", C.c_int32),
                ("radix_bits", C.c_int32)]


class TblField(C.Structure):
    _fields_ = [("field", C.c_int), ("kind", C.c_int), ("width", C.c_int), ("out", C.c_void_p)]


TBL_INT32, TBL_DATE, TBL_CENTS, TBL_HUNDREDTHS, TBL_CHAR, TBL_TEXT = 1, 2, 3, 4, 5, 6


class ServeCfg(C.Structure):
    _fields_ = [("lineitem_flat", C.c_void_p), ("customer", C.c_void_p), ("orders", C.c_void_p),
                ("lineitem", C.c_void_p), ("mode", C.c_int32), ("max_batch", C.c_int32),
                ("slots", C.c_int32), ("graphs", C.c_int32), ("hbm_budget", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("q6", Q6Params), ("q1", Q1Params),
                ("q3", Q3Params)]


class Response(C.Structure):
    _fields_ = [("id", C.c_uint64), ("kind", C.c_int32), ("status", C.c_int32),
                ("latency_ms", C.c_double), ("batch_size", C.c_int32), ("pad", C.c_int32),
                ("q6", Q6Result), ("q1", Q1Result), ("q3", Q3Result)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("batches", C.c_uint64), ("qps", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("mean_batch", C.c_double), ("inflight_peak_bytes", C.c_uint64),
                ("graph_captures", C.c_uint64), ("graph_replays", C.c_uint64),
                ("index_bytes", C.c_uint64), ("index_build_ms", C.c_double)]


REQ_Q6, REQ_Q1, REQ_Q3 = 0, 1, 2
SERVE_BATCHED, SERVE_SERIAL = 0, 1

PIPE_IDS = {"q6": 0, "q1": 1, "q3": 4, "q9": 5, "q13": 6, "join": 7}


def q1_row_dict(r):
    charge = (int(r.sum_charge_hi) << 64) | int(r.sum_charge_lo)
    return dict(returnflag=r.returnflag.decode(), linestatus=r.linestatus.decode(),
                sum_qty=r.sum_qty, sum_base_price=r.sum_base_price,
                sum_disc_price=r.sum_disc_price, sum_charge=charge, sum_disc=r.sum_disc,
                count=r.count, avg_qty=r.avg_qty, avg_price=r.avg_price, avg_disc=r.avg_disc)


P = C.c_void_p
I32P, I64P, U8P = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint8)
U32P, U64P = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)

# name -> (restype, argtypes); every symbol include/gdb.h declares.
PROTOTYPES = {
    "gdb_strerror": (C.c_char_p, [C.c_int]),
    "gdb_last_error": (C.c_char_p, []),
    "gdb_open": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gdb_close": (None, [P]),
    "gdb_device_info_get": (C.c_int, [P, C.POINTER(DeviceInfo)]),
    "gdb_hbm_used": (C.c_uint64, [P]),
    "gdb_debug_fail_allocs": (None, [P, C.c_int]),
    "gdb_debug_fail_allocs_after": (None, [P, C.c_int, C.c_int]),
    "gdb_release_scratch": (C.c_int, [P]),
    "gdb_bandwidth_probe": (C.c_int, [P, C.c_uint64, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "gdb_h2d_probe": (C.c_int, [P, P, C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "gdb_table_create": (C.c_int, [P, C.POINTER(ColDesc), C.c_int, C.c_uint64,
                                   C.POINTER(C.c_void_p)]),
    "gdb_table_free": (None, [P]),
    "gdb_table_rows": (C.c_uint64, [P]),
    "gdb_table_bytes": (C.c_uint64, [P]),
    "gdb_host_alloc": (C.c_int, [P, C.c_uint64, C.POINTER(C.c_void_p)]),
    "gdb_host_free": (None, [P, P]),
    "gdb_table_create_host": (C.c_int, [P, C.POINTER(ColDesc), C.c_int, C.c_uint64,
                                        C.POINTER(C.c_void_p)]),
    "gdb_table_is_host": (C.c_int, [P]),
    "gdb_set_chunk_rows": (C.c_int, [P, C.c_uint64]),
    "gdb_q6": (C.c_int, [P, P, C.POINTER(Q6Params), C.POINTER(Q6Result), C.POINTER(Timing)]),
    "gdb_q6_noalloc": (C.c_int, [P, P, C.POINTER(Q6Params), C.POINTER(Q6Result),
                                 C.POINTER(Timing)]),
    "gdb_q1": (C.c_int, [P, P, C.POINTER(Q1Params), C.POINTER(Q1Result), C.POINTER(Timing)]),
    "gdb_q3": (C.c_int, [P, P, P, P, C.POINTER(Q3Params), C.POINTER(Q3Result),
                         C.POINTER(Timing)]),
    "gdb_q9": (C.c_int, [P, P, P, P, P, P, C.POINTER(Q9Params), C.POINTER(Q9Result),
                         C.POINTER(Timing)]),
    "gdb_q13": (C.c_int, [P, P, P, C.POINTER(Q13Params), C.POINTER(Q13Result),
                          C.POINTER(Timing)]),
    "gdb_select_lt": (C.c_int, [P, P, C.c_char_p, C.c_int32, U64P, U32P, U64P,
                                C.POINTER(Timing)]),
    "gdb_sum": (C.c_int, [P, P, C.c_char_p, I64P, C.POINTER(Timing)]),
    "gdb_join_count": (C.c_int, [P, P, C.c_char_p, P, C.c_char_p, U64P, C.POINTER(Timing)]),
    "gdb_autotune": (C.c_int, [P, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.c_int,
                               C.POINTER(Tuning)]),
    "gdb_get_tuning": (C.c_int, [P, C.c_int, C.POINTER(Tuning)]),
    "gdb_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "gdb_set_kernel_timing": (C.c_int, [P, C.c_int]),
    "gdb_serve_start": (C.c_int, [P, C.POINTER(ServeCfg), C.POINTER(C.c_void_p)]),
    "gdb_submit": (C.c_int, [P, C.POINTER(Request), U64P]),
    "gdb_poll": (C.c_int, [P, C.POINTER(Response), C.c_int, C.c_int]),
    "gdb_serve_stats_get": (C.c_int, [P, C.POINTER(ServeStats)]),
    "gdb_serve_stop": (C.c_int, [P]),
    "gdb_serve_loadgen": (C.c_int, [P, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                    C.POINTER(ServeStats)]),
    "gdb_loadgen_request": (None, [C.c_uint64, C.c_uint64, C.c_double, C.POINTER(Request)]),
    "gdb_gen_lineitem_flat": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, I32P, I32P, I32P,
                                        I32P, I64P, U8P, U8P, I32P, C.c_int]),
    "gdb_gen_customer": (C.c_int, [C.c_uint64, C.c_uint64, I32P, U8P, C.c_int]),
    "gdb_gen_orders": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, I32P, I32P,
                                 I32P, I32P, U8P, C.c_int]),
    "gdb_gen_lineitem_rows": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
    "gdb_gen_lineitem_joined": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                          C.c_uint64, I32P, I32P, I32P, I32P, I32P, I64P, I32P,
                                          C.c_int]),
    "gdb_gen_part": (C.c_int, [C.c_uint64, C.c_uint64, I32P, U8P, C.c_int]),
    "gdb_gen_partsupp": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, I32P, I32P, I64P,
                                   C.c_int]),
    "gdb_gen_supplier": (C.c_int, [C.c_uint64, C.c_uint64, I32P, I32P, C.c_int]),
    "gdb_nation_name": (C.c_char_p, [C.c_int]),
    "gdb_segment_name": (C.c_char_p, [C.c_int]),
    "gdb_year_of_day": (C.c_int32, [C.c_int32]),
    "gdb_tbl_count_rows": (C.c_int, [C.c_char_p, C.c_int, U64P]),
    "gdb_tbl_parse": (C.c_int, [C.c_char_p, C.POINTER(TblField), C.c_int, C.c_uint64, C.c_int]),
}

_lib = None


def load(path=None):
    """Loads libgdb.so once and declares every prototype. Raises GdbError when
    the library has not been built: there is no fallback implementation."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise GdbError(GDB_E_UNSUPPORTED,
                       f"{path} is missing: build it with python -m gpudbmetalbenchmark_amd.build")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:  # reported by the export test; calling it raises AttributeError
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != GDB_OK:
        lib = load()
        raise GdbError(rc, f"{lib.gdb_strerror(rc).decode()}: {lib.gdb_last_error().decode()}")
    return rc


def tref(t):
    return C.byref(t) if t is not None else None


# ---- include/gdb_tpch.h ------------------------------------------------------
class TpchDb(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("lineitem", "orders", "customer", "part", "partsupp",
                                          "supplier", "nation")]


M3 = C.c_uint64 * 3


class Q2Params(C.Structure):
    _fields_ = [("size", C.c_int32), ("region", C.c_int32), ("top_k", C.c_int32),
                ("pad", C.c_int32), ("type_mask", M3)]


class Q2Row(C.Structure):
    _fields_ = [("s_acctbal", C.c_int64), ("suppkey", C.c_int32), ("nationkey", C.c_int32),
                ("partkey", C.c_int32), ("mfgr", C.c_int32)]


class Q2Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("pad", C.c_int32), ("matches", C.c_int64),
                ("rows", Q2Row * 100)]


class Q4Params(C.Structure):
    _fields_ = [("date_lo", C.c_int32), ("date_hi", C.c_int32)]


class Q4Result(C.Structure):
    _fields_ = [("counts", C.c_int64 * 5)]


class Q5Params(C.Structure):
    _fields_ = [("region", C.c_int32), ("date_lo", C.c_int32), ("date_hi", C.c_int32)]


class Q5Result(C.Structure):
    _fields_ = [("revenue", C.c_int64 * 25), ("has", C.c_uint32)]


class Q7Params(C.Structure):
    _fields_ = [("nation1", C.c_int32), ("nation2", C.c_int32), ("date_lo", C.c_int32),
                ("date_hi", C.c_int32)]


class Q7Result(C.Structure):
    _fields_ = [("revenue", (C.c_int64 * 8) * 2), ("has", C.c_uint32 * 2)]


class Q8Params(C.Structure):
    _fields_ = [("nation", C.c_int32), ("region", C.c_int32), ("type_code", C.c_int32),
                ("date_lo", C.c_int32), ("date_hi", C.c_int32)]


class Q8Result(C.Structure):
    _fields_ = [("nation_volume", C.c_int64 * 8), ("total_volume", C.c_int64 * 8),
                ("has", C.c_uint32)]


class Q10Params(C.Structure):
    _fields_ = [("date_lo", C.c_int32), ("date_hi", C.c_int32), ("top_k", C.c_int32),
                ("returnflag", C.c_char), ("pad", C.c_char * 3)]


class Q10Row(C.Structure):
    _fields_ = [("custkey", C.c_int32), ("nationkey", C.c_int32), ("revenue", C.c_int64),
                ("acctbal", C.c_int64)]


class Q10Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("pad", C.c_int32), ("groups", C.c_int64),
                ("rows", Q10Row * 100)]


class Q11Params(C.Structure):
    _fields_ = [("nation", C.c_int32), ("pad", C.c_int32), ("frac_num", C.c_int64),
                ("frac_den", C.c_int64)]


class Q11Row(C.Structure):
    _fields_ = [("partkey", C.c_int32), ("pad", C.c_int32), ("value", C.c_int64)]


class Q12Params(C.Structure):
    _fields_ = [("mode", C.c_int32 * 2), ("date_lo", C.c_int32), ("date_hi", C.c_int32),
                ("high_mask", C.c_uint32)]


class Q12Result(C.Structure):
    _fields_ = [("high", C.c_int64 * 2), ("low", C.c_int64 * 2)]


class Q14Params(C.Structure):
    _fields_ = [("date_lo", C.c_int32), ("date_hi", C.c_int32), ("type_mask", M3)]


class Q14Result(C.Structure):
    _fields_ = [("promo", C.c_int64), ("total", C.c_int64), ("promo_revenue", C.c_double)]


class Q15Params(C.Structure):
    _fields_ = [("date_lo", C.c_int32), ("date_hi", C.c_int32)]


class Q15Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("pad", C.c_int32), ("max_revenue", C.c_int64),
                ("suppkey", C.c_int32 * 100)]


class Q16Params(C.Structure):
    _fields_ = [("brand", C.c_int32), ("nsizes", C.c_int32), ("sizes", C.c_int32 * 8),
                ("type_mask", M3), ("word1", C.c_char * 32), ("word2", C.c_char * 32)]


class Q16Row(C.Structure):
    _fields_ = [("brand", C.c_int32), ("type", C.c_int32), ("size", C.c_int32),
                ("supplier_cnt", C.c_int32)]


class Q17Params(C.Structure):
    _fields_ = [("brand", C.c_int32), ("container", C.c_int32), ("frac_num", C.c_int32),
                ("frac_den", C.c_int32)]


class Q17Result(C.Structure):
    _fields_ = [("sum_price", C.c_int64), ("avg_yearly", C.c_double)]


class Q18Params(C.Structure):
    _fields_ = [("quantity", C.c_int32), ("top_k", C.c_int32)]


class Q18Row(C.Structure):
    _fields_ = [("custkey", C.c_int32), ("orderkey", C.c_int32), ("orderdate", C.c_int32),
                ("sum_qty", C.c_int32), ("totalprice", C.c_int64)]


class Q18Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("pad", C.c_int32), ("orders_over", C.c_int64),
                ("rows", Q18Row * 100)]


class Q19Block(C.Structure):
    _fields_ = [("brand", C.c_int32), ("qty_lo", C.c_int32), ("qty_hi", C.c_int32),
                ("size_lo", C.c_int32), ("size_hi", C.c_int32), ("pad", C.c_int32),
                ("container_mask", C.c_uint64)]


class Q19Params(C.Structure):
    _fields_ = [("nblocks", C.c_int32), ("mode_mask", C.c_uint32), ("instruct", C.c_int32),
                ("pad", C.c_int32), ("blocks", Q19Block * 3)]


class Q19Result(C.Structure):
    _fields_ = [("revenue", C.c_int64)]


class Q20Params(C.Structure):
    _fields_ = [("prefix", C.c_char * 32), ("date_lo", C.c_int32), ("date_hi", C.c_int32),
                ("nation", C.c_int32), ("frac_num", C.c_int32), ("frac_den", C.c_int32),
                ("pad", C.c_int32)]


class Q21Params(C.Structure):
    _fields_ = [("nation", C.c_int32), ("top_k", C.c_int32), ("status", C.c_char),
                ("pad", C.c_char * 7)]


class Q21Row(C.Structure):
    _fields_ = [("suppkey", C.c_int32), ("pad", C.c_int32), ("numwait", C.c_int64)]


class Q21Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("pad", C.c_int32), ("suppliers", C.c_int64),
                ("rows", Q21Row * 100)]


class Q22Params(C.Structure):
    _fields_ = [("ncodes", C.c_int32), ("codes", C.c_int32 * 7)]


class Q22Row(C.Structure):
    _fields_ = [("cntrycode", C.c_int32), ("pad", C.c_int32), ("numcust", C.c_int64),
                ("totacctbal", C.c_int64)]


class Q22Result(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("pad", C.c_int32), ("rows", Q22Row * 7)]


_DB = C.POINTER(TpchDb)
_T = C.POINTER(Timing)
PROTOTYPES.update({
    "gdb_q2": (C.c_int, [P, _DB, C.POINTER(Q2Params), C.POINTER(Q2Result), _T]),
    "gdb_q4": (C.c_int, [P, _DB, C.POINTER(Q4Params), C.POINTER(Q4Result), _T]),
    "gdb_q5": (C.c_int, [P, _DB, C.POINTER(Q5Params), C.POINTER(Q5Result), _T]),
    "gdb_q7": (C.c_int, [P, _DB, C.POINTER(Q7Params), C.POINTER(Q7Result), _T]),
    "gdb_q8": (C.c_int, [P, _DB, C.POINTER(Q8Params), C.POINTER(Q8Result), _T]),
    "gdb_q10": (C.c_int, [P, _DB, C.POINTER(Q10Params), C.POINTER(Q10Result), _T]),
    "gdb_q11": (C.c_int, [P, _DB, C.POINTER(Q11Params), C.POINTER(Q11Row), C.c_int64, I64P,
                          I64P, _T]),
    "gdb_q12": (C.c_int, [P, _DB, C.POINTER(Q12Params), C.POINTER(Q12Result), _T]),
    "gdb_q14": (C.c_int, [P, _DB, C.POINTER(Q14Params), C.POINTER(Q14Result), _T]),
    "gdb_q15": (C.c_int, [P, _DB, C.POINTER(Q15Params), C.POINTER(Q15Result), _T]),
    "gdb_q16": (C.c_int, [P, _DB, C.POINTER(Q16Params), C.POINTER(Q16Row), C.c_int64, I64P,
                          _T]),
    "gdb_q17": (C.c_int, [P, _DB, C.POINTER(Q17Params), C.POINTER(Q17Result), _T]),
    "gdb_q18": (C.c_int, [P, _DB, C.POINTER(Q18Params), C.POINTER(Q18Result), _T]),
    "gdb_q19": (C.c_int, [P, _DB, C.POINTER(Q19Params), C.POINTER(Q19Result), _T]),
    "gdb_q20": (C.c_int, [P, _DB, C.POINTER(Q20Params), I32P, C.c_int64, I64P, _T]),
    "gdb_q21": (C.c_int, [P, _DB, C.POINTER(Q21Params), C.POINTER(Q21Result), _T]),
    "gdb_q22": (C.c_int, [P, _DB, C.POINTER(Q22Params), C.POINTER(Q22Result), _T]),
})
