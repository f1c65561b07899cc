/*
 * gdb.h — C ABI of the MI355X (gfx950) query engine, libgdb.so.
 *
 * The reference drives each pipeline through a Metal buffer/pipeline layer:
 * runQ*Benchmark(MTL::Device*, MTL::CommandQueue*, MTL::Library*)
 * (reference src/infra.h:874-922), createPipeline (src/infra.h:178-186),
 * device->newBuffer(..., StorageModeShared) and commit/waitUntilCompleted
 * (e.g. src/q6.cpp:42-47, :89-90). This header replaces that layer: columnar
 * tables are copied once into HBM and stay resident, every pipeline is one
 * synchronous call on a resident table, and timing comes from hipEvents.
 *
 * Conventions
 *  - Every call returns GDB_OK (0) or a negative GDB_E_* code; no C++ exception
 *    crosses the ABI. gdb_last_error() gives a message for the calling thread.
 *  - Integer and money columns are int32 / int64 fixed point (cents, hundredths),
 *    so every sum and count is exact.
 *  - A context is used by one host thread at a time; the serving calls
 *    (gdb_serve_*) are the exception and are thread-safe.
 *  - Day numbers count from 1992-01-01 = day 0.
 */
#ifndef GDB_H
#define GDB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  GDB_OK = 0,
  GDB_E_INVALID = -1,     /* bad argument, unknown column, malformed parameters */
  GDB_E_NOMEM = -2,       /* device or pinned-host allocation failed */
  GDB_E_HIP = -3,         /* a HIP runtime call failed */
  GDB_E_ADMISSION = -4,   /* request footprint exceeds the HBM budget */
  GDB_E_UNSUPPORTED = -5, /* shape the engine does not handle */
  GDB_E_NODEVICE = -6,    /* no GPU visible */
  GDB_E_TIMEOUT = -7,     /* serving watchdog fired */
  GDB_E_CANCELLED = -8,   /* request cancelled before it ran */
  GDB_E_NOTFOUND = -9     /* unknown pipeline name / request id */
};

/* Replaces the stderr prints of createPipeline (reference src/infra.h:181-184). */
const char* gdb_strerror(int code);
const char* gdb_last_error(void);

/* ---- context (replaces MTL::CreateSystemDefaultDevice + newCommandQueue,
 *      reference src/main.cpp:76-89) ------------------------------------- */
typedef struct gdb_ctx gdb_ctx;
typedef struct gdb_table gdb_table;

typedef struct {
  char name[64];
  char arch[32];
  int cu_count;
  int xcd_count;           /* 8 on MI355X */
  uint64_t hbm_bytes;
  uint64_t l2_bytes;
  int lds_bytes_per_cu;
  int wave_size;
  double hbm_peak_gbps;    /* published peak used for roofline fractions */
} gdb_device_info;

/* hbm_budget_bytes = 0 means "all free HBM minus 4 GiB". */
int gdb_open(int device, uint64_t hbm_budget_bytes, gdb_ctx** out);
void gdb_close(gdb_ctx* ctx);
int gdb_device_info_get(gdb_ctx* ctx, gdb_device_info* out);
/* Bytes of HBM currently held by resident tables and arenas. */
uint64_t gdb_hbm_used(gdb_ctx* ctx);
/* Achievable HBM bandwidth: a plain streaming read and a plain copy over
 * `bytes` (rounded down to 4 KiB), averaged over `reps` runs after one warm-up.
 * copy_gbps counts read + write bytes. No counterpart in the reference: it is
 * the calibration kernel for counter-based traffic measurement. */
int gdb_bandwidth_probe(gdb_ctx* ctx, uint64_t bytes, int reps, double* read_gbps,
                        double* copy_gbps);
/* Host-to-device copy rate (GB/s) of `bytes` from `host` (pinned memory from
 * gdb_host_alloc gives the full PCIe rate) on the stream that streamed tables
 * are copied on; `reps` timed copies after one warm-up. The ceiling of a
 * streamed pipeline (the reference's SF100 paths, src/infra.h:816-869). */
int gdb_h2d_probe(gdb_ctx* ctx, const void* host, uint64_t bytes, int reps, double* gbps);
/* Gives the context's scratch arenas (partials, maps, hash tables, stream
 * slots; they grow on demand and are otherwise kept between calls) back to
 * the device after the work queued on the context has finished. Tables stay.
 * No counterpart in the reference, whose Metal buffers live only as long as
 * each query function. Like every call on a context, not concurrent with
 * another call on it (a server running on the context included). */
int gdb_release_scratch(gdb_ctx* ctx);
/* Test hook: the next `count` device allocations fail with GDB_E_NOMEM. */
void gdb_debug_fail_allocs(gdb_ctx* ctx, int count);
/* Test hook: the next `skip` device allocations succeed, then `count` fail. */
void gdb_debug_fail_allocs_after(gdb_ctx* ctx, int skip, int count);

/* ---- resident columnar tables (replaces newBuffer(ptr, bytes, Shared),
 *      reference src/q6.cpp:42-47, and the per-query column vectors) -------- */
enum { GDB_I32 = 1, GDB_I64 = 2, GDB_U8 = 3, GDB_BYTES = 4 /* fixed width */ };

typedef struct {
  const char* name;      /* e.g. "l_shipdate" (TPC-H column names) */
  int type;              /* GDB_I32 / GDB_I64 / GDB_U8 / GDB_BYTES */
  int width;             /* bytes per value for GDB_BYTES (55, 80), else ignored */
  const void* host_data; /* nrows values; copied to HBM by gdb_table_create */
} gdb_col_desc;

int gdb_table_create(gdb_ctx* ctx, const gdb_col_desc* cols, int ncols, uint64_t nrows,
                     gdb_table** out);
void gdb_table_free(gdb_table* t);
uint64_t gdb_table_rows(const gdb_table* t);
uint64_t gdb_table_bytes(const gdb_table* t);

/* ---- timing (replaces GPUEndTime()-GPUStartTime(), reference
 *      src/q6.cpp:92-94, and printTimingSummary, src/infra.h:511-516) ------- */
typedef struct {
  double gpu_ms;          /* hipEvent time of the whole pipeline on its stream */
  double main_kernel_ms;  /* hipEvent time of the dominant kernel alone */
  double post_ms;         /* host finalisation (sort, format) */
  uint64_t bytes;         /* algorithmic bytes read by the pipeline's scans */
  uint64_t main_kernel_bytes;
  uint64_t h2d_bytes;     /* bytes copied host->device for a streamed table */
  int32_t chunks;         /* chunks a streamed table was cut into (0: resident) */
  int32_t pad;
} gdb_timing;

/* Kernel timing (on by default): while on, every pipeline brackets its
 * dominant kernel with timing events and fills main_kernel_ms; off, it
 * records no event between its kernels (each one leaves the GPU idle for a
 * few microseconds) and sets main_kernel_ms to -1. No counterpart in the
 * reference, which times whole command buffers (src/q6.cpp:92-94). */
int gdb_set_kernel_timing(gdb_ctx* ctx, int on);

/* ---- host-resident tables streamed in chunks (replaces chunkedStreamLoop
 *      and ChunkConfig, reference src/infra.h:37-57, :805-869, used by the
 *      runQ*BenchmarkSF100 entry points, src/infra.h:874-922) ----------------
 * A host table references caller memory (no copy; keep it alive until
 * gdb_table_free). Passed as the fact table of gdb_q6/q1/q3/q9 or the orders
 * table of gdb_q13, it is copied in chunks through two device slots per column
 * on a copy stream while the previous chunk is processed; every other input
 * must be resident. Memory from gdb_host_alloc is pinned, so copies run at
 * full PCIe rate; pageable memory works but copies more slowly. */
int gdb_host_alloc(gdb_ctx* ctx, uint64_t bytes, void** out);
void gdb_host_free(gdb_ctx* ctx, void* p);
int gdb_table_create_host(gdb_ctx* ctx, const gdb_col_desc* cols, int ncols, uint64_t nrows,
                          gdb_table** out);
int gdb_table_is_host(const gdb_table* t);
/* Rows per streamed chunk; 0 = adaptive (1/8 of the HBM budget over two
 * slots, clamped to 1M..64M rows; the reference's ChunkConfig rule). */
int gdb_set_chunk_rows(gdb_ctx* ctx, uint64_t rows);

/* ---- pipelines (drop-in for runQ6/Q1/Q3/Q9/Q13Benchmark,
 *      reference src/infra.h:874-922) --------------------------------------- */

/* Q6 filter-scan-sum (reference src/q6.cpp:8-120, kernels/q6.metal:12-73,
 * sql/q6.sql). Table columns: l_shipdate, l_discount, l_quantity (int32),
 * l_extendedprice (int64 cents). Predicate: date_lo <= shipdate < date_hi,
 * disc_lo <= discount <= disc_hi, quantity < qty_lt.
 * revenue = SUM(extprice * discount) in cents*hundredths (1e-4 dollars). */
typedef struct { int32_t date_lo, date_hi, disc_lo, disc_hi, qty_lt; } gdb_q6_params;
typedef struct { int64_t revenue; uint64_t rows; } gdb_q6_result;
int gdb_q6(gdb_ctx* ctx, const gdb_table* lineitem, const gdb_q6_params* p, gdb_q6_result* out,
           gdb_timing* timing);

/* Q6 without allocation: no counterpart in the reference. For callers that
 * cannot allocate device or pinned memory during a query (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same kernels and result as gdb_q6 on a resident table, with their scratch
 * (per-workgroup partials and totals) in a fixed 1 MiB array in device memory
 * that each GPU holds until the process ends; no hipMalloc, hipFree or pinned
 * allocation happens in the call. A launch shape needing more than 1 MiB of
 * scratch fails with GDB_E_UNSUPPORTED, as does a host (streamed) table.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from several threads or contexts) race; serialise them. */
int gdb_q6_noalloc(gdb_ctx* ctx, const gdb_table* lineitem, const gdb_q6_params* p,
                   gdb_q6_result* out, gdb_timing* timing);

/* Q1 grouped aggregation (reference src/q1.cpp:8-173, kernels/q1.metal:39-121,
 * sql/q1.sql). Columns: l_shipdate, l_quantity, l_discount, l_tax (int32),
 * l_extendedprice (int64), l_returnflag, l_linestatus (uint8 chars).
 * Filter shipdate <= cutoff; groups (returnflag in {A,N,R}) x (linestatus in
 * {F,O}); rows with other codes belong to no group. Output rows are the
 * non-empty groups in (returnflag, linestatus) order. Units: sum_disc_price is
 * cents*hundredths, sum_charge cents*1e-4 held as a 128-bit (hi, lo) pair. */
typedef struct { int32_t cutoff; } gdb_q1_params;
typedef struct {
  char returnflag, linestatus;
  int64_t sum_qty, sum_base_price, sum_disc_price;
  uint64_t sum_charge_lo; int64_t sum_charge_hi;
  int64_t sum_disc, count;
  double avg_qty, avg_price, avg_disc; /* exact sums / count, in SQL units */
} gdb_q1_row;
typedef struct { int32_t nrows; gdb_q1_row rows[6]; } gdb_q1_result;
int gdb_q1(gdb_ctx* ctx, const gdb_table* lineitem, const gdb_q1_params* p, gdb_q1_result* out,
           gdb_timing* timing);

/* Q3 join + group + top-k (reference src/q3.cpp:8-173, kernels/q3.metal:43-340,
 * src/infra.h:538-564, sql/q3.sql). customer: c_custkey, c_mktsegment (uint8
 * code 0..4); orders: o_orderkey, o_custkey, o_orderdate, o_shippriority;
 * lineitem: l_orderkey, l_shipdate, l_discount (int32), l_extendedprice (int64).
 * Filter segment, orderdate < date, shipdate > date; revenue =
 * SUM(extprice*(100-discount)) per orderkey; top k by revenue desc, orderdate
 * asc, orderkey asc. `groups` = number of qualifying orderkeys. */
typedef struct { int32_t segment, date, top_k; } gdb_q3_params;
typedef struct { int32_t orderkey; int32_t orderdate; int32_t shippriority; int32_t pad; int64_t revenue; } gdb_q3_row;
typedef struct { int32_t nrows; int64_t groups; gdb_q3_row rows[100]; } gdb_q3_result;
int gdb_q3(gdb_ctx* ctx, const gdb_table* customer, const gdb_table* orders,
           const gdb_table* lineitem, const gdb_q3_params* p, gdb_q3_result* out,
           gdb_timing* timing);

/* Q9 multi-way join with substring filter (reference src/q9.cpp:8-241,
 * kernels/q9.metal:44-289, src/infra.h:733-776, sql/q9.sql). part: p_partkey,
 * p_name (55-byte field); supplier: s_suppkey, s_nationkey; partsupp:
 * ps_partkey, ps_suppkey (int32), ps_supplycost (int64 cents); orders:
 * o_orderkey, o_orderdate; lineitem: l_partkey, l_suppkey, l_orderkey,
 * l_quantity, l_discount (int32), l_extendedprice (int64).
 * amount = extprice*(100-discount) - supplycost*quantity*100 (cents*hundredths),
 * grouped by (nation, year(orderdate)), ordered nation name asc, year desc. */
typedef struct { char pattern[32]; } gdb_q9_params; /* substring, e.g. "green" */
typedef struct { int32_t nationkey; int32_t year; int64_t profit; } gdb_q9_row;
typedef struct { int32_t nrows; gdb_q9_row rows[25 * 16]; } gdb_q9_result;
int gdb_q9(gdb_ctx* ctx, const gdb_table* part, const gdb_table* supplier,
           const gdb_table* partsupp, const gdb_table* orders, const gdb_table* lineitem,
           const gdb_q9_params* p, gdb_q9_result* out, gdb_timing* timing);

/* Q13 left-outer count + histogram (reference src/q13.cpp:9-116,
 * kernels/q13.metal:33-167, src/infra.h:779-799, sql/q13.sql). customer:
 * c_custkey; orders: o_custkey (int32), o_comment (80-byte field, NUL padded).
 * An order counts when its comment does NOT match %word1%word2%. Output rows
 * (c_count, custdist) ordered custdist desc, c_count desc; no bin cap. */
typedef struct { char word1[32]; char word2[32]; } gdb_q13_params;
typedef struct { int32_t c_count; int32_t pad; int64_t custdist; } gdb_q13_row;
typedef struct { int32_t nrows; gdb_q13_row rows[512]; } gdb_q13_result;
int gdb_q13(gdb_ctx* ctx, const gdb_table* customer, const gdb_table* orders,
            const gdb_q13_params* p, gdb_q13_result* out, gdb_timing* timing);

/* ---- operator primitives (drop-in for the `selection`, `aggregation`,
 *      `join` microbenchmarks, reference src/microbenchmarks.cpp:51-251,
 *      kernels/microbenchmarks.metal:158-294) --------------------------------- */

/* Predicate scan `column < value` over an int32 column. Emits a selection
 * bitmap (bit r of word r/64 = row r passes; host buffer of ceil(n/64) u64, may
 * be NULL) and, if rowids != NULL, the passing row ids in ascending order (host
 * buffer of n u32). *count = number of passing rows. */
int gdb_select_lt(gdb_ctx* ctx, const gdb_table* t, const char* column, int32_t value,
                  uint64_t* bitmap, uint32_t* rowids, uint64_t* count, gdb_timing* timing);
/* SUM of an int32 or int64 column, exact in int64. */
int gdb_sum(gdb_ctx* ctx, const gdb_table* t, const char* column, int64_t* sum,
            gdb_timing* timing);
/* Equi-join count: number of (probe row, build row) pairs with equal keys,
 * build keys unique (reference hash_join_build/probe). */
int gdb_join_count(gdb_ctx* ctx, const gdb_table* build, const char* build_col,
                   const gdb_table* probe, const char* probe_col, uint64_t* matches,
                   gdb_timing* timing);

/* ---- serving (no counterpart in the reference, which runs one query per
 *      command buffer and has no request loop) ----------------------------------
 * A server owns the context while it runs (do not call other gdb_* functions
 * on the same context meanwhile). Requests are parameterised Q6 / Q1 / Q3
 * templates over the configured tables; gdb_submit and gdb_poll are
 * thread-safe. In GDB_SERVE_BATCHED mode a scheduler thread forms batches
 * continuously: pending flat-table requests (Q6, Q1) share one fused pass over
 * the fact table, pending Q3 requests run as one launch over resident join
 * indexes; up to `slots` batches are in flight on their own streams, and a
 * batch is admitted only while the HBM footprint of the in-flight batches
 * stays within `hbm_budget`. GDB_SERVE_SERIAL runs each request alone with the
 * plain pipeline (gdb_q6 / gdb_q1 / gdb_q3), one at a time. */
enum { GDB_SERVE_BATCHED = 0, GDB_SERVE_SERIAL = 1 };
enum { GDB_REQ_Q6 = 0, GDB_REQ_Q1 = 1, GDB_REQ_Q3 = 2 };
typedef struct gdb_server gdb_server;
typedef struct {
  const gdb_table* lineitem_flat;  /* Q6 / Q1 fact table (may be a host table) */
  const gdb_table* customer;       /* Q3 tables (resident), or NULL for no Q3 */
  const gdb_table* orders;
  const gdb_table* lineitem;
  int32_t mode;                    /* GDB_SERVE_BATCHED / GDB_SERVE_SERIAL */
  int32_t max_batch;               /* requests per batch, 1..64 */
  int32_t slots;                   /* batches in flight, 1..4 */
  int32_t graphs;                  /* 1: batch kernels of resident tables are captured
                                      once per launch shape and replayed as HIP graphs */
  uint64_t hbm_budget;             /* bytes for in-flight batch scratch; 0 = no limit */
} gdb_serve_cfg;
typedef struct {
  int32_t kind;                    /* GDB_REQ_* */
  int32_t pad;
  gdb_q6_params q6;
  gdb_q1_params q1;
  gdb_q3_params q3;
} gdb_request;
typedef struct {
  uint64_t id;
  int32_t kind;
  int32_t status;                  /* GDB_OK or an error code */
  double latency_ms;               /* submit -> result available */
  int32_t batch_size;              /* requests that shared this request's pass */
  int32_t pad;
  gdb_q6_result q6;
  gdb_q1_result q1;
  gdb_q3_result q3;
} gdb_response;
typedef struct {
  uint64_t submitted, completed, rejected, batches;
  double qps;                      /* completed / (last completion - first submit) */
  double p50_ms, p99_ms, mean_ms, max_ms;
  double mean_batch;
  uint64_t inflight_peak_bytes;
  uint64_t graph_captures;         /* launch shapes captured as graphs */
  uint64_t graph_replays;          /* batches launched as graph replays */
  uint64_t index_bytes;            /* HBM held by the Q3 join indexes built at start */
  double index_build_ms;           /* wall time gdb_serve_start spent building them */
} gdb_serve_stats;
int gdb_serve_start(gdb_ctx* ctx, const gdb_serve_cfg* cfg, gdb_server** out);
/* Thread-safe; GDB_E_ADMISSION if the request alone exceeds the budget. */
int gdb_submit(gdb_server* s, const gdb_request* req, uint64_t* id);
/* Up to max completed responses; waits up to timeout_ms for the first. */
int gdb_poll(gdb_server* s, gdb_response* out, int max, int timeout_ms);
int gdb_serve_stats_get(gdb_server* s, gdb_serve_stats* out);
/* Drains in-flight work, stops the scheduler and frees the server. */
int gdb_serve_stop(gdb_server* s);
/* Open-loop load: Poisson arrivals at `rate` requests/s for `seconds`, drawn
 * 50/30/20 from Q6/Q1/Q3 with seeded parameters (window start U[0,2160],
 * discount band centre U[2,9] -> [c-1, c+1], quantity cap 24 or 25, segment
 * U{0..4}; Q1 cutoff = start + 365, Q3 date = start + 182, top 10). Waits
 * for every response; fills stats for this run. q3_share = 0 drops Q3. */
int gdb_serve_loadgen(gdb_server* s, double rate, double seconds, uint64_t seed, double q3_share,
                      gdb_serve_stats* stats);
/* The request the load generator issues as its i-th request for `seed`. */
void gdb_loadgen_request(uint64_t seed, uint64_t i, double q3_share, gdb_request* out);

/* ---- autotuning (no counterpart in the reference: it hard-codes 2048 x 1024
 *      threadgroups, e.g. src/q6.cpp:72-74) ---------------------------------
 * A tuning entry is the launch shape of a pipeline's dominant kernel:
 * threads per workgroup, workgroups per CU (0 = occupancy limit) and vector
 * groups of 4 rows in flight per lane. gdb_autotune times every candidate on
 * the given tables (median of `reps` runs after one warm-up), keeps the fastest
 * for the context and returns it. Tables are passed in the pipeline's argument
 * order (q6/q1: lineitem; q3: customer, orders, lineitem; q9: part, supplier,
 * partsupp, orders, lineitem; q13: customer, orders; join: build, probe, on
 * each table's first column). For the join the candidates are strategies
 * and radix bits (the LDS table size follows from the bits). */
enum { GDB_PIPE_Q6 = 0, GDB_PIPE_Q1 = 1, GDB_PIPE_Q3 = 4, GDB_PIPE_Q9 = 5, GDB_PIPE_Q13 = 6,
       GDB_PIPE_JOIN = 7 };
typedef struct {
  int32_t pipeline;
  int32_t block;
  int32_t blocks_per_cu;
  int32_t unroll;
  double ms;           /* median pipeline time of the chosen shape */
  int32_t candidates;  /* shapes timed */
  int32_t radix_bits;  /* join only: partition bits 1..14; 0 = derived from the
                          build side (a direct-address filter over the build
                          keys' range when it spans <= 8x the build rows, else
                          LDS broadcast when it fits one table, else radix);
                          -1 = one global hash table; -2 = the direct-address
                          filter whenever the range spans <= 64x the build
                          rows + 2^20 values */
} gdb_tuning;
int gdb_autotune(gdb_ctx* ctx, int pipeline, const gdb_table* const* tables, int ntables,
                 int reps, gdb_tuning* best);
int gdb_get_tuning(gdb_ctx* ctx, int pipeline, gdb_tuning* out);
int gdb_set_tuning(gdb_ctx* ctx, const gdb_tuning* t);

/* ---- seeded synthetic tables (replace the dbgen .tbl loaders, reference
 *      src/infra.h:231-468). Host-side; callers own the output arrays. ------ */
int gdb_gen_lineitem_flat(uint64_t seed, uint64_t row0, uint64_t n, int32_t* shipdate,
                          int32_t* quantity, int32_t* discount, int32_t* tax, int64_t* extprice,
                          uint8_t* returnflag, uint8_t* linestatus, int32_t* receiptdate,
                          int nthreads);
int gdb_gen_customer(uint64_t seed, uint64_t n, int32_t* custkey, uint8_t* segment, int nthreads);
int gdb_gen_orders(uint64_t seed, uint64_t order0, uint64_t n, uint64_t ncustomers,
                   int32_t* orderkey, int32_t* custkey, int32_t* orderdate,
                   int32_t* shippriority, uint8_t* comment /* n*80 or NULL */, int nthreads);
/* Rows of the joined fact table generated by orders [order0, order0+n). */
uint64_t gdb_gen_lineitem_rows(uint64_t seed, uint64_t order0, uint64_t n);
int gdb_gen_lineitem_joined(uint64_t seed, uint64_t order0, uint64_t n, uint64_t nparts,
                            uint64_t nsuppliers, int32_t* orderkey, int32_t* partkey,
                            int32_t* suppkey, int32_t* quantity, int32_t* discount,
                            int64_t* extprice, int32_t* shipdate, int nthreads);
int gdb_gen_part(uint64_t seed, uint64_t n, int32_t* partkey, uint8_t* name /* n*55 */,
                 int nthreads);
int gdb_gen_partsupp(uint64_t seed, uint64_t nparts, uint64_t nsuppliers, int32_t* partkey,
                     int32_t* suppkey, int64_t* supplycost, int nthreads);
int gdb_gen_supplier(uint64_t seed, uint64_t n, int32_t* suppkey, int32_t* nationkey,
                     int nthreads);
/* ---- TPC-H .tbl ingest (replaces MappedFile, countLines and the
 *      parse*ColumnChunk loaders, reference src/infra.h:59-171, and the SF100
 *      chunk parsers, src/infra.h:262-361). Host-side; callers own the arrays.
 * A field is a 0-based '|'-separated column of every line. Decimals become
 * exact integers in hundredths (money in cents: "21168.23" -> 2116823;
 * discount and tax as int32: "0.06" -> 6), dates day numbers from
 * 1992-01-01, CHAR the first byte, TEXT a fixed-width NUL-padded field
 * (truncated to `width`). Malformed rows fail with GDB_E_INVALID naming the
 * row and field. */
enum {
  GDB_TBL_INT32 = 1,      /* int32_t (integer-valued decimals such as "17.00" accepted) */
  GDB_TBL_DATE = 2,       /* int32_t day number, YYYY-MM-DD */
  GDB_TBL_CENTS = 3,      /* int64_t hundredths */
  GDB_TBL_HUNDREDTHS = 4, /* int32_t hundredths */
  GDB_TBL_CHAR = 5,       /* uint8_t first byte */
  GDB_TBL_TEXT = 6        /* width bytes per row */
};
typedef struct {
  int field;  /* 0-based column of the line */
  int kind;   /* GDB_TBL_* */
  int width;  /* GDB_TBL_TEXT only */
  void* out;  /* rows values */
} gdb_tbl_field;
/* Rows of a .tbl file (non-empty lines). */
int gdb_tbl_count_rows(const char* path, int nthreads, uint64_t* rows);
/* Parses `rows` rows (as counted) into the fields' arrays, in parallel. */
int gdb_tbl_parse(const char* path, const gdb_tbl_field* fields, int nfields, uint64_t rows,
                  int nthreads);

/* Name of nation key 0..24 / market segment code 0..4 (TPC-H). */
const char* gdb_nation_name(int nationkey);
const char* gdb_segment_name(int code);
/* Calendar: year of a day number (day 0 = 1992-01-01). */
int32_t gdb_year_of_day(int32_t day);

#ifdef __cplusplus
}
#endif
#endif /* GDB_H */
