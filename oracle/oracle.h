/*
 * oracle.h — CPU restatement of the reference's five benchmarked pipelines.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only to check or to
 * time beside the GPU engine. The engine (libgdb.so) never links or calls it.
 *
 * Semantics are those of the reference's SQL (sql/q{1,3,6,9,13}.sql), run by
 * its CPU comparison engine (DuckDB, scripts/benchmark_duckdb.sh:66-103), with
 * decimal money held as exact int64 fixed point:
 *   extprice, supplycost : cents;  discount, tax : hundredths.
 * Every function takes plain column arrays and a thread count (std pthreads).
 */
#ifndef GDB_ORACLE_H
#define GDB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Q6: SUM(extprice*discount) over date_lo <= ship < date_hi,
 * disc_lo <= disc <= disc_hi, qty < qty_lt. */
void oracle_q6(uint64_t n, const int32_t* shipdate, const int32_t* discount,
               const int32_t* quantity, const int64_t* extprice, int32_t date_lo,
               int32_t date_hi, int32_t disc_lo, int32_t disc_hi, int32_t qty_lt,
               int nthreads, int64_t* revenue, uint64_t* rows);

/* Q1: per group g = rf*2+ls (rf A,N,R = 0,1,2; ls F,O = 0,1), out[g*5 + k]
 * for k = sum_qty, sum_base, sum_disc_price, sum_disc, count; the 128-bit
 * sum_charge in charge_lo[g] (low 64 bits) and charge_hi[g] (high 64 bits). */
/* Returns 0, or -1 (nothing written) when rows x max|price| x (100 + max|disc|)
 * reaches 2^63, where its int64 sums could wrap. */
int oracle_q1(uint64_t n, const int32_t* shipdate, const int32_t* quantity,
              const int32_t* discount, const int32_t* tax, const int64_t* extprice,
              const uint8_t* returnflag, const uint8_t* linestatus, int32_t cutoff,
              int nthreads, int64_t* out /* [30] */, uint64_t* charge_lo /* [6] */,
              int64_t* charge_hi /* [6] */);

/* Q3: top-k rows (orderkey, orderdate, shippriority, revenue) by revenue desc,
 * orderdate asc, orderkey asc; returns the number written, *groups = number of
 * qualifying orderkeys. */
int oracle_q3(uint64_t nc, const int32_t* c_custkey, const uint8_t* c_segment, uint64_t no,
              const int32_t* o_orderkey, const int32_t* o_custkey, const int32_t* o_orderdate,
              const int32_t* o_shippriority, uint64_t nl, const int32_t* l_orderkey,
              const int32_t* l_shipdate, const int32_t* l_discount, const int64_t* l_extprice,
              int32_t segment, int32_t date, int k, int nthreads, int32_t* out_orderkey,
              int32_t* out_orderdate, int32_t* out_shippriority, int64_t* out_revenue,
              int64_t* groups);

/* Q9: rows (nationkey, year, profit) ordered by nation name asc, year desc;
 * returns the number written (at most max_rows). */
int oracle_q9(uint64_t np, const int32_t* p_partkey, const uint8_t* p_name, int name_width,
              uint64_t ns, const int32_t* s_suppkey, const int32_t* s_nationkey, uint64_t nps,
              const int32_t* ps_partkey, const int32_t* ps_suppkey, const int64_t* ps_supplycost,
              uint64_t no, const int32_t* o_orderkey, const int32_t* o_orderdate, uint64_t nl,
              const int32_t* l_partkey, const int32_t* l_suppkey, const int32_t* l_orderkey,
              const int32_t* l_quantity, const int32_t* l_discount, const int64_t* l_extprice,
              const char* pattern, int nthreads, int32_t* out_nation, int32_t* out_year,
              int64_t* out_profit, int max_rows);

/* Q13: rows (c_count, custdist) ordered custdist desc, c_count desc; returns
 * the number written (at most max_rows). */
int oracle_q13(uint64_t nc, const int32_t* c_custkey, uint64_t no, const int32_t* o_custkey,
               const uint8_t* o_comment, int comment_width, const char* word1, const char* word2,
               int nthreads, int32_t* out_count, int64_t* out_custdist, int max_rows);

/* Primitives: selection `col < value` (bitmap of ceil(n/64) words, may be
 * NULL), SUM, equi-join pair count. */
uint64_t oracle_select_lt(uint64_t n, const int32_t* col, int32_t value, uint64_t* bitmap);
int64_t oracle_sum_i32(uint64_t n, const int32_t* col);
int64_t oracle_sum_i64(uint64_t n, const int64_t* col);
uint64_t oracle_join_count(uint64_t nb, const int32_t* build, uint64_t np, const int32_t* probe);

/* Substring helpers the pipelines use, exposed for direct tests. */
int oracle_contains(const uint8_t* s, int width, const char* pattern);
int oracle_like_two(const uint8_t* s, int width, const char* w1, const char* w2);
int32_t oracle_year_of_day(int32_t day);

#ifdef __cplusplus
}
#endif
#endif
