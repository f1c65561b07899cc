/*
 * gdb_tpch.h — C ABI of libgdb.so's remaining TPC-H pipelines: Q2, Q4, Q5,
 * Q7, Q8, Q10, Q11, Q12, Q14, Q15, Q16, Q17, Q18, Q19, Q20, Q21, Q22.
 *
 * Each gdb_qN replaces the reference's runQNBenchmark(MTL::Device*,
 * MTL::CommandQueue*, MTL::Library*) (reference src/infra.h:874-922,
 * registered by name in src/main.cpp:161-188). The conventions are those of
 * gdb.h: synchronous calls on resident tables, GDB_OK or a negative GDB_E_*
 * code, exact integer results, day numbers from 1992-01-01 = day 0.
 *
 * Tables are passed together as a gdb_tpch_db; a pipeline reads only the
 * tables and columns it needs (named as in TPC-H) and fails with
 * GDB_E_INVALID naming any that is missing. Low-cardinality text columns are
 * dictionary codes (uint8): p_type 0..149 (syllables s1*25 + s2*5 + s3 of
 * STANDARD SMALL MEDIUM LARGE ECONOMY PROMO / ANODIZED BURNISHED PLATED
 * POLISHED BRUSHED / TIN NICKEL BRASS STEEL COPPER), p_container 0..39
 * (SM LG MED JUMBO WRAP x CASE BOX BAG JAR PKG PACK CAN DRUM), p_brand
 * 10*M+N for Brand#MN, l_shipmode 0..6 (REG AIR, AIR, RAIL, SHIP, TRUCK,
 * MAIL, FOB), l_shipinstruct 0..3 (DELIVER IN PERSON, COLLECT COD, NONE,
 * TAKE BACK RETURN), o_orderpriority 0..4 (1-URGENT .. 5-LOW). A LIKE or IN
 * predicate over such a column is passed as the set of codes it selects, a
 * bit mask. Keys are direct-addressed: every key column must lie in
 * [0, 2^31 - 2] (GDB_E_UNSUPPORTED otherwise).
 *
 * Revenue sums SUM(l_extendedprice * (1 - l_discount)) are exact int64 in
 * 1e-4 dollars (cents * (100 - discount hundredths)).
 */
#ifndef GDB_TPCH_H
#define GDB_TPCH_H

#include "gdb.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const gdb_table* lineitem;  /* l_orderkey l_partkey l_suppkey l_quantity l_extendedprice
                                 l_discount l_returnflag l_shipdate l_commitdate
                                 l_receiptdate l_shipinstruct l_shipmode */
  const gdb_table* orders;    /* o_orderkey o_custkey o_orderstatus o_totalprice
                                 o_orderdate o_orderpriority */
  const gdb_table* customer;  /* c_custkey c_nationkey c_acctbal */
  const gdb_table* part;      /* p_partkey p_name p_mfgr p_brand p_type p_size p_container */
  const gdb_table* partsupp;  /* ps_partkey ps_suppkey ps_availqty ps_supplycost */
  const gdb_table* supplier;  /* s_suppkey s_nationkey s_acctbal s_comment */
  const gdb_table* nation;    /* n_nationkey n_regionkey (25 rows) */
} gdb_tpch_db;

/* Q2 minimum-cost supplier (reference src/q2.cpp, kernels/q2.metal:33-124).
 * Parts of size `size` whose type is in type_mask (LIKE '%BRASS'), partsupp
 * rows of suppliers in `region` with the part's minimum cost in that region.
 * Rows ordered s_acctbal desc, n_name, s_suppkey, p_partkey; top_k <= 100. */
typedef struct { int32_t size, region, top_k, pad; uint64_t type_mask[3]; } gdb_q2_params;
typedef struct { int64_t s_acctbal; int32_t suppkey, nationkey, partkey, mfgr; } gdb_q2_row;
typedef struct { int32_t nrows; int32_t pad; int64_t matches; gdb_q2_row rows[100]; } gdb_q2_result;
int gdb_q2(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q2_params* p, gdb_q2_result* out,
           gdb_timing* timing);

/* Q4 order priority checking (reference src/q4.cpp, kernels/q4.metal:21-90).
 * Orders dated in [date_lo, date_hi) with a line whose commitdate <
 * receiptdate, counted per o_orderpriority code 0..4 (counts[code]). */
typedef struct { int32_t date_lo, date_hi; } gdb_q4_params;
typedef struct { int64_t counts[5]; } gdb_q4_result;
int gdb_q4(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q4_params* p, gdb_q4_result* out,
           gdb_timing* timing);

/* Q5 local supplier volume (reference src/q5.cpp, kernels/q5.metal:19-125).
 * Revenue per nation of `region` from lines whose supplier's nation equals the
 * customer's, orders dated in [date_lo, date_hi). revenue[nationkey]; `has`
 * bit n set when nation n has a line. */
typedef struct { int32_t region, date_lo, date_hi; } gdb_q5_params;
typedef struct { int64_t revenue[25]; uint32_t has; } gdb_q5_result;
int gdb_q5(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q5_params* p, gdb_q5_result* out,
           gdb_timing* timing);

/* Q7 volume shipping (reference src/q7.cpp, kernels/q7.metal:13-109).
 * Lines shipped in [date_lo, date_hi] (inclusive) between nation1 and nation2
 * either way. revenue[d][y]: d = 0 for supplier nation1 -> customer nation2,
 * 1 for the reverse; y = year - 1992 (0..7); has[d] bit y when present. */
typedef struct { int32_t nation1, nation2, date_lo, date_hi; } gdb_q7_params;
typedef struct { int64_t revenue[2][8]; uint32_t has[2]; } gdb_q7_result;
int gdb_q7(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q7_params* p, gdb_q7_result* out,
           gdb_timing* timing);

/* Q8 national market share (reference src/q8.cpp, kernels/q8.metal:10-76).
 * Parts of type `type_code`, customers in `region`, orders dated in
 * [date_lo, date_hi] (inclusive); per order year y - 1992: total volume and the
 * volume supplied from `nation`. mkt_share = nation_volume / total_volume. */
typedef struct { int32_t nation, region, type_code, date_lo, date_hi; } gdb_q8_params;
typedef struct { int64_t nation_volume[8], total_volume[8]; uint32_t has; } gdb_q8_result;
int gdb_q8(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q8_params* p, gdb_q8_result* out,
           gdb_timing* timing);

/* Q10 returned item reporting (reference src/q10.cpp, kernels/q10.metal:9-54).
 * Revenue per customer of lines with l_returnflag = returnflag in orders dated
 * in [date_lo, date_hi); top_k (<= 100) by revenue desc, custkey asc.
 * `groups` = customers with such a line. */
typedef struct { int32_t date_lo, date_hi, top_k; char returnflag; char pad[3]; } gdb_q10_params;
typedef struct { int32_t custkey, nationkey; int64_t revenue, acctbal; } gdb_q10_row;
typedef struct { int32_t nrows; int32_t pad; int64_t groups; gdb_q10_row rows[100]; } gdb_q10_result;
int gdb_q10(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q10_params* p, gdb_q10_result* out,
            gdb_timing* timing);

/* Q11 important stock identification (reference src/q11.cpp,
 * kernels/q11.metal:27-59). value = SUM(supplycost * availqty) per partkey over
 * suppliers of `nation`; rows with value * frac_den > total * frac_num,
 * ordered value desc, partkey asc. Up to `cap` rows go to rows[]; *nrows is
 * the number that qualify. */
typedef struct { int32_t nation; int32_t pad; int64_t frac_num, frac_den; } gdb_q11_params;
typedef struct { int32_t partkey; int32_t pad; int64_t value; } gdb_q11_row;
int gdb_q11(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q11_params* p, gdb_q11_row* rows,
            int64_t cap, int64_t* nrows, int64_t* total, gdb_timing* timing);

/* Q12 shipping modes and order priority (reference src/q12.cpp,
 * kernels/q12.metal:22-100). Lines with shipmode mode[i], commitdate <
 * receiptdate, shipdate < commitdate, receiptdate in [date_lo, date_hi);
 * high[i] / low[i] = lines whose order priority is / is not in high_mask. */
typedef struct { int32_t mode[2], date_lo, date_hi; uint32_t high_mask; } gdb_q12_params;
typedef struct { int64_t high[2], low[2]; } gdb_q12_result;
int gdb_q12(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q12_params* p, gdb_q12_result* out,
            gdb_timing* timing);

/* Q14 promotion effect (reference src/q14.cpp, kernels/q14.metal:17-95).
 * Lines shipped in [date_lo, date_hi): promo = revenue of parts whose type is
 * in type_mask (LIKE 'PROMO%'), total = all revenue; promo_revenue =
 * 100 * promo / total. */
typedef struct { int32_t date_lo, date_hi; uint64_t type_mask[3]; } gdb_q14_params;
typedef struct { int64_t promo, total; double promo_revenue; } gdb_q14_result;
int gdb_q14(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q14_params* p, gdb_q14_result* out,
            gdb_timing* timing);

/* Q15 top supplier (reference src/q15.cpp, kernels/q15.metal:13-32). Revenue
 * per supplier over shipdate in [date_lo, date_hi); the suppliers whose
 * revenue equals the maximum, ordered by suppkey (up to 100). */
typedef struct { int32_t date_lo, date_hi; } gdb_q15_params;
typedef struct { int32_t nrows; int32_t pad; int64_t max_revenue; int32_t suppkey[100]; } gdb_q15_result;
int gdb_q15(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q15_params* p, gdb_q15_result* out,
            gdb_timing* timing);

/* Q16 parts/supplier relationship (reference src/q16.cpp,
 * kernels/q16.metal:12-66). Parts with p_brand != brand, type not in
 * type_mask (NOT LIKE 'MEDIUM POLISHED%'), size in sizes[0..nsizes); partsupp
 * rows whose supplier's comment does not match %word1%word2%. Rows (brand,
 * type, size, COUNT(DISTINCT suppkey)) ordered count desc, brand, type name,
 * size; up to `cap` to rows[], *nrows = groups. */
typedef struct {
  int32_t brand, nsizes; int32_t sizes[8]; uint64_t type_mask[3];
  char word1[32], word2[32];
} gdb_q16_params;
typedef struct { int32_t brand, type, size, supplier_cnt; } gdb_q16_row;
int gdb_q16(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q16_params* p, gdb_q16_row* rows,
            int64_t cap, int64_t* nrows, gdb_timing* timing);

/* Q17 small-quantity-order revenue (reference src/q17.cpp,
 * kernels/q17.metal:11-57). Lines of parts (brand, container) with quantity *
 * frac_den < frac_num * AVG(quantity over that part's lines); sum_price =
 * SUM(l_extendedprice) in cents; avg_yearly = sum_price / 7 in dollars. */
typedef struct { int32_t brand, container, frac_num, frac_den; } gdb_q17_params;
typedef struct { int64_t sum_price; double avg_yearly; } gdb_q17_result;
int gdb_q17(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q17_params* p, gdb_q17_result* out,
            gdb_timing* timing);

/* Q18 large volume customer (reference src/q18.cpp, kernels/q18.metal:9-61).
 * Orders whose lines' quantity sums to more than `quantity`, top_k (<= 100) by
 * totalprice desc, orderdate asc, orderkey asc. `orders_over` = all such. */
typedef struct { int32_t quantity, top_k; } gdb_q18_params;
typedef struct { int32_t custkey, orderkey, orderdate, sum_qty; int64_t totalprice; } gdb_q18_row;
typedef struct { int32_t nrows; int32_t pad; int64_t orders_over; gdb_q18_row rows[100]; } gdb_q18_result;
int gdb_q18(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q18_params* p, gdb_q18_result* out,
            gdb_timing* timing);

/* Q19 discounted revenue (reference src/q19.cpp, kernels/q19.metal:22-160).
 * Lines with shipmode in mode_mask and shipinstruct = instruct whose part
 * matches one of the blocks: p_brand = brand, p_container in container_mask,
 * size in [size_lo, size_hi], quantity in [qty_lo, qty_hi]. */
typedef struct { int32_t brand, qty_lo, qty_hi, size_lo, size_hi, pad; uint64_t container_mask; } gdb_q19_block;
typedef struct { int32_t nblocks; uint32_t mode_mask; int32_t instruct, pad; gdb_q19_block blocks[3]; } gdb_q19_params;
typedef struct { int64_t revenue; } gdb_q19_result;
int gdb_q19(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q19_params* p, gdb_q19_result* out,
            gdb_timing* timing);

/* Q20 potential part promotion (reference src/q20.cpp,
 * kernels/q20.metal:15-104). Suppliers of `nation` with a partsupp row of a
 * part whose name starts with `prefix`, where availqty * frac_den > frac_num
 * * SUM(quantity of that pair's lines shipped in [date_lo, date_hi)) and such
 * lines exist. suppkeys ascending (s_name order); up to `cap`. */
typedef struct { char prefix[32]; int32_t date_lo, date_hi, nation, frac_num, frac_den, pad; } gdb_q20_params;
int gdb_q20(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q20_params* p, int32_t* suppkeys,
            int64_t cap, int64_t* nrows, gdb_timing* timing);

/* Q21 suppliers who kept orders waiting (reference src/q21.cpp,
 * kernels/q21.metal:24-119). Late lines (receipt > commit) of orders with
 * o_orderstatus = status, by suppliers of `nation`, where the order has another
 * supplier and no other supplier's line is late; count per supplier, top_k (<=
 * 100) by count desc, suppkey asc. */
typedef struct { int32_t nation, top_k; char status; char pad[7]; } gdb_q21_params;
typedef struct { int32_t suppkey; int32_t pad; int64_t numwait; } gdb_q21_row;
typedef struct { int32_t nrows; int32_t pad; int64_t suppliers; gdb_q21_row rows[100]; } gdb_q21_result;
int gdb_q21(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q21_params* p, gdb_q21_result* out,
            gdb_timing* timing);

/* Q22 global sales opportunity (reference src/q22.cpp,
 * kernels/q22.metal:11-67). Country code = c_nationkey + 10 (the first two
 * characters of c_phone). Customers with a code in codes[0..ncodes), balance
 * above the average positive balance of those codes, and no orders; count and
 * balance sum per code (rows ordered by code). */
typedef struct { int32_t ncodes; int32_t codes[7]; } gdb_q22_params;
typedef struct { int32_t cntrycode; int32_t pad; int64_t numcust, totacctbal; } gdb_q22_row;
typedef struct { int32_t nrows; int32_t pad; gdb_q22_row rows[7]; } gdb_q22_result;
int gdb_q22(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q22_params* p, gdb_q22_result* out,
            gdb_timing* timing);

#ifdef __cplusplus
}
#endif
#endif /* GDB_TPCH_H */
