// gdb_bench's TPC-H side: the other 17 pipelines of the reference's registry
// (src/main.cpp:161-188: q2, q4, q5, q7, q8, q10, q11, q12, q14 .. q22) over a
// whole TPC-H database read from dbgen .tbl files (the reference's data/SF-N
// directories, or --data DIR). Each query runs with the reference SQL's
// constants (sql/qN.sql), 2 warm-ups + 10 timed runs, and prints a result line
// and the reference's `Qn | N rows` header with the printTimingSummary lines
// (src/infra.h:511-516), like the five pipelines of driver.cpp.
//
// Text with few values is dictionary-coded on load, as include/gdb_tpch.h
// expects: p_type, p_container, l_shipmode, l_shipinstruct, o_orderpriority,
// c_mktsegment by exact match against the TPC-H word lists, p_brand and p_mfgr
// as the number after '#'.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "gdb_tpch.h"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

void die(int rc, const char* what) {
  fprintf(stderr, "%s failed: %s (%s)\n", what, gdb_strerror(rc), gdb_last_error());
  exit(1);
}
#define TCHECK(x)                    \
  do {                               \
    int rc_ = (x);                   \
    if (rc_ != GDB_OK) die(rc_, #x); \
  } while (0)

// TPC-H word lists (specification clause 4.2.2.13); the code of a value is its index.
std::vector<std::string> type_names() {
  static const char* s1[6] = {"STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO"};
  static const char* s2[5] = {"ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED"};
  static const char* s3[5] = {"TIN", "NICKEL", "BRASS", "STEEL", "COPPER"};
  std::vector<std::string> v;
  for (auto a : s1)
    for (auto b : s2)
      for (auto c : s3) v.push_back(std::string(a) + " " + b + " " + c);
  return v;
}
std::vector<std::string> container_names() {
  static const char* c1[5] = {"SM", "LG", "MED", "JUMBO", "WRAP"};
  static const char* c2[8] = {"CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM"};
  std::vector<std::string> v;
  for (auto a : c1)
    for (auto b : c2) v.push_back(std::string(a) + " " + b);
  return v;
}
const std::vector<std::string> SHIPMODE = {"REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"};
const std::vector<std::string> SHIPINSTRUCT = {"DELIVER IN PERSON", "COLLECT COD", "NONE",
                                               "TAKE BACK RETURN"};
const std::vector<std::string> PRIORITY = {"1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED",
                                           "5-LOW"};
const std::vector<std::string> SEGMENT = {"AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD",
                                          "MACHINERY"};
const char* REGION[5] = {"AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"};

int nation_key(const char* name) {
  for (int k = 0; k < 25; ++k)
    if (!strcmp(gdb_nation_name(k), name)) return k;
  return -1;
}
int region_key(const char* name) {
  for (int k = 0; k < 5; ++k)
    if (!strcmp(REGION[k], name)) return k;
  return -1;
}

template <typename F>
void parallel_rows(uint64_t n, F f) {
  const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint64_t per = (n + T - 1) / T;
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) {
    const uint64_t a = (uint64_t)t * per, b = std::min(n, a + per);
    if (a < b) th.emplace_back([=] { f(a, b); });
  }
  for (auto& x : th) x.join();
}

// Fixed-width text rows -> codes (255 for values outside the list).
std::vector<uint8_t> dict_codes(const uint8_t* text, int width, uint64_t n,
                                const std::vector<std::string>& dict) {
  std::vector<uint8_t> out(n);
  parallel_rows(n, [&](uint64_t a, uint64_t b) {
    for (uint64_t r = a; r < b; ++r) {
      const char* s = (const char*)text + r * width;
      const size_t len = strnlen(s, width);
      uint8_t c = 255;
      for (size_t i = 0; i < dict.size(); ++i)
        if (dict[i].size() == len && !memcmp(dict[i].data(), s, len)) {
          c = (uint8_t)i;
          break;
        }
      out[r] = c;
    }
  });
  return out;
}

// 'Brand#23' -> 23, 'Manufacturer#4' -> 4.
std::vector<uint8_t> tail_numbers(const uint8_t* text, int width, uint64_t n) {
  std::vector<uint8_t> out(n);
  for (uint64_t r = 0; r < n; ++r) {
    const char* s = (const char*)text + r * width;
    const size_t len = strnlen(s, width);
    const char* h = (const char*)memrchr(s, '#', len);
    out[r] = h ? (uint8_t)atoi(h + 1) : 0;
  }
  return out;
}

// One table of the database: parsed columns kept alive until the upload.
struct Loaded {
  std::vector<std::vector<uint8_t>> store;
  std::vector<gdb_col_desc> cols;
  std::vector<std::string> names;
  uint64_t n = 0;
};

// kind: GDB_TBL_* or one of these for dictionary / '#'-number text.
enum { DICT_TYPE = 100, DICT_CONTAINER, DICT_MODE, DICT_INSTRUCT, DICT_PRIORITY, DICT_SEGMENT,
       TAIL_NUMBER };
struct Spec {
  int field;
  const char* name;
  int kind;
  int width;
};

Loaded load_table(const std::string& dir, const char* table, const std::vector<Spec>& spec) {
  Loaded L;
  const std::string path = dir + "/" + table + ".tbl";
  TCHECK(gdb_tbl_count_rows(path.c_str(), 0, &L.n));
  const uint64_t n = L.n;
  std::vector<gdb_tbl_field> f;
  L.store.resize(spec.size());
  for (size_t i = 0; i < spec.size(); ++i) {
    const Spec& s = spec[i];
    int kind = s.kind, width = s.width;
    size_t elem = 4;
    if (kind >= DICT_TYPE) {
      kind = GDB_TBL_TEXT;
      elem = 1;
    } else if (kind == GDB_TBL_CENTS) {
      elem = 8;
    } else if (kind == GDB_TBL_CHAR) {
      elem = 1;
    } else if (kind == GDB_TBL_TEXT) {
      elem = 1;
    }
    const uint64_t bytes = n * elem * (kind == GDB_TBL_TEXT ? width : 1) + 64;
    L.store[i].resize(bytes);
    f.push_back(gdb_tbl_field{s.field, kind, kind == GDB_TBL_TEXT ? width : 0, L.store[i].data()});
  }
  if (n) TCHECK(gdb_tbl_parse(path.c_str(), f.data(), (int)f.size(), n, 0));
  static const std::vector<std::string> types = type_names(), conts = container_names();
  for (size_t i = 0; i < spec.size(); ++i) {
    const Spec& s = spec[i];
    int type = GDB_I32, width = 0;
    if (s.kind >= DICT_TYPE) {
      const uint8_t* t = L.store[i].data();
      std::vector<uint8_t> codes =
          s.kind == DICT_TYPE        ? dict_codes(t, s.width, n, types)
          : s.kind == DICT_CONTAINER ? dict_codes(t, s.width, n, conts)
          : s.kind == DICT_MODE      ? dict_codes(t, s.width, n, SHIPMODE)
          : s.kind == DICT_INSTRUCT  ? dict_codes(t, s.width, n, SHIPINSTRUCT)
          : s.kind == DICT_PRIORITY  ? dict_codes(t, s.width, n, PRIORITY)
          : s.kind == DICT_SEGMENT   ? dict_codes(t, s.width, n, SEGMENT)
                                     : tail_numbers(t, s.width, n);
      codes.resize(n + 64);
      L.store[i] = std::move(codes);
      type = GDB_U8;
    } else if (s.kind == GDB_TBL_CENTS) {
      type = GDB_I64;
    } else if (s.kind == GDB_TBL_CHAR) {
      type = GDB_U8;
    } else if (s.kind == GDB_TBL_TEXT) {
      type = GDB_BYTES;
      width = s.width;
    }
    L.names.push_back(s.name);
    L.cols.push_back(gdb_col_desc{nullptr, type, width, L.store[i].data()});
  }
  for (size_t i = 0; i < L.cols.size(); ++i) L.cols[i].name = L.names[i].c_str();
  return L;
}

struct Db {
  gdb_table* t[7] = {};  // lineitem, orders, customer, part, partsupp, supplier, nation
  uint64_t rows[7] = {};
  double parse_ms = 0;
  Loaded streamed;       // a streamed lineitem's host columns (referenced, not copied)
  gdb_tpch_db view() const { return gdb_tpch_db{t[0], t[1], t[2], t[3], t[4], t[5], t[6]}; }
};

// stream: lineitem stays in host memory and the pipelines stream it in chunks
// (the reference's SF100 mode); the other tables are resident.
Db load_db(gdb_ctx* ctx, const std::string& dir, bool stream) {
  const auto t0 = Clock::now();
  const std::vector<std::pair<const char*, std::vector<Spec>>> tables = {
      {"lineitem",
       {{0, "l_orderkey", GDB_TBL_INT32, 0}, {1, "l_partkey", GDB_TBL_INT32, 0},
        {2, "l_suppkey", GDB_TBL_INT32, 0}, {4, "l_quantity", GDB_TBL_INT32, 0},
        {5, "l_extendedprice", GDB_TBL_CENTS, 0}, {6, "l_discount", GDB_TBL_HUNDREDTHS, 0},
        {7, "l_tax", GDB_TBL_HUNDREDTHS, 0}, {8, "l_returnflag", GDB_TBL_CHAR, 0},
        {9, "l_linestatus", GDB_TBL_CHAR, 0}, {10, "l_shipdate", GDB_TBL_DATE, 0},
        {11, "l_commitdate", GDB_TBL_DATE, 0}, {12, "l_receiptdate", GDB_TBL_DATE, 0},
        {13, "l_shipinstruct", DICT_INSTRUCT, 25}, {14, "l_shipmode", DICT_MODE, 10}}},
      {"orders",
       {{0, "o_orderkey", GDB_TBL_INT32, 0}, {1, "o_custkey", GDB_TBL_INT32, 0},
        {2, "o_orderstatus", GDB_TBL_CHAR, 0}, {3, "o_totalprice", GDB_TBL_CENTS, 0},
        {4, "o_orderdate", GDB_TBL_DATE, 0}, {5, "o_orderpriority", DICT_PRIORITY, 15}}},
      {"customer",
       {{0, "c_custkey", GDB_TBL_INT32, 0}, {3, "c_nationkey", GDB_TBL_INT32, 0},
        {5, "c_acctbal", GDB_TBL_CENTS, 0}, {6, "c_mktsegment", DICT_SEGMENT, 10}}},
      {"part",
       {{0, "p_partkey", GDB_TBL_INT32, 0}, {1, "p_name", GDB_TBL_TEXT, 55},
        {2, "p_mfgr", TAIL_NUMBER, 25}, {3, "p_brand", TAIL_NUMBER, 10},
        {4, "p_type", DICT_TYPE, 25}, {5, "p_size", GDB_TBL_INT32, 0},
        {6, "p_container", DICT_CONTAINER, 10}}},
      {"partsupp",
       {{0, "ps_partkey", GDB_TBL_INT32, 0}, {1, "ps_suppkey", GDB_TBL_INT32, 0},
        {2, "ps_availqty", GDB_TBL_INT32, 0}, {3, "ps_supplycost", GDB_TBL_CENTS, 0}}},
      {"supplier",
       {{0, "s_suppkey", GDB_TBL_INT32, 0}, {3, "s_nationkey", GDB_TBL_INT32, 0},
        {5, "s_acctbal", GDB_TBL_CENTS, 0}, {6, "s_comment", GDB_TBL_TEXT, 101}}},
      {"nation", {{0, "n_nationkey", GDB_TBL_INT32, 0}, {2, "n_regionkey", GDB_TBL_INT32, 0}}}};
  Db db;
  for (size_t i = 0; i < tables.size(); ++i) {
    Loaded L = load_table(dir, tables[i].first, tables[i].second);
    db.rows[i] = L.n;
    if (stream && i == 0) {
      db.streamed = std::move(L);
      for (size_t c = 0; c < db.streamed.cols.size(); ++c)
        db.streamed.cols[c].name = db.streamed.names[c].c_str();
      TCHECK(gdb_table_create_host(ctx, db.streamed.cols.data(), (int)db.streamed.cols.size(),
                                   db.streamed.n, &db.t[i]));
    } else {
      TCHECK(gdb_table_create(ctx, L.cols.data(), (int)L.cols.size(), L.n, &db.t[i]));
    }
  }
  db.parse_ms = ms_since(t0);
  return db;
}

void timing_lines(const char* q, uint64_t rows, double parse_ms, const gdb_timing& t) {
  printf("\n%s | %llu rows\n", q, (unsigned long long)rows);
  printf("  CPU Parsing (.tbl): %10.2f ms\n", parse_ms);
  printf("  GPU Execution:      %10.2f ms\n", t.gpu_ms);
  printf("  CPU Post Process:   %10.2f ms\n", t.post_ms);
  printf("  Total Execution:    %10.2f ms  (GPU + CPU post)\n", t.gpu_ms + t.post_ms);
  printf("  Bandwidth:          %10.2f GB/s\n", t.bytes / (t.gpu_ms / 1000.0) / 1e9);
}

gdb_timing timed(const std::function<void(gdb_timing*)>& fn) {
  std::vector<gdb_timing> ts;
  for (int i = 0; i < 12; ++i) {
    gdb_timing t{};
    fn(&t);
    if (i >= 2) ts.push_back(t);
  }
  std::sort(ts.begin(), ts.end(), [](const gdb_timing& a, const gdb_timing& b) { return a.gpu_ms < b.gpu_ms; });
  return ts[ts.size() / 2];
}

double dollars4(int64_t v) { return (double)v / 1e4; }  // cents * hundredths -> dollars

void mask_set(uint64_t* m, int c) { m[c / 64] |= 1ull << (c % 64); }

}  // namespace

// The TPC-H pipelines this translation unit runs.
const std::vector<std::string>& tpch_queries() {
  static const std::vector<std::string> q = {"q2", "q4", "q5", "q7", "q8", "q10", "q11", "q12",
                                             "q14", "q15", "q16", "q17", "q18", "q19", "q20",
                                             "q21", "q22"};
  return q;
}

// Runs `queries` over the database in `dir` (lineitem streamed from host memory
// when `stream`); returns a process exit code.
int run_tpch(gdb_ctx* ctx, const std::string& dir, const std::vector<std::string>& queries,
             bool stream) {
  Db db = load_db(ctx, dir, stream);
  const gdb_tpch_db v = db.view();
  const uint64_t L = db.rows[0];
  const std::vector<std::string> types = type_names(), conts = container_names();
  auto type_code = [&](const char* s) {
    return (int)(std::find(types.begin(), types.end(), s) - types.begin());
  };
  auto cont_code = [&](const char* s) {
    return (int)(std::find(conts.begin(), conts.end(), s) - conts.begin());
  };
  for (const std::string& q : queries) {
    printf("--- Running TPC-H Query %s Benchmark ---\n", q.c_str() + 1);
    gdb_timing t{};
    if (q == "q2") {
      gdb_q2_params p{15, region_key("EUROPE"), 100, 0, {0, 0, 0}};
      for (int c = 0; c < 150; ++c)
        if (types[c].size() >= 5 && types[c].compare(types[c].size() - 5, 5, "BRASS") == 0)
          mask_set(p.type_mask, c);
      gdb_q2_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q2(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q2: %d rows (of %lld minimum-cost offers)\n", r.nrows, (long long)r.matches);
      for (int i = 0; i < std::min(r.nrows, 5); ++i)
        printf("  %.2f Supplier#%09d %s part %d Manufacturer#%d\n", r.rows[i].s_acctbal / 100.0,
               r.rows[i].suppkey, gdb_nation_name(r.rows[i].nationkey), r.rows[i].partkey,
               r.rows[i].mfgr);
      timing_lines("Q2", db.rows[4], db.parse_ms, t);
    } else if (q == "q4") {
      gdb_q4_params p{547, 639};  // [1993-07-01, 1993-10-01)
      gdb_q4_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q4(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q4 Results:\n");
      for (int i = 0; i < 5; ++i)
        if (r.counts[i]) printf("  %-16s %lld\n", PRIORITY[i].c_str(), (long long)r.counts[i]);
      timing_lines("Q4", L, db.parse_ms, t);
    } else if (q == "q5") {
      gdb_q5_params p{region_key("ASIA"), 731, 1096};
      gdb_q5_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q5(ctx, &v, &p, &r, tm)); });
      std::vector<int> ns;
      for (int n = 0; n < 25; ++n)
        if ((r.has >> n) & 1) ns.push_back(n);
      std::sort(ns.begin(), ns.end(), [&](int a, int b) { return r.revenue[a] > r.revenue[b]; });
      printf("TPC-H Q5 Results:\n");
      for (int n : ns) printf("  %-15s %.4f\n", gdb_nation_name(n), dollars4(r.revenue[n]));
      timing_lines("Q5", L, db.parse_ms, t);
    } else if (q == "q7") {
      const int n1 = nation_key("FRANCE"), n2 = nation_key("GERMANY");
      gdb_q7_params p{n1, n2, 1096, 1826};
      gdb_q7_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q7(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q7 Results:\n");
      for (int d = 0; d < 2; ++d)
        for (int y = 0; y < 8; ++y)
          if ((r.has[d] >> y) & 1)
            printf("  %-8s %-8s %d %.4f\n", gdb_nation_name(d ? n2 : n1),
                   gdb_nation_name(d ? n1 : n2), 1992 + y, dollars4(r.revenue[d][y]));
      timing_lines("Q7", L, db.parse_ms, t);
    } else if (q == "q8") {
      gdb_q8_params p{nation_key("BRAZIL"), region_key("AMERICA"),
                      type_code("ECONOMY ANODIZED STEEL"), 1096, 1826};
      gdb_q8_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q8(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q8 Results:\n");
      for (int y = 0; y < 8; ++y)
        if ((r.has >> y) & 1)
          printf("  %d mkt_share %.6f\n", 1992 + y,
                 r.total_volume[y] ? (double)r.nation_volume[y] / r.total_volume[y] : 0.0);
      timing_lines("Q8", L, db.parse_ms, t);
    } else if (q == "q10") {
      gdb_q10_params p{639, 731, 20, 'R', {0, 0, 0}};
      gdb_q10_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q10(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q10: top %d of %lld customers\n", r.nrows, (long long)r.groups);
      for (int i = 0; i < std::min(r.nrows, 5); ++i)
        printf("  Customer#%09d revenue %.4f acctbal %.2f %s\n", r.rows[i].custkey,
               dollars4(r.rows[i].revenue), r.rows[i].acctbal / 100.0,
               gdb_nation_name(r.rows[i].nationkey));
      timing_lines("Q10", L, db.parse_ms, t);
    } else if (q == "q11") {
      gdb_q11_params p{nation_key("GERMANY"), 0, 1, 10000};
      std::vector<gdb_q11_row> rows(1000);
      int64_t n = 0, total = 0;
      t = timed([&](gdb_timing* tm) {
        TCHECK(gdb_q11(ctx, &v, &p, rows.data(), (int64_t)rows.size(), &n, &total, tm));
      });
      printf("TPC-H Q11: %lld parts above 0.0001 of %.2f\n", (long long)n, total / 100.0);
      for (int64_t i = 0; i < std::min<int64_t>(n, 5); ++i)
        printf("  part %d value %.2f\n", rows[i].partkey, rows[i].value / 100.0);
      timing_lines("Q11", db.rows[4], db.parse_ms, t);
    } else if (q == "q12") {
      const int mail = 5, ship = 3;
      gdb_q12_params p{{mail, ship}, 731, 1096, 3u};
      gdb_q12_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q12(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q12 Results:\n");
      for (int i = 0; i < 2; ++i)
        printf("  %-5s high %lld low %lld\n", SHIPMODE[p.mode[i]].c_str(),
               (long long)r.high[i], (long long)r.low[i]);
      timing_lines("Q12", L, db.parse_ms, t);
    } else if (q == "q14") {
      gdb_q14_params p{1339, 1369, {0, 0, 0}};
      for (int c = 0; c < 150; ++c)
        if (types[c].compare(0, 5, "PROMO") == 0) mask_set(p.type_mask, c);
      gdb_q14_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q14(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q14 Result: Promo Revenue = %.2f%%\n", r.promo_revenue);
      timing_lines("Q14", L, db.parse_ms, t);
    } else if (q == "q15") {
      gdb_q15_params p{1461, 1552};
      gdb_q15_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q15(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q15 Results:\n");
      for (int i = 0; i < r.nrows; ++i)
        printf("  Supplier#%09d total_revenue %.4f\n", r.suppkey[i], dollars4(r.max_revenue));
      timing_lines("Q15", L, db.parse_ms, t);
    } else if (q == "q16") {
      gdb_q16_params p{};
      p.brand = 45;
      const int sizes[8] = {49, 14, 23, 45, 19, 3, 36, 9};
      p.nsizes = 8;
      memcpy(p.sizes, sizes, sizeof sizes);
      for (int c = 0; c < 150; ++c)
        if (types[c].compare(0, 15, "MEDIUM POLISHED") == 0) mask_set(p.type_mask, c);
      strcpy(p.word1, "Customer");
      strcpy(p.word2, "Complaints");
      std::vector<gdb_q16_row> rows(1 << 16);
      int64_t n = 0;
      t = timed([&](gdb_timing* tm) {
        TCHECK(gdb_q16(ctx, &v, &p, rows.data(), (int64_t)rows.size(), &n, tm));
      });
      printf("TPC-H Q16: %lld groups\n", (long long)n);
      for (int64_t i = 0; i < std::min<int64_t>(n, 5); ++i)
        printf("  Brand#%d %-25s %2d %d\n", rows[i].brand, types[rows[i].type].c_str(),
               rows[i].size, rows[i].supplier_cnt);
      timing_lines("Q16", db.rows[4], db.parse_ms, t);
    } else if (q == "q17") {
      gdb_q17_params p{23, cont_code("MED BOX"), 1, 5};
      gdb_q17_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q17(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q17 Result: avg_yearly = %.2f\n", r.avg_yearly);
      timing_lines("Q17", L, db.parse_ms, t);
    } else if (q == "q18") {
      gdb_q18_params p{300, 100};
      gdb_q18_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q18(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q18: %d of %lld orders over 300\n", r.nrows, (long long)r.orders_over);
      for (int i = 0; i < std::min(r.nrows, 5); ++i)
        printf("  Customer#%09d order %d totalprice %.2f sum_qty %d\n", r.rows[i].custkey,
               r.rows[i].orderkey, r.rows[i].totalprice / 100.0, r.rows[i].sum_qty);
      timing_lines("Q18", L, db.parse_ms, t);
    } else if (q == "q19") {
      gdb_q19_params p{};
      p.nblocks = 3;
      p.mode_mask = (1u << 1) | (1u << 0);  // AIR, REG AIR
      p.instruct = 0;                       // DELIVER IN PERSON
      const char* sz[3] = {"SM", "MED", "LG"};
      const char* kinds[3][4] = {{"CASE", "BOX", "PACK", "PKG"}, {"BAG", "BOX", "PKG", "PACK"},
                                 {"CASE", "BOX", "PACK", "PKG"}};
      const int brand[3] = {12, 23, 34}, qlo[3] = {1, 10, 20}, shi[3] = {5, 10, 15};
      for (int b = 0; b < 3; ++b) {
        uint64_t m = 0;
        for (int k = 0; k < 4; ++k)
          m |= 1ull << cont_code((std::string(sz[b]) + " " + kinds[b][k]).c_str());
        p.blocks[b] = gdb_q19_block{brand[b], qlo[b], qlo[b] + 10, 1, shi[b], 0, m};
      }
      gdb_q19_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q19(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q19 Result: Revenue = %.4f\n", dollars4(r.revenue));
      timing_lines("Q19", L, db.parse_ms, t);
    } else if (q == "q20") {
      gdb_q20_params p{};
      strcpy(p.prefix, "forest");
      p.date_lo = 731;
      p.date_hi = 1096;
      p.nation = nation_key("CANADA");
      p.frac_num = 1;
      p.frac_den = 2;
      std::vector<int32_t> keys(1 << 16);
      int64_t n = 0;
      t = timed([&](gdb_timing* tm) {
        TCHECK(gdb_q20(ctx, &v, &p, keys.data(), (int64_t)keys.size(), &n, tm));
      });
      printf("TPC-H Q20: %lld suppliers\n", (long long)n);
      for (int64_t i = 0; i < std::min<int64_t>(n, 5); ++i) printf("  Supplier#%09d\n", keys[i]);
      timing_lines("Q20", L, db.parse_ms, t);
    } else if (q == "q21") {
      gdb_q21_params p{nation_key("SAUDI ARABIA"), 100, 'F', {0}};
      gdb_q21_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q21(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q21: %d of %lld suppliers\n", r.nrows, (long long)r.suppliers);
      for (int i = 0; i < std::min(r.nrows, 5); ++i)
        printf("  Supplier#%09d numwait %lld\n", r.rows[i].suppkey, (long long)r.rows[i].numwait);
      timing_lines("Q21", L, db.parse_ms, t);
    } else if (q == "q22") {
      gdb_q22_params p{7, {13, 31, 23, 29, 30, 18, 17}};
      gdb_q22_result r;
      t = timed([&](gdb_timing* tm) { TCHECK(gdb_q22(ctx, &v, &p, &r, tm)); });
      printf("TPC-H Q22 Results:\n");
      for (int i = 0; i < r.nrows; ++i)
        printf("  %d numcust %lld totacctbal %.2f\n", r.rows[i].cntrycode,
               (long long)r.rows[i].numcust, r.rows[i].totacctbal / 100.0);
      timing_lines("Q22", db.rows[2], db.parse_ms, t);
    }
    printf("\n");
  }
  for (gdb_table* x : db.t) gdb_table_free(x);
  return 0;
}
