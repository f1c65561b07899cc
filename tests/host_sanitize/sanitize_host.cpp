// Host-code driver for an AddressSanitizer / UndefinedBehaviorSanitizer build
// of the library's CPU-side sources: the seeded table generators (gen.cpp) and
// the .tbl ingest (tbl.cpp). Built with g++ alone (no GPU code), it calls every
// generator at edge sizes and parses valid and malformed .tbl text, checking
// return codes and a few values; any out-of-bounds access, overflow or other
// undefined behaviour is reported by the sanitizers and fails the run.
//
// Usage: sanitize_host <scratch directory>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gdb.h"

// tbl.cpp reports errors through the library's error slot; this stand-in keeps
// the last message so the driver can check that a malformed row is named.
namespace gdb {
static std::string g_last_error;
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace gdb

static int g_failures = 0;

#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      ++g_failures;                                                  \
    }                                                                \
  } while (0)

static std::string write_file(const std::string& dir, const std::string& name,
                              const std::string& text) {
  const std::string path = dir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::perror(path.c_str());
    std::exit(2);
  }
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return path;
}

static void generators() {
  for (uint64_t n : {0ull, 1ull, 63ull, 1000ull, 70001ull}) {
    for (int th : {1, 3, 0}) {
      std::vector<int32_t> sd(n), q(n), d(n), t(n), rd(n);
      std::vector<int64_t> ep(n);
      std::vector<uint8_t> rf(n), ls(n);
      CHECK(gdb_gen_lineitem_flat(7, 5, n, sd.data(), q.data(), d.data(), t.data(), ep.data(),
                                  rf.data(), ls.data(), rd.data(), th) == GDB_OK);
      CHECK(gdb_gen_lineitem_flat(7, 5, n, sd.data(), q.data(), d.data(), t.data(), ep.data(),
                                  rf.data(), ls.data(), nullptr, th) == GDB_OK);
      for (uint64_t i = 0; i < n; ++i) CHECK(q[i] >= 1 && q[i] <= 50 && d[i] >= 0 && d[i] <= 10);

      std::vector<int32_t> ck(n);
      std::vector<uint8_t> seg(n);
      CHECK(gdb_gen_customer(3, n, ck.data(), seg.data(), th) == GDB_OK);

      const uint64_t ncust = n ? n : 1;
      std::vector<int32_t> ok(n), oc(n), od(n), sp(n);
      std::vector<uint8_t> com(n * 80);
      CHECK(gdb_gen_orders(3, 11, n, ncust, ok.data(), oc.data(), od.data(), sp.data(), com.data(),
                           th) == GDB_OK);
      CHECK(gdb_gen_orders(3, 11, n, ncust, ok.data(), oc.data(), od.data(), sp.data(), nullptr,
                           th) == GDB_OK);
      CHECK(gdb_gen_orders(3, 11, n, 0, ok.data(), oc.data(), od.data(), sp.data(), nullptr, th) ==
            GDB_E_INVALID);

      const uint64_t rows = gdb_gen_lineitem_rows(3, 11, n);
      CHECK(rows >= n && rows <= 7 * n);
      std::vector<int32_t> lok(rows), lpk(rows), lsk(rows), lq(rows), ld(rows), lsd(rows);
      std::vector<int64_t> lep(rows);
      CHECK(gdb_gen_lineitem_joined(3, 11, n, 200, 10, lok.data(), lpk.data(), lsk.data(),
                                    lq.data(), ld.data(), lep.data(), lsd.data(), th) == GDB_OK);
      CHECK(gdb_gen_lineitem_joined(3, 11, n, 200, 3, lok.data(), lpk.data(), lsk.data(), lq.data(),
                                    ld.data(), lep.data(), lsd.data(), th) == GDB_E_INVALID);
      for (uint64_t i = 0; i < rows; ++i) CHECK(lpk[i] >= 1 && lpk[i] <= 200 && lsk[i] >= 1);

      std::vector<int32_t> pk(n);
      std::vector<uint8_t> pname(n * 55);
      CHECK(gdb_gen_part(3, n, pk.data(), pname.data(), th) == GDB_OK);

      std::vector<int32_t> ppk(4 * n), psk(4 * n);
      std::vector<int64_t> pcost(4 * n);
      CHECK(gdb_gen_partsupp(3, n, 10, ppk.data(), psk.data(), pcost.data(), th) == GDB_OK);
      CHECK(gdb_gen_partsupp(3, n, 2, ppk.data(), psk.data(), pcost.data(), th) == GDB_E_INVALID);

      std::vector<int32_t> sk(n), nk(n);
      CHECK(gdb_gen_supplier(3, n, sk.data(), nk.data(), th) == GDB_OK);
      for (uint64_t i = 0; i < n; ++i) CHECK(nk[i] >= 0 && nk[i] < 25);
    }
  }
}

struct Cols {
  std::vector<int32_t> key, date, disc;
  std::vector<int64_t> price;
  std::vector<uint8_t> flag, text;
  explicit Cols(uint64_t n) : key(n), date(n), disc(n), price(n), flag(n), text(n * 8) {}
  std::vector<gdb_tbl_field> fields() {
    return {{0, GDB_TBL_INT32, 0, key.data()},     {1, GDB_TBL_DATE, 0, date.data()},
            {2, GDB_TBL_HUNDREDTHS, 0, disc.data()}, {3, GDB_TBL_CENTS, 0, price.data()},
            {4, GDB_TBL_CHAR, 0, flag.data()},      {5, GDB_TBL_TEXT, 8, text.data()}};
  }
};

static int parse(const std::string& path, uint64_t* rows_out, Cols** cols_out, int threads) {
  uint64_t rows = 0;
  int rc = gdb_tbl_count_rows(path.c_str(), threads, &rows);
  if (rc != GDB_OK) return rc;
  *rows_out = rows;
  *cols_out = new Cols(rows);
  std::vector<gdb_tbl_field> f = (*cols_out)->fields();
  return gdb_tbl_parse(path.c_str(), f.data(), (int)f.size(), rows, threads);
}

static void tbl(const std::string& dir) {
  // Valid rows, a blank line, and a last line without its newline.
  const std::string good =
      "1|1992-01-01|0.06|21168.23|N|abc|\n"
      "-2147483648|1998-12-01|0.10|-0.5|R|a string longer than eight|\n"
      "\n"
      "2147483647|2000-02-29|.07|17|A||\n"
      "17.00|1995-06-17|0|0.01|F|x|";
  {
    uint64_t rows = 0;
    Cols* c = nullptr;
    for (int th : {1, 2, 8}) {
      CHECK(parse(write_file(dir, "good.tbl", good), &rows, &c, th) == GDB_OK);
      CHECK(rows == 4);
      CHECK(c->key[0] == 1 && c->key[1] == INT32_MIN && c->key[2] == INT32_MAX && c->key[3] == 17);
      CHECK(c->date[0] == 0 && c->date[1] == 2526);
      CHECK(c->disc[0] == 6 && c->disc[2] == 7 && c->disc[3] == 0);
      CHECK(c->price[0] == 2116823 && c->price[1] == -50 && c->price[2] == 1700 && c->price[3] == 1);
      CHECK(c->flag[1] == 'R' && c->flag[2] == 'A');
      CHECK(std::memcmp(&c->text[8], "a string", 8) == 0 && c->text[16] == 0);
      delete c;
      c = nullptr;
    }
  }
  // A file large enough to be split across threads, with CRLF-free long lines.
  {
    std::string big;
    for (int i = 0; i < 60000; ++i)
      big += std::to_string(i) + "|1994-03-0" + std::to_string(1 + i % 9) + "|0.0" +
             std::to_string(i % 10) + "|" + std::to_string(i) + "." + std::to_string(10 + i % 90) +
             "|O|" + std::string(i % 37, 'z') + "|\n";
    uint64_t rows = 0;
    Cols* c = nullptr;
    CHECK(parse(write_file(dir, "big.tbl", big), &rows, &c, 8) == GDB_OK);
    CHECK(rows == 60000);
    for (uint64_t i = 0; i < rows; ++i)
      CHECK(c->key[i] == (int32_t)i && c->price[i] == (int64_t)i * 100 + 10 + (int64_t)(i % 90));
    delete c;
  }
  // Malformed rows: each must fail with GDB_E_INVALID, never read past the line.
  const char* bad[] = {
      "1|1992-01-01|0.06|1.00|N\n",                  // too few fields
      "x|1992-01-01|0.06|1.00|N|t|\n",                // not an integer
      "2147483648|1992-01-01|0.06|1.00|N|t|\n",       // past INT32_MAX
      "-2147483649|1992-01-01|0.06|1.00|N|t|\n",      // past INT32_MIN
      "99999999999999999999999|1992-01-01|0|1|N|t|\n", // many digits
      "1.5|1992-01-01|0.06|1.00|N|t|\n",              // fractional integer
      "1|1992-13-01|0.06|1.00|N|t|\n",                // month 13
      "1|1992-01|0.06|1.00|N|t|\n",                   // short date
      "1|199999999999999999-01-01|0|1|N|t|\n",        // long year
      "1|1992-01-0199999999999999|0|1|N|t|\n",        // long day
      "1|1992-01-01|0.065|1.00|N|t|\n",               // three significant fraction digits
      "1|1992-01-01|0.06|9999999999999999999999.99|N|t|\n",  // past int64
      "1|1992-01-01|0.06|1.2.3|N|t|\n",              // two points
      "1|1992-01-01|0.06|-|N|t|\n",                   // sign only
      "1|1992-01-01|99999999999.99|1|N|t|\n",         // hundredths past int32
      "1|1992-01-01|0.06||N|t|\n",                    // empty decimal
  };
  const std::string first_line = good.substr(0, good.find('\n') + 1);
  for (const char* b : bad) {
    for (int th : {1, 4}) {
      uint64_t rows = 0;
      Cols* c = nullptr;
      gdb::g_last_error.clear();
      const int rc = parse(write_file(dir, "bad.tbl", first_line + b), &rows, &c, th);
      if (rc != GDB_E_INVALID) std::fprintf(stderr, "accepted: %s", b);
      CHECK(rc == GDB_E_INVALID);
      CHECK(gdb::g_last_error.find("row 2") != std::string::npos);
      delete c;
    }
  }
  // Row-count mismatch, bad field descriptions, a missing file, an empty file.
  {
    const std::string path = write_file(dir, "good.tbl", good);
    Cols c(8);
    std::vector<gdb_tbl_field> f = c.fields();
    CHECK(gdb_tbl_parse(path.c_str(), f.data(), (int)f.size(), 3, 1) == GDB_E_INVALID);
    CHECK(gdb_tbl_parse(path.c_str(), f.data(), (int)f.size(), 5, 1) == GDB_E_INVALID);
    gdb_tbl_field bad_kind = {0, 99, 0, c.key.data()};
    CHECK(gdb_tbl_parse(path.c_str(), &bad_kind, 1, 4, 1) == GDB_E_INVALID);
    gdb_tbl_field no_width = {5, GDB_TBL_TEXT, 0, c.text.data()};
    CHECK(gdb_tbl_parse(path.c_str(), &no_width, 1, 4, 1) == GDB_E_INVALID);
    gdb_tbl_field no_out = {0, GDB_TBL_INT32, 0, nullptr};
    CHECK(gdb_tbl_parse(path.c_str(), &no_out, 1, 4, 1) == GDB_E_INVALID);
    CHECK(gdb_tbl_parse(path.c_str(), nullptr, 0, 4, 1) == GDB_OK);
    uint64_t rows = 7;
    CHECK(gdb_tbl_count_rows((dir + "/missing.tbl").c_str(), 1, &rows) == GDB_E_INVALID);
    CHECK(gdb_tbl_count_rows(write_file(dir, "empty.tbl", "").c_str(), 4, &rows) == GDB_OK);
    CHECK(rows == 0);
    CHECK(gdb_tbl_count_rows(write_file(dir, "blank.tbl", "\n\n\n").c_str(), 4, &rows) == GDB_OK);
    CHECK(rows == 0);
    CHECK(gdb_tbl_count_rows(path.c_str(), 1, nullptr) == GDB_E_INVALID);
  }
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
    return 2;
  }
  generators();
  tbl(argv[1]);
  if (g_failures) {
    std::fprintf(stderr, "%d checks failed\n", g_failures);
    return 1;
  }
  std::printf("host sanitize: all checks passed\n");
  return 0;
}
