// TPC-H `.tbl` ingest: pipe-delimited text rows into the engine's exact
// columns (replaces the reference's MappedFile / loadColumnsMulti and the
// parse*ColumnChunk family, src/infra.h:59-171, and the SF100 chunk parsers,
// src/infra.h:262-361).
//
// The reference parses decimals with strtof into float32 and dates into
// YYYYMMDD integers (src/infra.h:137-156). Here a decimal becomes an exact
// int64 in hundredths ("21168.23" -> 2116823) and a date a day number from
// 1992-01-01, so every sum downstream stays exact. The file is memory-mapped
// once; a first pass counts the lines of each thread's byte range, a second
// parses the ranges in parallel, each thread writing its own rows.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gdb.h"
#include "internal.h"

using namespace gdb;

namespace {

struct Mapped {
  int fd = -1;
  const char* data = nullptr;
  size_t size = 0;
  ~Mapped() {
    if (data && size) munmap(const_cast<char*>(data), size);
    if (fd >= 0) close(fd);
  }
};

int map_file(const char* path, Mapped* m) {
  m->fd = open(path, O_RDONLY);
  if (m->fd < 0) return fail(GDB_E_INVALID, std::string("cannot open ") + path);
  struct stat st;
  if (fstat(m->fd, &st) != 0) return fail(GDB_E_INVALID, std::string("cannot stat ") + path);
  m->size = (size_t)st.st_size;
  if (m->size == 0) return GDB_OK;
  void* p = mmap(nullptr, m->size, PROT_READ, MAP_PRIVATE, m->fd, 0);
  if (p == MAP_FAILED) {
    m->size = 0;
    return fail(GDB_E_INVALID, std::string("cannot map ") + path);
  }
  madvise(p, m->size, MADV_SEQUENTIAL);
  m->data = static_cast<const char*>(p);
  return GDB_OK;
}

int threads_for(int nthreads, size_t bytes) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)nthreads, bytes / (1 << 20) + 1));
}

// Byte ranges [b[t], b[t+1]) that start at line starts.
std::vector<size_t> split_lines(const Mapped& m, int parts) {
  std::vector<size_t> b(parts + 1, m.size);
  b[0] = 0;
  for (int t = 1; t < parts; ++t) {
    size_t p = m.size / parts * t;
    while (p < m.size && m.data[p - 1] != '\n') ++p;
    b[t] = std::max(p, b[t - 1]);
  }
  return b;
}

// A row is a line with at least one byte before its newline (a last line
// without a newline counts too).
uint64_t count_range(const Mapped& m, size_t lo, size_t hi) {
  uint64_t n = 0;
  size_t start = lo;
  for (size_t i = lo; i < hi; ++i) {
    if (m.data[i] == '\n') {
      n += i > start ? 1 : 0;
      start = i + 1;
    }
  }
  if (hi == m.size && hi > start) ++n;
  return n;
}

bool parse_int(const char* s, const char* e, int32_t* out) {
  bool neg = false;
  if (s < e && (*s == '-' || *s == '+')) neg = *s++ == '-';
  if (s == e) return false;
  int64_t v = 0;
  for (; s < e && *s >= '0' && *s <= '9'; ++s) {
    v = v * 10 + (*s - '0');
    if (v > 2147483648ll) return false;
  }
  if (s < e) {  // an integer-valued decimal such as "17.00"
    if (*s != '.') return false;
    for (++s; s < e; ++s)
      if (*s != '0') return false;
  }
  v = neg ? -v : v;
  if (v < INT32_MIN || v > INT32_MAX) return false;
  *out = (int32_t)v;
  return true;
}

// Decimal with at most two fraction digits -> hundredths, exact.
bool parse_hundredths(const char* s, const char* e, int64_t* out) {
  bool neg = false;
  if (s < e && (*s == '-' || *s == '+')) neg = *s++ == '-';
  if (s == e) return false;
  int64_t v = 0;
  bool digits = false;
  for (; s < e && *s >= '0' && *s <= '9'; ++s) {
    if (v > (INT64_MAX - 9) / 1000) return false;
    v = v * 10 + (*s - '0');
    digits = true;
  }
  int frac = 0;
  if (s < e) {
    if (*s != '.') return false;
    for (++s; s < e; ++s) {
      if (*s < '0' || *s > '9') return false;
      if (frac == 2) {
        if (*s != '0') return false;  // more than two significant fraction digits
        continue;
      }
      v = v * 10 + (*s - '0');
      ++frac;
      digits = true;
    }
  }
  if (!digits) return false;
  for (; frac < 2; ++frac) v *= 10;
  *out = neg ? -v : v;
  return true;
}

// Days from 1970-01-01 of a proleptic Gregorian date.
int64_t days_from_civil(int64_t y, int m, int d) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const int64_t yoe = y - era * 400;
  const int64_t doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}

bool parse_date(const char* s, const char* e, int32_t* out) {
  int v[3] = {0, 0, 0};
  int k = 0;
  const char* p = s;
  for (; k < 3; ++k) {
    const char* q = p;
    while (p < e && *p >= '0' && *p <= '9' && p - q < 5) v[k] = v[k] * 10 + (*p++ - '0');
    if (p == q || p - q > 4) return false;  // at most four digits per part
    if (k < 2) {
      if (p == e || *p != '-') return false;
      ++p;
    }
  }
  if (p != e || v[1] < 1 || v[1] > 12 || v[2] < 1 || v[2] > 31) return false;
  *out = (int32_t)(days_from_civil(v[0], v[1], v[2]) - days_from_civil(1992, 1, 1));
  return true;
}

struct Field {
  gdb_tbl_field f;
  int order;  // position in the caller's array
};

// Parses the rows of [lo, hi), the first being row `row0`.
int parse_range(const Mapped& m, size_t lo, size_t hi, uint64_t row0, uint64_t rows,
                const std::vector<Field>& fields, int max_field, std::string* err) {
  const char* p = m.data + lo;
  const char* end = m.data + hi;
  uint64_t r = row0;
  while (p < end) {
    const char* eol = static_cast<const char*>(memchr(p, '\n', (size_t)(end - p)));
    if (!eol) eol = end;
    if (eol > p) {
      if (r >= rows) {
        *err = "more rows than counted";
        return GDB_E_INVALID;
      }
      const char* s = p;
      int idx = 0;
      size_t k = 0;
      while (k < fields.size() && idx <= max_field) {
        const char* bar = static_cast<const char*>(memchr(s, '|', (size_t)(eol - s)));
        const char* fe = bar ? bar : eol;
        while (k < fields.size() && fields[k].f.field == idx) {
          const gdb_tbl_field& f = fields[k].f;
          bool ok = true;
          switch (f.kind) {
            case GDB_TBL_INT32: ok = parse_int(s, fe, static_cast<int32_t*>(f.out) + r); break;
            case GDB_TBL_DATE: ok = parse_date(s, fe, static_cast<int32_t*>(f.out) + r); break;
            case GDB_TBL_CENTS: ok = parse_hundredths(s, fe, static_cast<int64_t*>(f.out) + r); break;
            case GDB_TBL_HUNDREDTHS: {
              int64_t v = 0;
              ok = parse_hundredths(s, fe, &v) && v >= INT32_MIN && v <= INT32_MAX;
              static_cast<int32_t*>(f.out)[r] = (int32_t)v;
              break;
            }
            case GDB_TBL_CHAR: static_cast<uint8_t*>(f.out)[r] = fe > s ? (uint8_t)*s : 0; break;
            case GDB_TBL_TEXT: {
              uint8_t* dst = static_cast<uint8_t*>(f.out) + r * (uint64_t)f.width;
              const size_t len = std::min<size_t>((size_t)(fe - s), (size_t)f.width);
              memcpy(dst, s, len);
              memset(dst + len, 0, (size_t)f.width - len);
              break;
            }
            default: ok = false;
          }
          if (!ok) {
            *err = "row " + std::to_string(r + 1) + " field " + std::to_string(idx) + ": cannot parse '" +
                   std::string(s, std::min<size_t>((size_t)(fe - s), 40)) + "'";
            return GDB_E_INVALID;
          }
          ++k;
        }
        if (!bar) break;
        s = bar + 1;
        ++idx;
      }
      if (k < fields.size()) {
        *err = "row " + std::to_string(r + 1) + " has too few fields";
        return GDB_E_INVALID;
      }
      ++r;
    }
    p = eol + 1;
  }
  return GDB_OK;
}

}  // namespace

extern "C" {

int gdb_tbl_count_rows(const char* path, int nthreads, uint64_t* rows) {
  if (!path || !rows) return fail(GDB_E_INVALID, "null argument");
  Mapped m;
  GDB_TRY(map_file(path, &m));
  const int T = threads_for(nthreads, m.size);
  const std::vector<size_t> b = split_lines(m, T);
  std::vector<uint64_t> n(T, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) th.emplace_back([&, t] { n[t] = count_range(m, b[t], b[t + 1]); });
  for (auto& x : th) x.join();
  uint64_t s = 0;
  for (uint64_t v : n) s += v;
  *rows = s;
  return GDB_OK;
}

int gdb_tbl_parse(const char* path, const gdb_tbl_field* fields, int nfields, uint64_t rows,
                  int nthreads) {
  if (!path || (!fields && nfields > 0) || nfields < 0) return fail(GDB_E_INVALID, "null argument");
  std::vector<Field> fs;
  int max_field = -1;
  for (int i = 0; i < nfields; ++i) {
    const gdb_tbl_field& f = fields[i];
    if (!f.out || f.field < 0 || f.kind < GDB_TBL_INT32 || f.kind > GDB_TBL_TEXT ||
        (f.kind == GDB_TBL_TEXT && f.width <= 0))
      return fail(GDB_E_INVALID, "bad field description " + std::to_string(i));
    fs.push_back(Field{f, i});
    max_field = std::max(max_field, f.field);
  }
  std::stable_sort(fs.begin(), fs.end(),
                   [](const Field& a, const Field& b) { return a.f.field < b.f.field; });
  Mapped m;
  GDB_TRY(map_file(path, &m));
  const int T = threads_for(nthreads, m.size);
  const std::vector<size_t> b = split_lines(m, T);
  std::vector<uint64_t> n(T, 0);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back([&, t] { n[t] = count_range(m, b[t], b[t + 1]); });
    for (auto& x : th) x.join();
  }
  uint64_t total = 0;
  std::vector<uint64_t> row0(T, 0);
  for (int t = 0; t < T; ++t) {
    row0[t] = total;
    total += n[t];
  }
  if (total != rows)
    return fail(GDB_E_INVALID, std::string(path) + " has " + std::to_string(total) + " rows, " +
                                   std::to_string(rows) + " expected");
  std::vector<int> rc(T, GDB_OK);
  std::vector<std::string> err(T);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, t] { rc[t] = parse_range(m, b[t], b[t + 1], row0[t], rows, fs, max_field, &err[t]); });
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < T; ++t)
    if (rc[t] != GDB_OK) return fail(rc[t], std::string(path) + ": " + err[t]);
  return GDB_OK;
}

}  // extern "C"
