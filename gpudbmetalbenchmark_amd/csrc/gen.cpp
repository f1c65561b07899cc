// Host-side seeded table generator behind the gdb_gen_* entry points. Each
// entry point splits its row range over `nthreads` std::threads; because
// every cell is a pure function of (seed, stream, key) the split does not
// change the output.
#include <algorithm>
#include <thread>
#include <vector>

#include "gdb.h"
#include "gen_core.h"

using namespace gdbgen;

namespace {

template <class F>
void parallel_rows(uint64_t n, int nthreads, F&& body) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  uint64_t per = (n + nthreads - 1) / (uint64_t)nthreads;
  if (n < 65536 || nthreads == 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    uint64_t lo = (uint64_t)t * per, hi = std::min(n, lo + per);
    if (lo >= hi) break;
    th.emplace_back([=, &body] { body(lo, hi); });
  }
  for (auto& x : th) x.join();
}

const int32_t YEAR_START[] = {0, 366, 731, 1096, 1461, 1827, 2192, 2557, 2922, 3288, 3653};

}  // namespace

extern "C" {

int32_t gdb_year_of_day(int32_t day) {
  int y = 0;
  while (y + 1 < (int)(sizeof(YEAR_START) / sizeof(YEAR_START[0])) && day >= YEAR_START[y + 1]) ++y;
  return 1992 + y;
}

const char* gdb_nation_name(int k) {
  static const char* const N[25] = {
      "ALGERIA", "ARGENTINA", "BRAZIL", "CANADA", "EGYPT", "ETHIOPIA", "FRANCE",
      "GERMANY", "INDIA", "INDONESIA", "IRAN", "IRAQ", "JAPAN", "JORDAN", "KENYA",
      "MOROCCO", "MOZAMBIQUE", "PERU", "CHINA", "ROMANIA", "SAUDI ARABIA", "VIETNAM",
      "RUSSIA", "UNITED KINGDOM", "UNITED STATES"};
  return (k >= 0 && k < 25) ? N[k] : "";
}

const char* gdb_segment_name(int c) {
  static const char* const S[5] = {"AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD",
                                   "MACHINERY"};
  return (c >= 0 && c < 5) ? S[c] : "";
}

int gdb_gen_lineitem_flat(uint64_t seed, uint64_t row0, uint64_t n, int32_t* shipdate,
                          int32_t* quantity, int32_t* discount, int32_t* tax, int64_t* extprice,
                          uint8_t* returnflag, uint8_t* linestatus, int32_t* receiptdate,
                          int nthreads) {
  const uint64_t bs = stream_base(seed, LI_SHIP), bq = stream_base(seed, LI_QTY),
                 bd = stream_base(seed, LI_DISC), br = stream_base(seed, LI_RETAIL),
                 bt = stream_base(seed, LI_TAX), bc = stream_base(seed, LI_RECEIPT),
                 bf = stream_base(seed, LI_RFLAG);
  parallel_rows(n, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t j = lo; j < hi; ++j) {
      const uint64_t r = row0 + j;
      int32_t s = (int32_t)uni(draw(bs, r), 0, SHIP_MAX);
      int32_t q = (int32_t)uni(draw(bq, r), 1, 50);
      if (shipdate) shipdate[j] = s;
      if (quantity) quantity[j] = q;
      if (discount) discount[j] = (int32_t)uni(draw(bd, r), 0, 10);
      if (tax) tax[j] = (int32_t)uni(draw(bt, r), 0, 8);
      if (extprice) extprice[j] = (int64_t)q * uni(draw(br, r), RETAIL_LO, RETAIL_HI);
      int32_t rc = s + (int32_t)uni(draw(bc, r), 1, 30);
      if (receiptdate) receiptdate[j] = rc;
      if (linestatus) linestatus[j] = (uint8_t)(s < STATUS_DAY ? 'F' : 'O');
      if (returnflag)
        returnflag[j] = (uint8_t)(rc < STATUS_DAY ? (uni(draw(bf, r), 0, 1) ? 'R' : 'A') : 'N');
    }
  });
  return GDB_OK;
}

int gdb_gen_customer(uint64_t seed, uint64_t n, int32_t* custkey, uint8_t* segment,
                     int nthreads) {
  const uint64_t b = stream_base(seed, C_SEG);
  parallel_rows(n, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      if (custkey) custkey[i] = (int32_t)(i + 1);
      if (segment) segment[i] = (uint8_t)uni(draw(b, i), 0, NSEGMENTS - 1);
    }
  });
  return GDB_OK;
}

int gdb_gen_orders(uint64_t seed, uint64_t order0, uint64_t n, uint64_t ncustomers,
                   int32_t* orderkey, int32_t* custkey, int32_t* orderdate,
                   int32_t* shippriority, uint8_t* comment, int nthreads) {
  if (ncustomers == 0) return GDB_E_INVALID;
  const uint64_t bc = stream_base(seed, O_CUST);
  parallel_rows(n, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t j = lo; j < hi; ++j) {
      const uint64_t i = order0 + j;
      if (orderkey) orderkey[j] = orderkey_of(i);
      if (custkey) custkey[j] = (int32_t)uni(draw(bc, i), 1, (int64_t)ncustomers);
      if (orderdate) orderdate[j] = orderdate_of(seed, i);
      if (shippriority) shippriority[j] = 0;
      if (comment) order_comment(seed, i, comment + j * COMMENT_W);
    }
  });
  return GDB_OK;
}

uint64_t gdb_gen_lineitem_rows(uint64_t seed, uint64_t order0, uint64_t n) {
  uint64_t total = 0;
  for (uint64_t j = 0; j < n; ++j) total += (uint64_t)lines_of_order(seed, order0 + j);
  return total;
}

int gdb_gen_lineitem_joined(uint64_t seed, uint64_t order0, uint64_t n, uint64_t nparts,
                            uint64_t nsuppliers, int32_t* orderkey, int32_t* partkey,
                            int32_t* suppkey, int32_t* quantity, int32_t* discount,
                            int64_t* extprice, int32_t* shipdate, int nthreads) {
  if (nparts == 0 || nsuppliers < 4) return GDB_E_INVALID;
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  // Row offset of every order: prefix sum of the per-order line counts.
  std::vector<uint64_t> off(n + 1, 0);
  {
    int T = (n < 65536) ? 1 : nthreads;
    uint64_t per = (n + T - 1) / (uint64_t)T;
    std::vector<uint64_t> part_sum(T, 0);
    auto count = [&](int t) {
      uint64_t lo = (uint64_t)t * per, hi = std::min(n, lo + per), s = 0;
      for (uint64_t j = lo; j < hi; ++j) {
        off[j + 1] = (uint64_t)lines_of_order(seed, order0 + j);
        s += off[j + 1];
      }
      part_sum[t] = s;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(count, t);
    for (auto& x : th) x.join();
    uint64_t base = 0;
    for (int t = 0; t < T; ++t) {
      uint64_t lo = (uint64_t)t * per, hi = std::min(n, lo + per);
      uint64_t run = base;
      for (uint64_t j = lo; j < hi; ++j) {
        run += off[j + 1];
        off[j + 1] = run;
      }
      base += part_sum[t];
    }
  }
  const uint64_t bs = stream_base(seed, J_SHIP), bq = stream_base(seed, J_QTY),
                 bd = stream_base(seed, J_DISC), br = stream_base(seed, J_RETAIL),
                 bp = stream_base(seed, J_PART), bx = stream_base(seed, J_SUPPIDX);
  parallel_rows(n, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t j = lo; j < hi; ++j) {
      const uint64_t i = order0 + j;
      const int32_t ok = orderkey_of(i), od = orderdate_of(seed, i);
      for (uint64_t r = off[j], l = 0; r < off[j + 1]; ++r, ++l) {
        const uint64_t key = i * 8 + l;
        int32_t q = (int32_t)uni(draw(bq, key), 1, 50);
        int64_t p = uni(draw(bp, key), 1, (int64_t)nparts);
        if (orderkey) orderkey[r] = ok;
        if (partkey) partkey[r] = (int32_t)p;
        if (suppkey) suppkey[r] = supplier_of(p, (int)uni(draw(bx, key), 0, 3), (int64_t)nsuppliers);
        if (quantity) quantity[r] = q;
        if (discount) discount[r] = (int32_t)uni(draw(bd, key), 0, 10);
        if (extprice) extprice[r] = (int64_t)q * uni(draw(br, key), RETAIL_LO, RETAIL_HI);
        if (shipdate) shipdate[r] = od + (int32_t)uni(draw(bs, key), 1, 121);
      }
    }
  });
  return GDB_OK;
}

int gdb_gen_part(uint64_t seed, uint64_t n, int32_t* partkey, uint8_t* name, int nthreads) {
  parallel_rows(n, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      if (partkey) partkey[i] = (int32_t)(i + 1);
      if (name) part_name(seed, i, name + i * PNAME_W);
    }
  });
  return GDB_OK;
}

int gdb_gen_partsupp(uint64_t seed, uint64_t nparts, uint64_t nsuppliers, int32_t* partkey,
                     int32_t* suppkey, int64_t* supplycost, int nthreads) {
  if (nsuppliers < 4) return GDB_E_INVALID;
  const uint64_t b = stream_base(seed, PS_COST);
  parallel_rows(nparts * 4, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t r = lo; r < hi; ++r) {
      int64_t p = (int64_t)(r / 4) + 1;
      if (partkey) partkey[r] = (int32_t)p;
      if (suppkey) suppkey[r] = supplier_of(p, (int)(r % 4), (int64_t)nsuppliers);
      if (supplycost) supplycost[r] = uni(draw(b, r), 100, 100000);
    }
  });
  return GDB_OK;
}

int gdb_gen_supplier(uint64_t seed, uint64_t n, int32_t* suppkey, int32_t* nationkey,
                     int nthreads) {
  const uint64_t b = stream_base(seed, S_NATION);
  parallel_rows(n, nthreads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      if (suppkey) suppkey[i] = (int32_t)(i + 1);
      if (nationkey) nationkey[i] = (int32_t)uni(draw(b, i), 0, NNATIONS - 1);
    }
  });
  return GDB_OK;
}

}  // extern "C"
