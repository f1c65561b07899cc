// TPC-H pipelines built around one fact scan: Q14, Q19, Q12, Q4, Q15, Q17, Q22.
//
// Each replaces the reference's runQNBenchmark host function and its kernels
// (reference src/qN.cpp, kernels/qN.metal, cited per pipeline). A dimension
// table becomes a bitmap or byte map over its dense key; the fact table is
// streamed once with dwordx4 loads, the most selective column first.
#include "tpch_common.h"

using namespace gdb;
using namespace gdbk;
using namespace gdbt;

namespace {

// ---- Q14 (reference kernels/q14.metal:17-95, src/q14.cpp:8-120) ----------------

__global__ __launch_bounds__(256) void q14_scan(const int32_t* __restrict__ ship,
                                                const int32_t* __restrict__ pk,
                                                const int32_t* __restrict__ disc,
                                                const int64_t* __restrict__ price, uint32_t n,
                                                int32_t lo, int32_t hi,
                                                const uint32_t* __restrict__ promo,
                                                int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t red[4];
  int64_t sp = 0, st = 0;
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ship, g); }, [&](uint32_t g, const int4& s) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = lane_of(s, j);
      m |= (uint32_t)((4 * g + j < n) & (v >= lo) & (v < hi)) << j;
    }
    if (!m) return;
    const int4 p = ld4nt(pk, g), d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
    const uint32_t pm = bits4(promo, p);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const int64_t rev = lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j));
      st += rev;
      if ((pm >> j) & 1) sp += rev;
    }
  });
  block_add(sp, red, out);
  block_add(st, red, out + 1);
}

// ---- Q19 (reference kernels/q19.metal:22-160, src/q19.cpp) ---------------------

struct Q19Part {
  int32_t nblocks;
  int32_t brand[3], size_lo[3], size_hi[3];
  uint64_t cont[3];
};
struct Q19Line {
  int32_t nblocks;
  uint32_t modes;
  int32_t instruct;
  int32_t qlo[3], qhi[3];
};

// map[partkey] = bit b set when the part satisfies block b's part predicates.
__global__ __launch_bounds__(256) void q19_part(const int32_t* __restrict__ pk,
                                                const uint8_t* __restrict__ brand,
                                                const uint8_t* __restrict__ cont,
                                                const int32_t* __restrict__ size, uint32_t n,
                                                Q19Part q, uint8_t* __restrict__ map) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (int b = 0; b < q.nblocks; ++b)
      bits |= (uint32_t)((brand[i] == q.brand[b]) & ((q.cont[b] >> (cont[i] & 63)) & 1) &
                         (cont[i] < 64) & (size[i] >= q.size_lo[b]) & (size[i] <= q.size_hi[b]))
              << b;
    map[pk[i]] = (uint8_t)bits;
  }
}

__global__ __launch_bounds__(256) void q19_scan(const uint8_t* __restrict__ mode,
                                                const uint8_t* __restrict__ instr,
                                                const int32_t* __restrict__ pk,
                                                const int32_t* __restrict__ qty,
                                                const int32_t* __restrict__ disc,
                                                const int64_t* __restrict__ price, uint32_t n,
                                                Q19Line q, const uint8_t* __restrict__ map,
                                                int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t red[4];
  int64_t sum = 0;
  walk2(quads(n), [&](uint32_t g) { return make_uint2(ld_u8x4nt(mode, g), ld_u8x4nt(instr, g)); },
        [&](uint32_t g, const uint2& w) {
    const uint32_t mw = w.x, iw = w.y;
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      m |= (uint32_t)((4 * g + j < n) & ((q.modes >> (byte_of(mw, j) & 31)) & 1) &
                      (byte_of(mw, j) < 32) & (byte_of(iw, j) == q.instruct))
           << j;
    if (!m) return;
    const int4 p = ld4nt(pk, g), qq = ld4nt(qty, g);
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const uint32_t pb = map[lane_of(p, j)];
      const int qv = lane_of(qq, j);
      bool hit = false;
      for (int b = 0; b < q.nblocks; ++b)
        hit |= ((pb >> b) & 1) && qv >= q.qlo[b] && qv <= q.qhi[b];
      m2 |= (uint32_t)hit << j;
    }
    if (!m2) return;
    const int4 d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((m2 >> j) & 1) sum += lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j));
  });
  block_add(sum, red, out);
}

// ---- Q12 (reference kernels/q12.metal:22-100, src/q12.cpp) ---------------------

// high[orderkey] = priority in mask, plain byte stores (no atomics).
__global__ __launch_bounds__(256) void order_prio_bytes(const int32_t* __restrict__ ok,
                                                        const uint8_t* __restrict__ prio,
                                                        uint32_t n, uint32_t mask,
                                                        uint8_t* __restrict__ high) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    high[ok[i]] = prio[i] < 32 && ((mask >> prio[i]) & 1);
}

__global__ __launch_bounds__(256) void q12_scan(const uint8_t* __restrict__ mode,
                                                const int32_t* __restrict__ receipt,
                                                const int32_t* __restrict__ commit,
                                                const int32_t* __restrict__ ship,
                                                const int32_t* __restrict__ ok, uint32_t n,
                                                int32_t m0, int32_t m1, int32_t lo, int32_t hi,
                                                const uint8_t* __restrict__ high,
                                                int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t red[4];
  int64_t c[4] = {0, 0, 0, 0};  // [mode index][high]
  constexpr int QB = 4;
  walk_batch<QB>(quads(n), [&](const uint32_t* g, const bool* live) {
    uint32_t mw[QB], m[QB];
    int4 r[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      mw[b] = live[b] ? ld_u8x4nt(mode, g[b]) : 0u;
      r[b] = live[b] ? ld4nt(receipt, g[b]) : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      m[b] = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int md = byte_of(mw[b], j), rv = lane_of(r[b], j);
        m[b] |= (uint32_t)(live[b] && (4 * g[b] + j < n) && ((md == m0) | (md == m1)) &&
                           (rv >= lo) && (rv < hi)) << j;
      }
    }
    int4 cm[QB], s[QB], o[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (m[b]) {
        cm[b] = ld4nt(commit, g[b]);
        s[b] = ld4nt(ship, g[b]);
        o[b] = ld4nt(ok, g[b]);
      }
    uint32_t hv[QB][4];
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (m[b]) gather4(high, o[b], hv[b]);
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      if (!m[b]) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int cv = lane_of(cm[b], j);
        if (!((m[b] >> j) & 1) || !(cv < lane_of(r[b], j)) || !(lane_of(s[b], j) < cv)) continue;
        const int mi = byte_of(mw[b], j) == m0 ? 0 : 1;
        c[mi * 2 + (hv[b][j] ? 1 : 0)] += 1;
      }
    }
  });
#pragma unroll
  for (int i = 0; i < 4; ++i) block_add(c[i], red, out + i);
}

// ---- Q4 (reference kernels/q4.metal:21-90, src/q4.cpp) -------------------------

// late[orderkey] = 1 when a line of the order has commitdate < receiptdate.
__global__ __launch_bounds__(256) void q4_late(const int32_t* __restrict__ ok,
                                               const int32_t* __restrict__ commit,
                                               const int32_t* __restrict__ receipt, uint32_t n,
                                               uint8_t* __restrict__ late) {
  struct Ld {
    int4 c, r, o;
  };
  walk2(quads(n), [&](uint32_t g) { return Ld{ld4nt(commit, g), ld4nt(receipt, g), ld4nt(ok, g)}; },
        [&](uint32_t g, const Ld& ld) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (4 * g + j < n && lane_of(ld.c, j) < lane_of(ld.r, j)) late[lane_of(ld.o, j)] = 1;
  });
}

__global__ __launch_bounds__(256) void q4_count(const int32_t* __restrict__ date,
                                                const int32_t* __restrict__ ok,
                                                const uint8_t* __restrict__ prio, uint32_t n,
                                                int32_t lo, int32_t hi,
                                                const uint8_t* __restrict__ late,
                                                int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t red[4];
  int64_t c[5] = {0, 0, 0, 0, 0};
  // Four groups per lane per trip; date, key and priority load together for all
  // of them, then the late flags of the window's orders (one dependent phase).
  constexpr int QB = 4;
  walk_batch<QB>(quads(n), [&](const uint32_t* g, const bool* live) {
    int4 d[QB], o[QB];
    uint32_t pw[QB], m[QB], lt[QB][4];
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (live[b]) {
        d[b] = ld4nt(date, g[b]);
        o[b] = ld4nt(ok, g[b]);
        pw[b] = ld_u8x4nt(prio, g[b]);
      }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      m[b] = 0;
      if (!live[b]) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int v = lane_of(d[b], j);
        m[b] |= (uint32_t)((4 * g[b] + j < n) & (v >= lo) & (v < hi)) << j;
      }
      if (m[b]) gather4(late, o[b], lt[b]);
    }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      if (!m[b]) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (!((m[b] >> j) & 1) || !lt[b][j]) continue;
        const uint32_t pr = byte_of(pw[b], j);
#pragma unroll
        for (int q = 0; q < 5; ++q) c[q] += pr == (uint32_t)q;
      }
    }
  });
#pragma unroll
  for (int i = 0; i < 5; ++i) block_add(c[i], red, out + i);
}

// ---- Q15 (reference kernels/q15.metal:13-32, src/q15.cpp) ----------------------

__global__ __launch_bounds__(256) void q15_scan(const int32_t* __restrict__ ship,
                                                const int32_t* __restrict__ sk,
                                                const int32_t* __restrict__ disc,
                                                const int64_t* __restrict__ price, uint32_t n,
                                                int32_t lo, int32_t hi, int64_t* __restrict__ rev,
                                                uint8_t* __restrict__ seen) {
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ship, g); }, [&](uint32_t g, const int4& s) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = lane_of(s, j);
      m |= (uint32_t)((4 * g + j < n) & (v >= lo) & (v < hi)) << j;
    }
    if (!m) return;
    const int4 k = ld4nt(sk, g), d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const uint32_t key = (uint32_t)lane_of(k, j);
      add_i64(&rev[key], lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j)));
      seen[key] = 1;
    }
  });
}

// Order-preserving map of int64 onto uint64 (for atomicMax).
__device__ __forceinline__ unsigned long long ord64(int64_t v) {
  return (unsigned long long)v ^ 0x8000000000000000ull;
}

__global__ __launch_bounds__(256) void q15_max(const int64_t* __restrict__ rev,
                                               const uint8_t* __restrict__ seen, uint32_t n,
                                               unsigned long long* __restrict__ best) {
  best = acc_copy(best);  // this workgroup's copy (folded by unsigned max)
  unsigned long long m = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (seen[i]) m = max(m, ord64(rev[i]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(best, m);
}

__global__ __launch_bounds__(256) void q15_collect(const int64_t* __restrict__ rev,
                                                   const uint8_t* __restrict__ seen, uint32_t n,
                                                   const unsigned long long* __restrict__ best,
                                                   uint32_t* __restrict__ cnt,
                                                   int32_t* __restrict__ keys, uint32_t cap) {
  const unsigned long long b = *best;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (seen[i] && ord64(rev[i]) == b) {
      const uint32_t pos = atomicAdd(cnt, 1u);
      if (pos < cap) keys[pos] = (int32_t)i;
    }
}

// ---- Q17 (reference kernels/q17.metal:11-57, src/q17.cpp) ----------------------

// bits[partkey] (and the folded filter) = (p_brand, p_container) = (brand, container).
__global__ __launch_bounds__(256) void q17_part(const int32_t* __restrict__ pk,
                                                const uint8_t* __restrict__ brand,
                                                const uint8_t* __restrict__ cont, uint32_t n,
                                                int32_t b, int32_t c, uint32_t* __restrict__ bits,
                                                uint32_t* __restrict__ fold) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (brand[i] == b && cont[i] == c) {
      bit_set(bits, (uint32_t)pk[i]);
      fold_set(fold, (uint32_t)pk[i]);
    }
}

// Bits of the 4 rows of group g whose part is selected (LDS filter, then bitmap).
__device__ __forceinline__ uint32_t q17_sel(const int4& p, uint32_t g, uint32_t n,
                                            const uint32_t* lfold, const uint32_t* bits) {
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t key = (uint32_t)lane_of(p, j);
    m |= (uint32_t)((4 * g + j < n) && fold_test(lfold, key) && bit_test(bits, key)) << j;
  }
  return m;
}

// Pass 1: per selected part, the quantity sum and line count of its lines.
__global__ __launch_bounds__(FOLD_BLOCK) void q17_stats(const int32_t* __restrict__ pk,
                                                        const int32_t* __restrict__ qty,
                                                        uint32_t n,
                                                        const uint32_t* __restrict__ sel,
                                                        const uint32_t* __restrict__ fold,
                                                        int64_t* __restrict__ sumq,
                                                        uint32_t* __restrict__ cnt) {
  __shared__ __align__(16) uint32_t lfold[FOLD_WORDS];
  fold_load(lfold, fold);
  walk2(quads(n), [&](uint32_t g) { return ld4nt(pk, g); }, [&](uint32_t g, const int4& p) {
    const uint32_t m = q17_sel(p, g, n, lfold, sel);
    if (!m) return;
    const int4 q = ld4nt(qty, g);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const uint32_t key = (uint32_t)lane_of(p, j);
      add_i64(&sumq[key], lane_of(q, j));
      atomicAdd(&cnt[key], 1u);
    }
  });
}

// Pass 2: SUM(extendedprice) of the selected lines with quantity * count * den <
// sum * num (quantity below num/den of the part's average).
__global__ __launch_bounds__(FOLD_BLOCK) void q17_sum(const int32_t* __restrict__ pk,
                                                      const int32_t* __restrict__ qty,
                                                      const int64_t* __restrict__ price,
                                                      uint32_t n, const uint32_t* __restrict__ sel,
                                                      const uint32_t* __restrict__ fold,
                                                      const int64_t* __restrict__ sumq,
                                                      const uint32_t* __restrict__ cnt,
                                                      int32_t num, int32_t den,
                                                      int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ __align__(16) uint32_t lfold[FOLD_WORDS];
  __shared__ int64_t red[FOLD_BLOCK / 64];
  fold_load(lfold, fold);
  int64_t sum = 0;
  walk2(quads(n), [&](uint32_t g) { return ld4nt(pk, g); }, [&](uint32_t g, const int4& p) {
    const uint32_t m = q17_sel(p, g, n, lfold, sel);
    if (!m) return;
    const int4 q = ld4nt(qty, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const uint32_t key = (uint32_t)lane_of(p, j);
      if ((int64_t)lane_of(q, j) * cnt[key] * den < sumq[key] * num) sum += lane_of(a, b, j);
    }
  });
  block_add(sum, red, out);
}

// ---- Q22 (reference kernels/q22.metal:11-67, src/q22.cpp) ----------------------

struct Q22Codes {
  int8_t idx[64];  // country code -> position in the code list, -1 if absent
};

__global__ __launch_bounds__(256) void mark_bytes(const int32_t* __restrict__ key, uint32_t n,
                                                  uint8_t* __restrict__ map) {
  constexpr int QB = 8;  // groups per lane per trip, loads issued for all first
  walk_batch<QB>(quads(n), [&](const uint32_t* g, const bool* live) {
    int4 k[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (live[b]) k[b] = ld4nt(key, g[b]);
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (live[b])
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (4 * g[b] + j < n) map[lane_of(k[b], j)] = 1;
  });
}

__device__ __forceinline__ int q22_slot(const Q22Codes& c, int32_t nation) {
  const int code = nation + 10;
  return (code >= 0 && code < 64) ? c.idx[code] : -1;
}

__global__ __launch_bounds__(256) void q22_avg(const int32_t* __restrict__ nat,
                                               const int64_t* __restrict__ bal, uint32_t n,
                                               Q22Codes c, int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t red[4];
  int64_t s = 0, k = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (q22_slot(c, nat[i]) >= 0 && bal[i] > 0) {
      s += bal[i];
      k += 1;
    }
  block_add(s, red, out);
  block_add(k, red, out + 1);
}

__global__ __launch_bounds__(256) void q22_final(const int32_t* __restrict__ ck,
                                                 const int32_t* __restrict__ nat,
                                                 const int64_t* __restrict__ bal, uint32_t n,
                                                 Q22Codes c, const uint8_t* __restrict__ has,
                                                 uint32_t has_n, int64_t* __restrict__ out) {
  __shared__ int64_t bins[14];
  if (threadIdx.x < 14) bins[threadIdx.x] = 0;
  __syncthreads();
  const int64_t sum = out[0], cnt = out[1];  // folded by acc_fold before this launch
  out = acc_copy(out);  // this workgroup's copy of the result
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int slot = q22_slot(c, nat[i]);
    if (slot < 0 || cnt == 0) continue;
    const int64_t b = bal[i];
    const uint32_t key = (uint32_t)ck[i];
    if (b * cnt > sum && !(key < has_n && has[key])) {
      lds_add_i64(&bins[2 * slot], 1);
      lds_add_i64(&bins[2 * slot + 1], b);
    }
  }
  __syncthreads();
  if (threadIdx.x < 14 && bins[threadIdx.x]) add_i64(&out[2 + threadIdx.x], bins[threadIdx.x]);
}

}  // namespace

extern "C" {

int gdb_q14(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q14_params* p, gdb_q14_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *ls, *lp, *ld, *le, *pp, *pt;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipdate", GDB_I32, &ls));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_partkey", GDB_I32, &lp));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_type", GDB_U8, &pt));
  const uint64_t nl = db->lineitem->nrows, np = db->part->nrows;
  uint64_t s1, s2;
  GDB_TRY(key_span(pp, np, &s1));
  GDB_TRY(key_span(lp, nl, &s2));
  const uint64_t words = (std::max(s1, s2) + 31) / 32;
  uint32_t* bits;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_BITMAP, words * 4, (void**)&bits));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(bits, 0, words * 4, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (np)
    code_bits<<<blocks_for(ctx, code_bits, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const uint8_t*)pt->dev, (uint32_t)np, mask4(p->type_mask), bits);
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {ls, lp, ld, le}, [&](const void* const* d, uint32_t rows) {
    q14_scan<<<blocks_for(ctx, q14_scan, quads(rows)), BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int64_t*)d[3],
        rows, p->date_lo, p->date_hi, bits, dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(acc_finish(ctx, dout, 0, 2));
  GDB_HIP(hipMemcpyAsync(host, dout, 16, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: l_shipdate, l_partkey, l_discount (4 B), l_extendedprice (8 B);
  // p_partkey (4 B), p_type (1 B). The scan must read l_shipdate of every row.
  GDB_TRY(timing_finish(ctx, timing, 20 * nl + 5 * np, 4 * nl, &rs));
  out->promo = host[0];
  out->total = host[1];
  out->promo_revenue = host[1] ? 100.0 * (double)host[0] / (double)host[1] : 0.0;
  return GDB_OK;
}

int gdb_q19(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q19_params* p, gdb_q19_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  if (p->nblocks < 0 || p->nblocks > 3) return fail(GDB_E_INVALID, "Q19 takes 0..3 blocks");
  const Column *lm, *li, *lp, *lq, *ld, *le, *pp, *pb, *pc, *ps;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipmode", GDB_U8, &lm));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipinstruct", GDB_U8, &li));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_partkey", GDB_I32, &lp));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_quantity", GDB_I32, &lq));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_brand", GDB_U8, &pb));
  GDB_TRY(col(db->part, "part", "p_container", GDB_U8, &pc));
  GDB_TRY(col(db->part, "part", "p_size", GDB_I32, &ps));
  const uint64_t nl = db->lineitem->nrows, np = db->part->nrows;
  uint64_t s1, s2;
  GDB_TRY(key_span(pp, np, &s1));
  GDB_TRY(key_span(lp, nl, &s2));
  const uint64_t span = std::max(s1, s2);
  uint8_t* map;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, span, (void**)&map));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  Q19Part qp{};
  Q19Line ql{};
  qp.nblocks = ql.nblocks = p->nblocks;
  ql.modes = p->mode_mask;
  ql.instruct = p->instruct;
  for (int b = 0; b < p->nblocks; ++b) {
    qp.brand[b] = p->blocks[b].brand;
    qp.cont[b] = p->blocks[b].container_mask;
    qp.size_lo[b] = p->blocks[b].size_lo;
    qp.size_hi[b] = p->blocks[b].size_hi;
    ql.qlo[b] = p->blocks[b].qty_lo;
    ql.qhi[b] = p->blocks[b].qty_hi;
  }
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(map, 0, span, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (np)
    q19_part<<<blocks_for(ctx, q19_part, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const uint8_t*)pb->dev, (const uint8_t*)pc->dev,
        (const int32_t*)ps->dev, (uint32_t)np, qp, map);
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lm, li, lp, lq, ld, le},
                     [&](const void* const* d, uint32_t rows) {
    if (!p->nblocks) return GDB_OK;
    q19_scan<<<blocks_for(ctx, q19_scan, quads(rows)), BLOCK, 0, s>>>(
        (const uint8_t*)d[0], (const uint8_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
        (const int32_t*)d[4], (const int64_t*)d[5], rows, ql, map, dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(acc_finish(ctx, dout, 0, 1));
  GDB_HIP(hipMemcpyAsync(host, dout, 8, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: shipmode, shipinstruct (1 B), partkey, quantity, discount (4 B),
  // extendedprice (8 B); part key, brand, container, size. Mandatory: 2 B per line.
  GDB_TRY(timing_finish(ctx, timing, 22 * nl + 10 * np, 2 * nl, &rs));
  out->revenue = host[0];
  return GDB_OK;
}

int gdb_q12(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q12_params* p, gdb_q12_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *lm, *lr, *lc, *ls, *lo, *oo, *op;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipmode", GDB_U8, &lm));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_receiptdate", GDB_I32, &lr));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_commitdate", GDB_I32, &lc));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipdate", GDB_I32, &ls));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_orderpriority", GDB_U8, &op));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows;
  uint64_t s1, s2;
  GDB_TRY(key_span(oo, no, &s1));
  GDB_TRY(key_span(lo, nl, &s2));
  const uint64_t span = std::max(s1, s2);
  uint8_t* high;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, span, (void**)&high));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(high, 0, span, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (no)
    order_prio_bytes<<<blocks_for(ctx, order_prio_bytes, no), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const uint8_t*)op->dev, (uint32_t)no, p->high_mask, high);
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lm, lr, lc, ls, lo},
                     [&](const void* const* d, uint32_t rows) {
    q12_scan<<<blocks_for(ctx, q12_scan, quads(rows)), BLOCK, 0, s>>>(
        (const uint8_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
        (const int32_t*)d[4], rows, p->mode[0], p->mode[1], p->date_lo, p->date_hi, high, dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(acc_finish(ctx, dout, 0, 4));
  GDB_HIP(hipMemcpyAsync(host, dout, 32, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: shipmode (1 B), receipt, commit, ship, orderkey (4 B); o_orderkey,
  // o_orderpriority. Mandatory: shipmode and receiptdate, 5 B per line.
  GDB_TRY(timing_finish(ctx, timing, 17 * nl + 5 * no, 5 * nl, &rs));
  out->low[0] = host[0];
  out->high[0] = host[1];
  out->low[1] = host[2];
  out->high[1] = host[3];
  return GDB_OK;
}

int gdb_q4(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q4_params* p, gdb_q4_result* out,
           gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *lo, *lc, *lr, *oo, *od, *op;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_commitdate", GDB_I32, &lc));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_receiptdate", GDB_I32, &lr));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_orderdate", GDB_I32, &od));
  GDB_TRY(col(db->orders, "orders", "o_orderpriority", GDB_U8, &op));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows;
  uint64_t s1, s2;
  GDB_TRY(key_span(oo, no, &s1));
  GDB_TRY(key_span(lo, nl, &s2));
  const uint64_t span = std::max(s1, s2);
  uint8_t* late;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, span, (void**)&late));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(late, 0, span, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lo, lc, lr}, [&](const void* const* d, uint32_t rows) {
    q4_late<<<blocks_for(ctx, q4_late, quads(rows)), BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], rows, late);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  if (no)
    q4_count<<<blocks_for(ctx, q4_count, quads(no)), BLOCK, 0, s>>>(
        (const int32_t*)od->dev, (const int32_t*)oo->dev, (const uint8_t*)op->dev, (uint32_t)no,
        p->date_lo, p->date_hi, late, dout);
  GDB_HIP(hipGetLastError());
  GDB_TRY(acc_finish(ctx, dout, 0, 5));
  GDB_HIP(hipMemcpyAsync(host, dout, 40, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: l_orderkey, l_commitdate, l_receiptdate (12 B per line, all read by
  // the dominant q4_late scan); o_orderkey, o_orderdate (4 B), o_orderpriority (1 B).
  GDB_TRY(timing_finish(ctx, timing, 12 * nl + 9 * no, 12 * nl, &rs));
  for (int i = 0; i < 5; ++i) out->counts[i] = host[i];
  return GDB_OK;
}

int gdb_q15(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q15_params* p, gdb_q15_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *ls, *lk, *ld, *le;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipdate", GDB_I32, &ls));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_suppkey", GDB_I32, &lk));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  const uint64_t nl = db->lineitem->nrows;
  uint64_t span;
  GDB_TRY(key_span(lk, nl, &span));
  int64_t* rev;
  uint8_t* seen;
  uint64_t* dout;
  int64_t* host;
  GDB_TRY(arena(ctx, A_AGG, span * 8, (void**)&rev));
  GDB_TRY(arena(ctx, A_MAP, span, (void**)&seen));
  GDB_TRY(arena(ctx, A_OUT, 512 + ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64 + 4 * 100, (void**)&host));
  unsigned long long* best = (unsigned long long*)dout;
  uint32_t* cnt = (uint32_t*)(dout + 1);
  int32_t* keys = (int32_t*)(dout + 2);
  int64_t* acc = (int64_t*)(dout + 64);  // per-workgroup copies of the maximum
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(rev, 0, span * 8, s));
  GDB_TRY(fill(seen, 0, span, s));
  GDB_TRY(fill(dout, 0, 16, s));
  GDB_TRY(fill(acc, 0, ACC_BYTES, s));
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {ls, lk, ld, le}, [&](const void* const* d, uint32_t rows) {
    q15_scan<<<blocks_for(ctx, q15_scan, quads(rows)), BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int64_t*)d[3],
        rows, p->date_lo, p->date_hi, rev, seen);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  q15_max<<<blocks_for(ctx, q15_max, span), BLOCK, 0, s>>>(rev, seen, (uint32_t)span,
                                                           (unsigned long long*)acc);
  GDB_TRY(acc_finish(ctx, acc, 0, 1, 0, 1u, (int64_t*)best));
  q15_collect<<<blocks_for(ctx, q15_collect, span), BLOCK, 0, s>>>(rev, seen, (uint32_t)span,
                                                                   best, cnt, keys, 100);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, dout, 16 + 4 * 100, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: l_shipdate, l_suppkey, l_discount (4 B), l_extendedprice (8 B).
  GDB_TRY(timing_finish(ctx, timing, 20 * nl, 4 * nl, &rs));
  const uint64_t b = (uint64_t)host[0];
  const uint32_t n = (uint32_t)((const uint64_t*)host)[1];
  out->nrows = (int32_t)std::min<uint32_t>(n, 100);
  out->max_revenue = b ? (int64_t)(b ^ 0x8000000000000000ull) : 0;
  const int32_t* hk = (const int32_t*)(host + 2);
  for (int i = 0; i < out->nrows; ++i) out->suppkey[i] = hk[i];
  std::sort(out->suppkey, out->suppkey + out->nrows);
  return GDB_OK;
}

int gdb_q17(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q17_params* p, gdb_q17_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  if (p->frac_den <= 0 || p->frac_num < 0) return fail(GDB_E_INVALID, "Q17 fraction must be >= 0");
  const Column *lp, *lq, *le, *pp, *pb, *pc;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_partkey", GDB_I32, &lp));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_quantity", GDB_I32, &lq));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_brand", GDB_U8, &pb));
  GDB_TRY(col(db->part, "part", "p_container", GDB_U8, &pc));
  const uint64_t nl = db->lineitem->nrows, np = db->part->nrows;
  uint64_t s1, s2;
  GDB_TRY(key_span(pp, np, &s1));
  GDB_TRY(key_span(lp, nl, &s2));
  const uint64_t span = std::max(s1, s2), words = (span + 31) / 32;
  const uint64_t words4 = (words + 3) & ~3ull;  // the folded filter starts 16-byte aligned
  uint32_t *bits, *fold, *cnt;
  int64_t *sumq, *dout, *host;
  GDB_TRY(arena(ctx, A_BITMAP, (words4 + FOLD_WORDS) * 4, (void**)&bits));
  fold = bits + words4;
  GDB_TRY(arena(ctx, A_AGG, span * 8, (void**)&sumq));
  GDB_TRY(arena(ctx, A_MAP, span * 4, (void**)&cnt));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(bits, 0, (words4 + FOLD_WORDS) * 4, s));
  GDB_TRY(fill(sumq, 0, span * 8, s));
  GDB_TRY(fill(cnt, 0, span * 4, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (np)
    q17_part<<<blocks_for(ctx, q17_part, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const uint8_t*)pb->dev, (const uint8_t*)pc->dev, (uint32_t)np,
        p->brand, p->container, bits, fold);
  GDB_HIP(hipGetLastError());
  // pass 1 over all of lineitem, then pass 2 (streamed twice for a host table)
  RunStats rs, rs2;
  GDB_TRY(scan_table(ctx, db->lineitem, {lp, lq}, [&](const void* const* d, uint32_t rows) {
    q17_stats<<<fold_blocks(ctx, (const void*)q17_stats, quads(rows)), FOLD_BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], rows, bits, fold, sumq, cnt);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  auto pass2 = [&](const void* const* d, uint32_t rows) -> int {
    q17_sum<<<fold_blocks(ctx, (const void*)q17_sum, quads(rows)), FOLD_BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int64_t*)d[2], rows, bits, fold, sumq,
        cnt, p->frac_num, p->frac_den, dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  if (!db->lineitem->host) {  // resident: launched directly, ev[1..2] keep timing pass 1
    const void* d[3] = {lp->dev, lq->dev, le->dev};
    if (nl) GDB_TRY(pass2(d, (uint32_t)nl));
  } else {
    GDB_TRY(scan_table(ctx, db->lineitem, {lp, lq, le}, pass2, &rs2));
  }
  rs.kernel_ms += rs2.kernel_ms;
  rs.h2d_bytes += rs2.h2d_bytes;
  rs.chunks += rs2.chunks;
  GDB_TRY(acc_finish(ctx, dout, 0, 2));
  GDB_HIP(hipMemcpyAsync(host, dout, 16, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: l_partkey, l_quantity (4 B), l_extendedprice (8 B); p_partkey,
  // p_brand, p_container. Mandatory per pass: l_partkey, 4 B per line.
  GDB_TRY(timing_finish(ctx, timing, 16 * nl + 6 * np, 4 * nl, &rs));
  out->sum_price = host[0];
  out->avg_yearly = (double)host[0] / 100.0 / 7.0;
  return GDB_OK;
}

int gdb_q22(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q22_params* p, gdb_q22_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  if (p->ncodes < 0 || p->ncodes > 7) return fail(GDB_E_INVALID, "Q22 takes 0..7 country codes");
  const Column *ck, *cn, *cb, *oc;
  GDB_TRY(col(db->customer, "customer", "c_custkey", GDB_I32, &ck));
  GDB_TRY(col(db->customer, "customer", "c_nationkey", GDB_I32, &cn));
  GDB_TRY(col(db->customer, "customer", "c_acctbal", GDB_I64, &cb));
  GDB_TRY(col(db->orders, "orders", "o_custkey", GDB_I32, &oc));
  const uint64_t nc = db->customer->nrows, no = db->orders->nrows;
  uint64_t s1, s2;
  GDB_TRY(key_span(ck, nc, &s1));
  GDB_TRY(key_span(oc, no, &s2));
  const uint64_t span = std::max(s1, s2);
  Q22Codes codes;
  for (int i = 0; i < 64; ++i) codes.idx[i] = -1;
  for (int i = 0; i < p->ncodes; ++i) {
    const int c = p->codes[i];
    if (c < 0 || c >= 64) return fail(GDB_E_INVALID, "Q22 country codes must lie in 0..63");
    if (codes.idx[c] < 0) codes.idx[c] = (int8_t)i;
  }
  uint8_t* has;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, span, (void**)&has));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(has, 0, span, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  GDB_HIP(kernel_mark(ctx, 1, s));
  if (no)
    mark_bytes<<<blocks_for(ctx, mark_bytes, quads(no)), BLOCK, 0, s>>>((const int32_t*)oc->dev,
                                                                         (uint32_t)no, has);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 2, s));
  if (nc) {
    q22_avg<<<blocks_for(ctx, q22_avg, nc), BLOCK, 0, s>>>(
        (const int32_t*)cn->dev, (const int64_t*)cb->dev, (uint32_t)nc, codes, dout);
    GDB_TRY(acc_finish(ctx, dout, 0, 2));  // q22_final reads the average's sum and count
    q22_final<<<blocks_for(ctx, q22_final, nc), BLOCK, 0, s>>>(
        (const int32_t*)ck->dev, (const int32_t*)cn->dev, (const int64_t*)cb->dev, (uint32_t)nc,
        codes, has, (uint32_t)span, dout);
  }
  GDB_HIP(hipGetLastError());
  GDB_TRY(acc_finish(ctx, dout, 2, 16));
  GDB_HIP(hipMemcpyAsync(host, dout, 16 * 8, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: c_custkey, c_nationkey (4 B), c_acctbal (8 B); o_custkey (4 B, the
  // dominant mark_bytes scan).
  GDB_TRY(timing_finish(ctx, timing, 16 * nc + 4 * no, 4 * no));
  // rows ordered by country code
  int order[7];
  for (int i = 0; i < p->ncodes; ++i) order[i] = i;
  std::sort(order, order + p->ncodes, [&](int a, int b) { return p->codes[a] < p->codes[b]; });
  out->nrows = 0;
  for (int t = 0; t < p->ncodes; ++t) {
    const int i = order[t];
    if (codes.idx[p->codes[i]] != i) continue;  // duplicate code
    if (host[2 + 2 * i] == 0) continue;
    gdb_q22_row& r = out->rows[out->nrows++];
    r.cntrycode = p->codes[i];
    r.pad = 0;
    r.numcust = host[2 + 2 * i];
    r.totacctbal = host[3 + 2 * i];
  }
  return GDB_OK;
}

}  // extern "C"
