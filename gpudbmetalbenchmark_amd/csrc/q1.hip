// Q1 low-cardinality grouped aggregation (reference kernels/q1.metal:39-121,
// host src/q1.cpp:8-173, SQL sql/q1.sql).
//
// Layout: one pass over 7 columns (26 B/row) with dwordx4 loads; each thread
// owns groups of 4 rows. Groups are (returnflag, linestatus) -> bin 0..5 (the
// reference's bins, kernels/q1.metal:27-32); rows that fail the filter or carry
// another code are skipped (their lanes sit out the LDS adds).
//
// Accumulation is lane-private in LDS: every lane has its own 6x4 table of u64
// counters, updated with ds_add_u64 at an address picked by the bin, so a row
// costs 4 LDS adds and no per-bin compare/select chain. 12 KiB per wave lets
// three 256-thread workgroups share a CU (a seventh, dump bin cost a third of
// that occupancy and 12% of the scan time). The 4 counters per bin:
//   0: qty | disc << 21 | 1 << 42   (count, SUM(disc), SUM(qty) packed)
//   1: price                        (SUM(extprice), cents)
//   2: price * (100 - disc)         (SUM disc_price, cents*hundredths)
//   3: price * (100 - disc) * tax   (so SUM charge = 100*[2] + [3], exact)
// Packing is exact while a lane sees < 2^14 rows and every row has
// qty <= 127, disc <= 100; prices must be < 2^32 and tax <= 100. The kernel
// counts rows outside that domain and the host rejects such tables rather
// than return a wrong sum; before the launch the host also refuses a table
// whose rows x max(price) x 100 x max(tax) reaches 2^63, since the sums of
// price*(100-disc) and price*(100-disc)*tax are int64. At block end the lane tables are unpacked and
// reduced to 36 per-block partials [bin][cnt, qty, disc, base, dp, dpt], k-major.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "internal.h"
#include "kcommon.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int NBIN = 6;        // (returnflag, linestatus) groups
constexpr int NVAL = 4;
constexpr int K = 6 * 6 + 1;   // partial columns: 6 bins x 6 values + domain-violation count
constexpr uint32_t MAX_ROWS_PER_LANE = 1u << 14;
constexpr uint64_t M21 = (1ull << 21) - 1;

__device__ __forceinline__ uint32_t rf_code(uint32_t c) {
  return c == 'A' ? 0u : c == 'N' ? 1u : c == 'R' ? 2u : 3u;
}
__device__ __forceinline__ uint32_t ls_code(uint32_t c) { return c == 'F' ? 0u : c == 'O' ? 1u : 2u; }

struct Q1Row {
  int32_t ship, qty, disc, tax;
  int64_t price;
  uint32_t rf, ls;
};

__device__ __forceinline__ void q1_row(const Q1Row& r, int32_t cutoff,
                                       unsigned long long* __restrict__ lane_tab,
                                       uint32_t& bad) {
  const uint32_t rc = rf_code(r.rf), lc = ls_code(r.ls);
  const bool ok = (r.ship <= cutoff) & (rc < 3u) & (lc < 2u);
  const uint32_t bin = rc * 2u + lc;
  const uint32_t q = (uint32_t)r.qty, d = (uint32_t)r.disc, t = (uint32_t)r.tax;
  const uint64_t pr = (uint64_t)r.price;
  bad += ((q > 127u) | (d > 100u) | (t > 100u) | ((pr >> 32) != 0)) ? 1u : 0u;
  const uint64_t packed = (uint64_t)(q | (d << 21)) | (1ull << 42);
  const uint64_t dp = (uint64_t)(uint32_t)pr * (uint64_t)(100u - d);
  const uint64_t dpt = dp * (uint64_t)t;
  if (!ok) return;
  unsigned long long* b = lane_tab + bin * (NVAL * WAVE);
  __hip_atomic_fetch_add(b + 0 * WAVE, (unsigned long long)packed, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(b + 1 * WAVE, (unsigned long long)pr, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(b + 2 * WAVE, (unsigned long long)dp, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(b + 3 * WAVE, (unsigned long long)dpt, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct Q1Cols {
  const int32_t *ship, *qty, *disc, *tax;
  const int64_t* price;
  const uint8_t *rf, *ls;
};

__device__ __forceinline__ void q1_group(const Q1Cols& c, uint32_t g, int32_t cutoff,
                                         unsigned long long* lane_tab, uint32_t& bad) {
  const int4 s = ld4nt(c.ship, g), q = ld4nt(c.qty, g), d = ld4nt(c.disc, g), t = ld4nt(c.tax, g);
  const longlong2 a = ld2nt(c.price, 2 * g), b = ld2nt(c.price, 2 * g + 1);
  const uint32_t rf = ld_u8x4nt(c.rf, g), ls = ld_u8x4nt(c.ls, g);
  q1_row({s.x, q.x, d.x, t.x, a.x, rf & 0xff, ls & 0xff}, cutoff, lane_tab, bad);
  q1_row({s.y, q.y, d.y, t.y, a.y, (rf >> 8) & 0xff, (ls >> 8) & 0xff}, cutoff, lane_tab, bad);
  q1_row({s.z, q.z, d.z, t.z, b.x, (rf >> 16) & 0xff, (ls >> 16) & 0xff}, cutoff, lane_tab, bad);
  q1_row({s.w, q.w, d.w, t.w, b.y, rf >> 24, ls >> 24}, cutoff, lane_tab, bad);
}

// LDS: blockDim/64 waves x NBIN x NVAL x 64 lanes x 8 B (12 KiB per wave).
template <int U>
__global__ __launch_bounds__(256) void q1_scan(Q1Cols c, uint32_t n, int32_t cutoff,
                                               int64_t* __restrict__ partials) {
  extern __shared__ unsigned long long tab[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  unsigned long long* wtab = tab + (size_t)w * NBIN * NVAL * WAVE;
  for (int i = lane; i < NBIN * NVAL * WAVE; i += WAVE) wtab[i] = 0ull;
  __syncthreads();
  unsigned long long* lane_tab = wtab + lane;

  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t bad = 0;
  for (; g + (U - 1) * nthr < nvec; g += U * nthr) {
#pragma unroll
    for (int u = 0; u < U; ++u) q1_group(c, g + u * nthr, cutoff, lane_tab, bad);
  }
  for (; g < nvec; g += nthr) q1_group(c, g, cutoff, lane_tab, bad);
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const uint32_t r = nvec * 4 + threadIdx.x;
    q1_row({c.ship[r], c.qty[r], c.disc[r], c.tax[r], c.price[r], c.rf[r], c.ls[r]}, cutoff,
           lane_tab, bad);
  }
  __syncthreads();

  // Unpack and reduce: wave w reduces its own table; red[w][k] holds the wave totals.
  __shared__ int64_t red[4][K];
  for (int bin = 0; bin < 6; ++bin) {
    const unsigned long long* bt = wtab + bin * NVAL * WAVE;
    const uint64_t pk = bt[lane];
    const int64_t cnt = wave_sum((int64_t)(pk >> 42));
    const int64_t qty = wave_sum((int64_t)(pk & M21));
    const int64_t dsc = wave_sum((int64_t)((pk >> 21) & M21));
    const int64_t base = wave_sum((int64_t)bt[WAVE + lane]);
    const int64_t dp = wave_sum((int64_t)bt[2 * WAVE + lane]);
    const int64_t dpt = wave_sum((int64_t)bt[3 * WAVE + lane]);
    if (lane == 0) {
      red[w][bin * 6 + 0] = cnt;
      red[w][bin * 6 + 1] = qty;
      red[w][bin * 6 + 2] = dsc;
      red[w][bin * 6 + 3] = base;
      red[w][bin * 6 + 4] = dp;
      red[w][bin * 6 + 5] = dpt;
    }
  }
  const int64_t wb = wave_sum((int64_t)bad);
  if (lane == 0) red[w][K - 1] = wb;
  __syncthreads();
  if (threadIdx.x < K) {
    int64_t s = 0;
    for (int i = 0; i < nw; ++i) s += red[i][threadIdx.x];
    partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

}  // namespace

extern "C" int gdb_q1(gdb_ctx* ctx, const gdb_table* li, const gdb_q1_params* p,
                      gdb_q1_result* out, gdb_timing* timing) {
  if (!ctx || !li || !p || !out) return fail(GDB_E_INVALID, "null argument");
  const Column* c[7];
  GDB_TRY(column(li, "l_shipdate", GDB_I32, &c[0]));
  GDB_TRY(column(li, "l_quantity", GDB_I32, &c[1]));
  GDB_TRY(column(li, "l_discount", GDB_I32, &c[2]));
  GDB_TRY(column(li, "l_tax", GDB_I32, &c[3]));
  GDB_TRY(column(li, "l_extendedprice", GDB_I64, &c[4]));
  GDB_TRY(column(li, "l_returnflag", GDB_U8, &c[5]));
  GDB_TRY(column(li, "l_linestatus", GDB_U8, &c[6]));
  const uint64_t n = li->nrows;
  // Rows in the per-row domain still sum into int64: the largest term is
  // price*(100-disc)*tax (disc >= 0 there), which bounds the other sums.
  GDB_TRY(sum_fits(n, (unsigned __int128)magnitude(c[4]) * 100u * std::max<uint64_t>(1, magnitude(c[3])),
                   "Q1 SUM(extprice*(100-disc)*tax)"));
  GDB_HIP(hipSetDevice(ctx->device));
  const Tuning& tu = ctx->tune[T_Q1];
  const int block = (tu.block == 64 || tu.block == 128 || tu.block == 256) ? tu.block : 256;
  const size_t lds = (size_t)(block / 64) * NBIN * NVAL * WAVE * sizeof(unsigned long long);
  const void* kfn = tu.unroll >= 4   ? (const void*)q1_scan<4>
                    : tu.unroll == 1 ? (const void*)q1_scan<1>
                                     : (const void*)q1_scan<2>;
  const int resident = resident_blocks(ctx, kfn, block, lds, tu.blocks_per_cu);
  // Worst case per launch: enough blocks that no lane sees more than
  // MAX_ROWS_PER_LANE rows (the packed counters' capacity).
  const uint64_t max_launch_rows = li->host ? std::min<uint64_t>(n, 1ull << 32) : n;
  const uint64_t need_max = (max_launch_rows + (uint64_t)block * MAX_ROWS_PER_LANE - 1) /
                            ((uint64_t)block * MAX_ROWS_PER_LANE);
  const int max_blocks = std::max<int>(resident, (int)need_max);
  int64_t *partials, *dout, *host, *hdev;
  GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * K * max_blocks, (void**)&partials));
  GDB_TRY(arena(ctx, A_OUT, sizeof(int64_t) * K, (void**)&dout));
  GDB_TRY(result_buf(ctx, sizeof(int64_t) * K, (void**)&host, (void**)&hdev));
  hipStream_t s = ctx->stream;
  int launches = 0;
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const uint64_t need = ((uint64_t)rows + (uint64_t)block * MAX_ROWS_PER_LANE - 1) /
                          ((uint64_t)block * MAX_ROWS_PER_LANE);
    int blocks = std::max<int>(resident, (int)need);
    blocks = std::max(1, std::min<int>(blocks, std::max<int>((int)need, (int)((rows / 4 + block - 1) / block))));
    const Q1Cols q{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                   (const int32_t*)d[3], (const int64_t*)d[4], (const uint8_t*)d[5],
                   (const uint8_t*)d[6]};
    if (tu.unroll >= 4)
      q1_scan<4><<<blocks, block, lds, s>>>(q, rows, p->cutoff, partials);
    else if (tu.unroll == 1)
      q1_scan<1><<<blocks, block, lds, s>>>(q, rows, p->cutoff, partials);
    else
      q1_scan<2><<<blocks, block, lds, s>>>(q, rows, p->cutoff, partials);
    GDB_HIP(hipGetLastError());
    if (launches++ == 0)
      sum_columns<0><<<K, 256, 0, s>>>(partials, (uint32_t)blocks, dout, hdev);
    else
      sum_columns<1><<<K, 256, 0, s>>>(partials, (uint32_t)blocks, dout, hdev);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  RunStats rs;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  if (n == 0) {
    GDB_HIP(kernel_mark(ctx, 1, s));
    GDB_HIP(kernel_mark(ctx, 2, s));
  } else {
    GDB_TRY(run_over(ctx, li, c, 7, consume, &rs));
  }
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  GDB_TRY(timing_finish(ctx, timing, 26ull * n, 26ull * n, &rs));
  if (n == 0) memset(host, 0, sizeof(int64_t) * K);
  if (host[K - 1] != 0)
    return fail(GDB_E_UNSUPPORTED,
                "Q1 input outside the exact domain (quantity 0..127, discount 0..100, tax "
                "0..100, extendedprice 0..2^32-1): " + std::to_string(host[K - 1]) + " rows");
  memset(out, 0, sizeof(*out));
  static const char RF[3] = {'A', 'N', 'R'}, LS[2] = {'F', 'O'};
  for (int bin = 0; bin < 6; ++bin) {
    const int64_t* v = host + bin * 6;
    if (v[0] == 0) continue;
    gdb_q1_row& r = out->rows[out->nrows++];
    r.returnflag = RF[bin / 2];
    r.linestatus = LS[bin % 2];
    r.count = v[0];
    r.sum_qty = v[1];
    r.sum_disc = v[2];
    r.sum_base_price = v[3];
    r.sum_disc_price = v[4];
    const __int128 charge = (__int128)v[4] * 100 + (__int128)v[5];
    r.sum_charge_lo = (uint64_t)charge;
    r.sum_charge_hi = (int64_t)(charge >> 64);
    r.avg_qty = (double)r.sum_qty / (double)r.count;
    r.avg_price = (double)r.sum_base_price / 100.0 / (double)r.count;
    r.avg_disc = (double)r.sum_disc / 100.0 / (double)r.count;
  }
  return GDB_OK;
}
