// Q13 left-outer count + histogram (reference kernels/q13.metal:33-167, host
// src/q13.cpp:9-116, post-processing src/infra.h:779-799, SQL sql/q13.sql).
//
//   match     (dominant) tiles of 256 orders: their 80-byte comments are copied
//             global->LDS with contiguous dwordx4 loads (the tile is one
//             20 KiB run), then each lane searches its comment in LDS for
//             word1 ... word2 (strmatch.h). The custkeys of orders whose
//             comment does NOT match go to the block's own output region, with
//             a per-block histogram over custkey buckets of 16384 keys
//   scan      exclusive scan of the [bucket][block] counts (scan_util)
//   scatter   every block moves its custkeys to their bucket's segment
//   bucket    one block per bucket counts its segment in LDS and writes the
//             bucket's per-custkey counts: no global atomics anywhere (random
//             per-order global atomics, the reference's kernels/q13.metal:125,
//             run at ~10 G/s on MI355X and dominated this pipeline)
// A host-resident (streamed) orders table uses q13_count instead: one global
// atomic per qualifying order into the zeroed counters.
//   histogram per customer row, c_count = counts[custkey] into an LDS-private
//             1024-bin histogram flushed with global atomics; larger counts go
//             to an overflow list the host folds in, so there is no bin cap
//             (the reference caps at 256 bins, kernels/q13.metal:145)
// Customers without qualifying orders land in bin 0: the LEFT OUTER JOIN.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "internal.h"
#include "kcommon.h"
#include "scan_util.h"
#include "strmatch.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int TILE = 256;
constexpr int CW = 80;         // comment width
constexpr int SLOT = 84;       // LDS bytes per comment: 21 dwords, odd stride (no bank conflicts)
constexpr int NBINS = 1024;
constexpr int BUCKET_SHIFT = 14;                 // 16384 custkeys per bucket (64 KiB of LDS counters)
constexpr uint32_t MAX_BUCKETS = 4096;

__global__ __launch_bounds__(256) void q13_clear(uint32_t* __restrict__ counts, uint32_t ncounts,
                                                 uint32_t* __restrict__ hist) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncounts + NBINS + 4; i += stride) {
    if (i < ncounts) counts[i] = 0u;
    else hist[i - ncounts] = 0u;  // NBINS bins + overflow counter + pad
  }
}

__global__ __launch_bounds__(TILE) void q13_count(const uint8_t* __restrict__ comment,
                                                  const int32_t* __restrict__ custkey, uint32_t n,
                                                  Pattern w1, Pattern w2, uint32_t maxck,
                                                  uint32_t* __restrict__ counts) {
  __shared__ __align__(16) uint8_t tile[TILE * SLOT + 16];
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  uint8_t* mine = tile + threadIdx.x * SLOT;
  // Zero the 4 pad bytes after every comment once; the copies never touch them.
  *reinterpret_cast<uint32_t*>(mine + CW) = 0u;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t row0 = t * TILE;
    const uint32_t rows = min((uint32_t)TILE, n - row0);
    const uint32_t chunks = rows * (CW / 16);
    const int4* src = reinterpret_cast<const int4*>(comment + (size_t)row0 * CW);
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < chunks; c += TILE) {
      const int4 v = src[c];
      const uint32_t r = c / (CW / 16), off = (c % (CW / 16)) * 16;
      uint32_t* d = reinterpret_cast<uint32_t*>(tile + r * SLOT + off);
      d[0] = (uint32_t)v.x;
      d[1] = (uint32_t)v.y;
      d[2] = (uint32_t)v.z;
      d[3] = (uint32_t)v.w;
    }
    __syncthreads();
    if (threadIdx.x < rows) {
      const int len = field_len(mine, CW);
      const int p1 = find_from(mine, len, 0, w1);
      const bool matches = p1 >= 0 && find_from(mine, len, p1 + w1.len, w2) >= 0;
      if (!matches) {
        const uint32_t ck = (uint32_t)custkey[row0 + threadIdx.x];
        if (ck <= maxck) atomicAdd(&counts[ck], 1u);
      }
    }
  }
}

// Copies tile t (256 comments) into LDS; returns the rows in the tile.
__device__ __forceinline__ uint32_t q13_load_tile(const uint8_t* __restrict__ comment, uint32_t n,
                                                  uint32_t t, uint8_t* tile) {
  const uint32_t row0 = t * TILE;
  const uint32_t rows = min((uint32_t)TILE, n - row0);
  const uint32_t chunks = rows * (CW / 16);
  const int4* src = reinterpret_cast<const int4*>(comment + (size_t)row0 * CW);
  for (uint32_t c = threadIdx.x; c < chunks; c += TILE) {
    const int4 v = ld4nt(reinterpret_cast<const int32_t*>(src), c);
    const uint32_t r = c / (CW / 16), off = (c % (CW / 16)) * 16;
    uint32_t* d = reinterpret_cast<uint32_t*>(tile + r * SLOT + off);
    d[0] = (uint32_t)v.x;
    d[1] = (uint32_t)v.y;
    d[2] = (uint32_t)v.z;
    d[3] = (uint32_t)v.w;
  }
  return rows;
}

// Pass A: block b owns tiles [b*tpb, (b+1)*tpb); the custkeys of its counted
// orders go to region b (keys + b*tpb*TILE), with a per-bucket histogram
// written k-major to hist[k*G + b].
__global__ __launch_bounds__(TILE) void q13_match(const uint8_t* __restrict__ comment,
                                                  const int32_t* __restrict__ custkey, uint32_t n,
                                                  Pattern w1, Pattern w2, uint32_t maxck,
                                                  uint32_t tpb, uint32_t nb,
                                                  uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ region_count,
                                                  uint32_t* __restrict__ hist) {
  __shared__ __align__(16) uint8_t tile[TILE * SLOT + 16];
  __shared__ uint32_t lhist[MAX_BUCKETS];
  __shared__ uint32_t scratch[17];
  for (uint32_t i = threadIdx.x; i < nb; i += TILE) lhist[i] = 0u;
  uint8_t* mine = tile + threadIdx.x * SLOT;
  *reinterpret_cast<uint32_t*>(mine + CW) = 0u;
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  const uint32_t t0 = blockIdx.x * tpb, t1 = min(ntiles, t0 + tpb);
  uint32_t* out = keys + (size_t)blockIdx.x * tpb * TILE;
  uint32_t count = 0;  // block-uniform
  for (uint32_t t = t0; t < t1; ++t) {
    __syncthreads();
    const uint32_t rows = q13_load_tile(comment, n, t, tile);
    __syncthreads();
    bool q = false;
    uint32_t ck = 0;
    if (threadIdx.x < rows) {
      const int len = field_len(mine, CW);
      const int p1 = find_from(mine, len, 0, w1);
      const bool matches = p1 >= 0 && find_from(mine, len, p1 + w1.len, w2) >= 0;
      if (!matches) {
        ck = (uint32_t)custkey[t * TILE + threadIdx.x];
        q = ck <= maxck;
      }
    }
    uint32_t total;
    const uint32_t pos = block_excl_scan(q ? 1u : 0u, scratch, &total);
    if (q) {
      out[count + pos] = ck;
      atomicAdd(&lhist[ck >> BUCKET_SHIFT], 1u);
    }
    count += total;
  }
  __syncthreads();
  if (threadIdx.x == 0) region_count[blockIdx.x] = count;
  for (uint32_t k = threadIdx.x; k < nb; k += TILE) hist[(size_t)k * gridDim.x + blockIdx.x] = lhist[k];
}

// Pass C: region b's custkeys to their bucket segments (offs = scanned hist).
__global__ __launch_bounds__(TILE) void q13_scatter(const uint32_t* __restrict__ keys, uint32_t tpb,
                                                    const uint32_t* __restrict__ region_count,
                                                    const uint32_t* __restrict__ offs, uint32_t nb,
                                                    uint32_t* __restrict__ sorted) {
  __shared__ uint32_t base[MAX_BUCKETS];
  for (uint32_t k = threadIdx.x; k < nb; k += TILE) base[k] = offs[(size_t)k * gridDim.x + blockIdx.x];
  __syncthreads();
  const uint32_t* in = keys + (size_t)blockIdx.x * tpb * TILE;
  const uint32_t cnt = region_count[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < cnt; i += TILE) {
    const uint32_t ck = in[i];
    sorted[atomicAdd(&base[ck >> BUCKET_SHIFT], 1u)] = ck;
  }
}

// Pass D: block k counts bucket k's custkeys in LDS and writes their counts.
__global__ __launch_bounds__(1024) void q13_bucket_count(const uint32_t* __restrict__ sorted,
                                                         const uint32_t* __restrict__ offs,
                                                         uint32_t nblocks_a,
                                                         const uint32_t* __restrict__ total,
                                                         uint32_t nb, uint32_t maxck,
                                                         uint32_t* __restrict__ counts) {
  __shared__ uint32_t c[1u << BUCKET_SHIFT];
  for (uint32_t i = threadIdx.x; i < (1u << BUCKET_SHIFT); i += blockDim.x) c[i] = 0u;
  __syncthreads();
  const uint32_t k = blockIdx.x;
  const uint32_t begin = offs[(size_t)k * nblocks_a];
  const uint32_t end = k + 1 < nb ? offs[(size_t)(k + 1) * nblocks_a] : *total;
  for (uint32_t i = begin + threadIdx.x; i < end; i += blockDim.x)
    atomicAdd(&c[sorted[i] & ((1u << BUCKET_SHIFT) - 1)], 1u);
  __syncthreads();
  const uint32_t key0 = k << BUCKET_SHIFT;
  for (uint32_t i = threadIdx.x; i < (1u << BUCKET_SHIFT); i += blockDim.x)
    if (key0 + i <= maxck) counts[key0 + i] = c[i];
}

__global__ __launch_bounds__(256) void q13_histogram(const int32_t* __restrict__ custkey,
                                                     uint32_t nc,
                                                     const uint32_t* __restrict__ counts,
                                                     uint32_t* __restrict__ hist,
                                                     uint32_t* __restrict__ overflow) {
  __shared__ uint32_t lh[NBINS];
  for (int i = threadIdx.x; i < NBINS; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) {
    const uint32_t c = counts[(uint32_t)custkey[i]];
    if (c < NBINS) {
      atomicAdd(&lh[c], 1u);
    } else {
      const uint32_t pos = atomicAdd(&hist[NBINS], 1u);
      overflow[pos] = c;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NBINS; i += blockDim.x)
    if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

bool make_pattern(const char* s, Pattern* p) {
  const size_t n = strnlen(s, 32);
  if (n == 0 || n > 31) return false;
  memset(p, 0, sizeof(*p));
  memcpy(p->b, s, n);
  p->len = (int)n;
  return true;
}

}  // namespace

extern "C" int gdb_q13(gdb_ctx* ctx, const gdb_table* customer, const gdb_table* orders,
                       const gdb_q13_params* prm, gdb_q13_result* out, gdb_timing* timing) {
  if (!ctx || !customer || !orders || !prm || !out) return fail(GDB_E_INVALID, "null argument");
  Pattern w1, w2;
  if (!make_pattern(prm->word1, &w1) || !make_pattern(prm->word2, &w2))
    return fail(GDB_E_INVALID, "Q13 words must be 1..31 bytes");
  const Column *cck, *ock, *ocom;
  GDB_TRY(column(customer, "c_custkey", GDB_I32, &cck));
  GDB_TRY(column(orders, "o_custkey", GDB_I32, &ock));
  GDB_TRY(column(orders, "o_comment", GDB_BYTES, &ocom));
  if (ocom->width != CW) return fail(GDB_E_UNSUPPORTED, "o_comment must be an 80-byte field");
  GDB_TRY(require_resident(customer, "Q13 customer"));
  if (customer->nrows && cck->minv < 0)
    return fail(GDB_E_UNSUPPORTED, "Q13 needs non-negative custkey values");
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t nc = (uint32_t)customer->nrows;
  const uint64_t no = orders->nrows;
  const uint32_t maxck = (uint32_t)std::max<int64_t>(0, cck->maxv);
  const uint32_t ncounts = maxck + 1;
  const uint32_t nb = (maxck >> BUCKET_SHIFT) + 1;
  const bool partitioned = !orders->host && nb <= MAX_BUCKETS;
  uint32_t *counts, *hist, *overflow, *host;
  {
    char* b;
    GDB_TRY(arena(ctx, A_MAP, (size_t)ncounts * 4 + 256, (void**)&b));
    counts = (uint32_t*)b;
    GDB_TRY(arena(ctx, A_AGG, (NBINS + 4) * 4 + 256, (void**)&b));
    hist = (uint32_t*)b;
    GDB_TRY(arena(ctx, A_AUX, (size_t)std::max<uint32_t>(nc, 1) * 4, (void**)&b));
    overflow = (uint32_t*)b;
  }
  GDB_TRY(pinned(ctx, (NBINS + 4) * 4, (void**)&host));
  hipStream_t st = ctx->stream;
  RunStats rs;
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  if (partitioned) {
    const uint32_t n = (uint32_t)no;
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    uint32_t G = (uint32_t)std::max(1, resident_blocks(ctx, (const void*)q13_match, TILE, 0,
                                                       ctx->tune[T_Q13].blocks_per_cu));
    G = std::max<uint32_t>(1, std::min<uint32_t>(G, ntiles));
    const uint32_t tpb = std::max<uint32_t>(1, (ntiles + G - 1) / G);
    G = std::max<uint32_t>(1, (ntiles + tpb - 1) / tpb);
    const uint32_t nhist = nb * G;
    uint32_t *keys, *sorted, *rcount, *bh, *tsum, *total;
    char* b;
    GDB_TRY(arena(ctx, A_AUX2, (size_t)G * tpb * TILE * 4 + 256, (void**)&keys));
    GDB_TRY(arena(ctx, A_HT, (size_t)std::max<uint32_t>(n, 1) * 4 + 256, (void**)&sorted));
    const uint32_t ntsum = scan_tile_sums_needed(nhist);
    GDB_TRY(arena(ctx, A_PARTIALS, ((size_t)G + nhist + ntsum + 64) * 4, (void**)&b));
    rcount = (uint32_t*)b;
    bh = rcount + G;
    tsum = bh + nhist;
    total = tsum + ntsum;
    q13_clear<<<5, 256, 0, st>>>(counts, 0, hist);
    GDB_HIP(hipEventRecord(ctx->ev[1], st));
    q13_match<<<G, TILE, 0, st>>>((const uint8_t*)ocom->dev, (const int32_t*)ock->dev, n, w1, w2,
                                  maxck, tpb, nb, keys, rcount, bh);
    GDB_HIP(hipGetLastError());
    GDB_HIP(hipEventRecord(ctx->ev[2], st));
    GDB_TRY(exclusive_scan_u32(ctx, bh, nhist, bh, total, tsum, st));
    q13_scatter<<<G, TILE, 0, st>>>(keys, tpb, rcount, bh, nb, sorted);
    q13_bucket_count<<<nb, 1024, 0, st>>>(sorted, bh, G, total, nb, maxck, counts);
    GDB_HIP(hipGetLastError());
  } else {
    q13_clear<<<std::min<uint32_t>(4096, (ncounts + NBINS + 4 + 255) / 256), 256, 0, st>>>(
        counts, ncounts, hist);
    const int cresident =
        resident_blocks(ctx, (const void*)q13_count, TILE, 0, ctx->tune[T_Q13].blocks_per_cu);
    auto consume = [&](const void* const* d, uint32_t rows) -> int {
      const int blocks = std::max(1, std::min<int>(cresident, (int)((rows + TILE - 1) / TILE)));
      q13_count<<<blocks, TILE, 0, st>>>((const uint8_t*)d[0], (const int32_t*)d[1], rows, w1,
                                          w2, maxck, counts);
      GDB_HIP(hipGetLastError());
      return GDB_OK;
    };
    const Column* oc[2] = {ocom, ock};
    if (no) {
      GDB_TRY(run_over(ctx, orders, oc, 2, consume, &rs));
    } else {
      GDB_HIP(hipEventRecord(ctx->ev[1], st));
      GDB_HIP(hipEventRecord(ctx->ev[2], st));
    }
  }
  if (nc)
    q13_histogram<<<std::min<uint32_t>(1024, (nc + 255) / 256), 256, 0, st>>>(
        (const int32_t*)cck->dev, nc, counts, hist, overflow);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, hist, (NBINS + 1) * 4, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  GDB_TRY(timing_finish(ctx, timing, 84ull * no + 4ull * nc, 84ull * no, &rs));
  std::map<uint32_t, int64_t> h;
  for (int i = 0; i < NBINS; ++i)
    if (host[i]) h[i] += host[i];
  const uint32_t nover = host[NBINS];
  if (nover) {
    std::vector<uint32_t> ov(nover);
    GDB_HIP(hipMemcpy(ov.data(), overflow, nover * 4ull, hipMemcpyDeviceToHost));
    for (uint32_t c : ov) h[c] += 1;
  }
  if (h.size() > sizeof(out->rows) / sizeof(out->rows[0]))
    return fail(GDB_E_UNSUPPORTED, "more distinct c_count values than the result holds");
  std::vector<gdb_q13_row> rows;
  for (auto& kv : h) rows.push_back(gdb_q13_row{(int32_t)kv.first, 0, kv.second});
  std::sort(rows.begin(), rows.end(), [](const gdb_q13_row& a, const gdb_q13_row& b) {
    if (a.custdist != b.custdist) return a.custdist > b.custdist;
    return a.c_count > b.c_count;
  });
  memset(out, 0, sizeof(*out));
  out->nrows = (int32_t)rows.size();
  std::copy(rows.begin(), rows.end(), out->rows);
  return GDB_OK;
}
