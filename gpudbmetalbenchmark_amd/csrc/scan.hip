// Predicate-scan pipelines: Q6 filter-scan-sum (reference kernels/q6.metal:12-73,
// host src/q6.cpp:8-120) and the selection / aggregation primitives (reference
// kernels/microbenchmarks.metal:158-218, host src/microbenchmarks.cpp:51-131).
//
// Every scan streams its columns with dwordx4 loads: each thread owns groups of
// 4 consecutive rows (16 B of each int32 column, 32 B of an int64 column) and
// walks the table with a grid stride sized to the resident wave count, so one
// launch keeps all CUs busy and each column byte is read exactly once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "internal.h"
#include "kcommon.h"

using namespace gdb;
using namespace gdbk;

namespace {

struct Q6P {
  int32_t dlo, dhi, clo, chi, qlt;
};

__device__ __forceinline__ void q6_row(int32_t s, int32_t d, int32_t q, int64_t pr, const Q6P& p,
                                       int64_t& rev, uint32_t& cnt) {
  const bool ok = (s >= p.dlo) & (s < p.dhi) & (d >= p.clo) & (d <= p.chi) & (q < p.qlt);
  rev += ok ? pr * (int64_t)d : 0;
  cnt += ok ? 1u : 0u;
}

__device__ __forceinline__ void q6_group(const int32_t* __restrict__ ship,
                                         const int32_t* __restrict__ disc,
                                         const int32_t* __restrict__ qty,
                                         const int64_t* __restrict__ price, uint32_t g,
                                         const Q6P& p, int64_t& rev, uint32_t& cnt) {
  const int4 s = ld4nt(ship, g), d = ld4nt(disc, g), q = ld4nt(qty, g);
  const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
  q6_row(s.x, d.x, q.x, a.x, p, rev, cnt);
  q6_row(s.y, d.y, q.y, a.y, p, rev, cnt);
  q6_row(s.z, d.z, q.z, b.x, p, rev, cnt);
  q6_row(s.w, d.w, q.w, b.y, p, rev, cnt);
}

// Q6 scan: per block, partials[0][block] = SUM(price*disc), partials[1][block] =
// selected rows. Rows [4*nvec, n) are handled by the first threads of block 0.
template <int U>
__global__ __launch_bounds__(256) void q6_scan(const int32_t* __restrict__ ship,
                                               const int32_t* __restrict__ disc,
                                               const int32_t* __restrict__ qty,
                                               const int64_t* __restrict__ price, uint32_t n,
                                               Q6P p, int64_t* __restrict__ partials) {
  __shared__ int64_t red[4];
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  int64_t rev = 0;
  uint32_t cnt = 0;
  for (; g + (U - 1) * nthr < nvec; g += U * nthr) {
#pragma unroll
    for (int u = 0; u < U; ++u) q6_group(ship, disc, qty, price, g + u * nthr, p, rev, cnt);
  }
  for (; g < nvec; g += nthr) q6_group(ship, disc, qty, price, g, p, rev, cnt);
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const uint32_t r = nvec * 4 + threadIdx.x;
    q6_row(ship[r], disc[r], qty[r], price[r], p, rev, cnt);
  }
  const int64_t brev = block_sum(rev, red);
  const int64_t bcnt = block_sum((int64_t)cnt, red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = brev;
    partials[gridDim.x + blockIdx.x] = bcnt;
  }
}

// SUM of an int32 or int64 column.
template <typename T>
__global__ __launch_bounds__(256) void sum_scan(const T* __restrict__ col, uint32_t n,
                                                int64_t* __restrict__ partials) {
  __shared__ int64_t red[4];
  constexpr uint32_t PER = 16 / sizeof(T);  // values per dwordx4
  const uint32_t nvec = n / PER;
  const uint32_t nthr = gridDim.x * blockDim.x;
  int64_t s = 0;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nvec; g += nthr) {
    const int4 v = reinterpret_cast<const int4*>(col)[g];
    if constexpr (sizeof(T) == 4) {
      s += (int64_t)v.x + v.y + v.z + v.w;
    } else {
      s += (int64_t)(((uint64_t)(uint32_t)v.y << 32) | (uint32_t)v.x) +
           (int64_t)(((uint64_t)(uint32_t)v.w << 32) | (uint32_t)v.z);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < n % PER) s += (int64_t)col[nvec * PER + threadIdx.x];
  s = block_sum(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Selection `col < value`, pass 1: the bitmap (bit r of word r/64) and one
// count per block. Block b owns rows [b*chunk, (b+1)*chunk), chunk a multiple
// of 1024, so pass 2 can emit row ids in ascending order. Each lane tests 4
// rows from one dwordx4 load; 16 lanes' nibbles make one bitmap word.
__global__ __launch_bounds__(256) void select_lt_bitmap(const int32_t* __restrict__ col,
                                                        uint32_t n, int32_t value,
                                                        uint32_t chunk,
                                                        uint64_t* __restrict__ bitmap,
                                                        uint32_t* __restrict__ block_counts) {
  __shared__ int64_t red[4];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t begin = blockIdx.x * chunk;
  const uint32_t end = min(n, begin + chunk);
  uint32_t cnt = 0;
  for (uint32_t base = begin; base < end; base += 4 * blockDim.x) {
    const uint32_t r0 = base + 4 * threadIdx.x;
    uint32_t nib = 0;
    if (r0 + 3 < end) {
      const int4 v = ld4(col, r0 >> 2);
      nib = (uint32_t)(v.x < value) | ((uint32_t)(v.y < value) << 1) |
            ((uint32_t)(v.z < value) << 2) | ((uint32_t)(v.w < value) << 3);
    } else {
      for (uint32_t j = 0; j < 4; ++j)
        if (r0 + j < end && col[r0 + j] < value) nib |= 1u << j;
    }
    cnt += __popc(nib);
    // Gather 16 nibbles into a 64-bit word: lane 16k+i contributes bits 4i..4i+3.
    uint64_t word = (uint64_t)nib << (4 * (lane & 15));
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) word |= __shfl_xor((unsigned long long)word, off, 64);
    if ((lane & 15) == 0 && r0 < end) bitmap[r0 >> 6] = word;
  }
  const int64_t b = block_sum((int64_t)cnt, red);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = (uint32_t)b;
}

// Exclusive scan of the block counts (one block; nblocks <= 256 * 64).
__global__ __launch_bounds__(256) void scan_counts(const uint32_t* __restrict__ counts,
                                                   uint32_t nblocks, uint32_t* __restrict__ offs,
                                                   uint64_t* __restrict__ total) {
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nblocks; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nblocks ? counts[i] : 0;
    const uint32_t incl = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = carry;
    for (int k = 0; k < w; ++k) pre += wsum[k];
    if (i < nblocks) offs[i] = pre + incl - v;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry = pre + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Selection pass 2: each lane expands one bitmap word of its block's range into
// ascending row ids at the block's offset plus the wave/block prefix.
__global__ __launch_bounds__(256) void select_emit(const uint64_t* __restrict__ bitmap,
                                                   uint32_t n, uint32_t chunk,
                                                   const uint32_t* __restrict__ offs,
                                                   uint32_t* __restrict__ rowids) {
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = offs[blockIdx.x];
  __syncthreads();
  const uint32_t begin = blockIdx.x * chunk;
  const uint32_t end = min(n, begin + chunk);
  const uint32_t wbegin = begin >> 6, wend = (end + 63) >> 6;
  for (uint32_t base = wbegin; base < wend; base += blockDim.x) {
    const uint32_t wi = base + threadIdx.x;
    const uint64_t word = wi < wend ? bitmap[wi] : 0ull;
    const uint32_t c = (uint32_t)__popcll(word);
    const uint32_t incl = wave_incl_scan(c);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pos = carry;
    for (int k = 0; k < w; ++k) pos += wsum[k];
    pos += incl - c;
    uint64_t m = word;
    while (m) {
      const int b = __ffsll((unsigned long long)m) - 1;
      rowids[pos++] = wi * 64 + (uint32_t)b;
      m &= m - 1;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry = pos;
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int gdb_q6(gdb_ctx* ctx, const gdb_table* li, const gdb_q6_params* p, gdb_q6_result* out,
           gdb_timing* timing) {
  if (!ctx || !li || !p || !out) return fail(GDB_E_INVALID, "null argument");
  const Column* c[4];
  GDB_TRY(column(li, "l_shipdate", GDB_I32, &c[0]));
  GDB_TRY(column(li, "l_discount", GDB_I32, &c[1]));
  GDB_TRY(column(li, "l_quantity", GDB_I32, &c[2]));
  GDB_TRY(column(li, "l_extendedprice", GDB_I64, &c[3]));
  GDB_TRY(sum_fits(li->nrows, (unsigned __int128)magnitude(c[3]) * magnitude(c[1]),
                   "Q6 SUM(extprice*discount)"));
  GDB_HIP(hipSetDevice(ctx->device));
  const uint64_t n = li->nrows;
  const Tuning& tu = ctx->tune[T_Q6];
  const int block = 256;
  const int resident = resident_blocks(ctx, (const void*)q6_scan<2>, block, 0, tu.blocks_per_cu);
  int64_t *partials, *dout, *host, *hdev;
  GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * 2 * resident, (void**)&partials));
  GDB_TRY(arena(ctx, A_OUT, 64, (void**)&dout));
  GDB_TRY(result_buf(ctx, 64, (void**)&host, (void**)&hdev));
  const Q6P kp{p->date_lo, p->date_hi, p->disc_lo, p->disc_hi, p->qty_lt};
  hipStream_t s = ctx->stream;
  int launches = 0;
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const int blocks = std::max(1, std::min<int>(resident, (int)((rows / 4 + block - 1) / block)));
    const int32_t *sp = (const int32_t*)d[0], *dp = (const int32_t*)d[1], *qp = (const int32_t*)d[2];
    const int64_t* pp = (const int64_t*)d[3];
    if (tu.unroll >= 4)
      q6_scan<4><<<blocks, block, 0, s>>>(sp, dp, qp, pp, rows, kp, partials);
    else if (tu.unroll == 1)
      q6_scan<1><<<blocks, block, 0, s>>>(sp, dp, qp, pp, rows, kp, partials);
    else
      q6_scan<2><<<blocks, block, 0, s>>>(sp, dp, qp, pp, rows, kp, partials);
    GDB_HIP(hipGetLastError());
    // partials is [2][blocks]; the finaliser reads column k at k*blocks.
    if (launches++ == 0)
      sum_columns<0><<<2, 256, 0, s>>>(partials, (uint32_t)blocks, dout, hdev);
    else
      sum_columns<1><<<2, 256, 0, s>>>(partials, (uint32_t)blocks, dout, hdev);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  RunStats rs;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  if (n == 0) {
    GDB_HIP(kernel_mark(ctx, 1, s));
    GDB_HIP(kernel_mark(ctx, 2, s));
  } else {
    GDB_TRY(run_over(ctx, li, c, 4, consume, &rs));
  }
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  GDB_TRY(timing_finish(ctx, timing, 20ull * n, 20ull * n, &rs));
  if (n == 0) memset(host, 0, 2 * sizeof(int64_t));
  out->revenue = host[0];
  out->rows = (uint64_t)host[1];
  return GDB_OK;
}

// ---- Q6 without allocation --------------------------------------------------
// gdb_q6_noalloc runs the same kernels as gdb_q6 with their scratch (the
// per-workgroup partials and the two totals) in `scratch`, a fixed array in
// device memory that every GPU the library runs on holds until the process
// ends, instead of the context's arenas and pinned staging. Its address on a
// device comes from hipGetSymbolAddress once per device (gdb_open calls
// noalloc_scratch_init with the device current); host code never uses the
// array's name as an address. Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race.
}  // extern "C"

namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = 1u << 20;  // 1 MiB per device
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int64_t* g_scratch[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) int64_t scratch[NOALLOC_SCRATCH_BYTES / sizeof(int64_t)];

namespace gdb {
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(GDB_E_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch[device]) return GDB_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return fail(GDB_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  g_scratch[device] = (int64_t*)p;
  return GDB_OK;
}
}  // namespace gdb

extern "C" {

int gdb_q6_noalloc(gdb_ctx* ctx, const gdb_table* li, const gdb_q6_params* p, gdb_q6_result* out,
                   gdb_timing* timing) {
  if (!ctx || !li || !p || !out) return fail(GDB_E_INVALID, "null argument");
  const Column* c[4];
  GDB_TRY(column(li, "l_shipdate", GDB_I32, &c[0]));
  GDB_TRY(column(li, "l_discount", GDB_I32, &c[1]));
  GDB_TRY(column(li, "l_quantity", GDB_I32, &c[2]));
  GDB_TRY(column(li, "l_extendedprice", GDB_I64, &c[3]));
  GDB_TRY(require_resident(li, "the lineitem table of gdb_q6_noalloc"));  // streaming allocates
  GDB_TRY(sum_fits(li->nrows, (unsigned __int128)magnitude(c[3]) * magnitude(c[1]),
                   "Q6 SUM(extprice*discount)"));
  if (li->nrows >= (1ull << 32)) return fail(GDB_E_UNSUPPORTED, "gdb_q6_noalloc takes below 2^32 rows");
  GDB_HIP(hipSetDevice(ctx->device));
  // Set up by gdb_open; retried here (device current) if that failed.
  GDB_TRY(noalloc_scratch_init(ctx->device));
  int64_t* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    base = g_scratch[ctx->device];
  }
  const uint32_t n = (uint32_t)li->nrows;
  const Tuning& tu = ctx->tune[T_Q6];
  const int block = 256;
  const int resident = resident_blocks(ctx, (const void*)q6_scan<2>, block, 0, tu.blocks_per_cu);
  const int blocks = std::max(1, std::min<int>(resident, (int)((n / 4 + block - 1) / block)));
  // partials [2][blocks], then the two totals at a 256-byte boundary
  const size_t part_words = ((size_t)2 * blocks + 31) & ~size_t(31);
  const size_t need = (part_words + 2) * sizeof(int64_t);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(GDB_E_UNSUPPORTED, "gdb_q6_noalloc: this launch shape needs " + std::to_string(need) +
                                       " bytes of scratch, above its " +
                                       std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte maximum");
  int64_t* partials = base;
  int64_t* dout = base + part_words;
  const Q6P kp{p->date_lo, p->date_hi, p->disc_lo, p->disc_hi, p->qty_lt};
  hipStream_t s = ctx->stream;
  const int32_t *sp = (const int32_t*)c[0]->dev, *dp = (const int32_t*)c[1]->dev,
                *qp = (const int32_t*)c[2]->dev;
  const int64_t* pp = (const int64_t*)c[3]->dev;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(kernel_mark(ctx, 1, s));
  if (n == 0) {
    GDB_TRY(fill(dout, 0, 2 * sizeof(int64_t), s));  // a fill kernel, not a memset node
  } else {
    // Every workgroup writes both of its partials and the finaliser
    // overwrites both totals, so nothing in the scratch needs clearing.
    if (tu.unroll >= 4)
      q6_scan<4><<<blocks, block, 0, s>>>(sp, dp, qp, pp, n, kp, partials);
    else if (tu.unroll == 1)
      q6_scan<1><<<blocks, block, 0, s>>>(sp, dp, qp, pp, n, kp, partials);
    else
      q6_scan<2><<<blocks, block, 0, s>>>(sp, dp, qp, pp, n, kp, partials);
    GDB_HIP(hipGetLastError());
    sum_columns<0><<<2, 256, 0, s>>>(partials, (uint32_t)blocks, dout);
    GDB_HIP(hipGetLastError());
  }
  GDB_HIP(kernel_mark(ctx, 2, s));
  int64_t h[2] = {0, 0};  // the caller's stack, not pinned staging
  GDB_HIP(hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  GDB_HIP(hipStreamSynchronize(s));
  GDB_TRY(timing_finish(ctx, timing, 20ull * n, 20ull * n, nullptr));
  out->revenue = h[0];
  out->rows = (uint64_t)h[1];
  return GDB_OK;
}

int gdb_sum(gdb_ctx* ctx, const gdb_table* t, const char* name, int64_t* sum, gdb_timing* timing) {
  if (!ctx || !t || !sum) return fail(GDB_E_INVALID, "null argument");
  const Column* c;
  GDB_TRY(column(t, name, 0, &c));
  if (c->type != GDB_I32 && c->type != GDB_I64)
    return fail(GDB_E_INVALID, "gdb_sum needs an int32 or int64 column");
  GDB_TRY(require_resident(t, "gdb_sum's table"));
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t n = (uint32_t)t->nrows;
  const int block = 256;
  int blocks = resident_blocks(ctx, (const void*)sum_scan<int32_t>, block, 0,
                               ctx->tune[T_SUM].blocks_per_cu);
  blocks = std::max(1, std::min<int>(blocks, (int)((n / 2 + block - 1) / block)));
  int64_t *partials, *dout, *host;
  GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * blocks, (void**)&partials));
  GDB_TRY(arena(ctx, A_OUT, 64, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(kernel_mark(ctx, 1, s));
  if (c->type == GDB_I32)
    sum_scan<int32_t><<<blocks, block, 0, s>>>((const int32_t*)c->dev, n, partials);
  else
    sum_scan<int64_t><<<blocks, block, 0, s>>>((const int64_t*)c->dev, n, partials);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 2, s));
  sum_columns<<<1, 256, 0, s>>>(partials, (uint32_t)blocks, dout);
  GDB_HIP(hipMemcpyAsync(host, dout, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  const uint64_t bytes = (uint64_t)n * c->width;
  GDB_TRY(timing_finish(ctx, timing, bytes, bytes));
  *sum = host[0];
  return GDB_OK;
}

int gdb_select_lt(gdb_ctx* ctx, const gdb_table* t, const char* name, int32_t value,
                  uint64_t* bitmap, uint32_t* rowids, uint64_t* count, gdb_timing* timing) {
  if (!ctx || !t || !count) return fail(GDB_E_INVALID, "null argument");
  const Column* c;
  GDB_TRY(column(t, name, GDB_I32, &c));
  GDB_TRY(require_resident(t, "gdb_select_lt's table"));
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t n = (uint32_t)t->nrows;
  const int block = 256;
  const uint32_t nwords = (n + 63) / 64;
  int target = resident_blocks(ctx, (const void*)select_lt_bitmap, block, 0, 0);
  // Chunk: a multiple of 4*block rows, so every bitmap word lies in one block.
  uint32_t chunk = (uint32_t)(((uint64_t)n + target - 1) / target);
  chunk = std::max<uint32_t>(1024, (chunk + 1023) / 1024 * 1024);
  const uint32_t nblocks = std::max<uint32_t>(1, (n + chunk - 1) / chunk);
  uint64_t* dbm;
  uint32_t *dcounts, *doffs, *drows;
  uint64_t *dtotal, *host;
  GDB_TRY(arena(ctx, A_BITMAP, sizeof(uint64_t) * (nwords + 1), (void**)&dbm));
  GDB_TRY(arena(ctx, A_PARTIALS, sizeof(uint32_t) * 2 * nblocks + 64, (void**)&dcounts));
  doffs = dcounts + nblocks;
  GDB_TRY(arena(ctx, A_OUT, 64, (void**)&dtotal));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(kernel_mark(ctx, 1, s));
  select_lt_bitmap<<<nblocks, block, 0, s>>>((const int32_t*)c->dev, n, value, chunk, dbm, dcounts);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 2, s));
  scan_counts<<<1, 256, 0, s>>>(dcounts, nblocks, doffs, dtotal);
  GDB_HIP(hipMemcpyAsync(host, dtotal, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  GDB_HIP(hipStreamSynchronize(s));
  const uint64_t total = host[0];
  if (rowids && total) {
    GDB_TRY(arena(ctx, A_AUX, sizeof(uint32_t) * total, (void**)&drows));
    select_emit<<<nblocks, block, 0, s>>>(dbm, n, chunk, doffs, drows);
    GDB_HIP(hipGetLastError());
    GDB_HIP(hipMemcpyAsync(rowids, drows, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, s));
  }
  if (bitmap && nwords)
    GDB_HIP(hipMemcpyAsync(bitmap, dbm, sizeof(uint64_t) * nwords, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  const uint64_t bytes = (uint64_t)n * 4;
  GDB_TRY(timing_finish(ctx, timing, bytes, bytes + nwords * 8ull));
  *count = total;
  return GDB_OK;
}

}  // extern "C"
