// Shared host and device pieces of the TPC-H pipelines in tpch_*.hip.
//
// Layout and strategy (all pipelines):
//  - Dimension predicates become direct-address maps over dense integer keys:
//    bitmaps (1 bit per key, built with atomicOr) for yes/no tests, byte maps
//    (plain stores) for small payloads such as a nation or a year, int32 /
//    int64 maps for keys and sums. A map covers [0, max key] of the column it
//    is built from (the range is recorded when the table is uploaded).
//  - Fact scans stream their columns with nontemporal dwordx4 loads, 4 rows
//    per lane per vector group, over ceil(n/4) groups: the last group reads the
//    zeroed column padding and masks rows >= n. The most selective column is
//    loaded first and the rest only for groups with a surviving row, so
//    64-byte lines with no survivor are never fetched.
//  - Small group tables (<= 64 bins) live in LDS per workgroup and are merged
//    with one global atomic per non-empty bin.
//  - Top-k uses the LDS block top-k of topk.h in two stages.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <string>
#include <vector>

#include "gdb_tpch.h"
#include "internal.h"
#include "kcommon.h"
#include "topk.h"

namespace gdbt {

using gdb::Column;

constexpr int BLOCK = 256;

// ---- host ------------------------------------------------------------------

// Column `name` of table `t` (the `what` table of the query), with type check.
inline int col(const gdb_table* t, const char* what, const char* name, int type,
               const Column** out) {
  if (!t) return gdb::fail(GDB_E_INVALID, std::string("missing the ") + what + " table");
  GDB_TRY(gdb::column(t, name, type, out));
  return gdb::require_resident(t, what);
}

// A lineitem column: the table may be host-resident (streamed in chunks).
inline int scol(const gdb_table* t, const char* what, const char* name, int type,
                const Column** out) {
  if (!t) return gdb::fail(GDB_E_INVALID, std::string("missing the ") + what + " table");
  return gdb::column(t, name, type, out);
}

// Runs `fn(dev, rows)` over columns of table t: once over the resident columns
// (bracketed by the dominant-kernel events ev[1], ev[2]) or, for a host table,
// chunk by chunk through two device slots with copies overlapped (the reference's
// chunkedStreamLoop). Chunk times add into *rs.
inline int scan_table(gdb_ctx* ctx, const gdb_table* t, std::initializer_list<const Column*> cols,
                      const gdb::Consume& fn, gdb::RunStats* rs) {
  if (t->nrows == 0) {
    GDB_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    GDB_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    return GDB_OK;
  }
  std::vector<const Column*> v(cols);
  gdb::RunStats one;
  GDB_TRY(gdb::run_over(ctx, t, v.data(), (int)v.size(), fn, &one));
  rs->kernel_ms += one.kernel_ms;
  rs->h2d_bytes += one.h2d_bytes;
  rs->chunks += one.chunks;
  return GDB_OK;
}

// Entries of a direct-address map over key column c: max key + 1. Keys must be
// non-negative and below 2^31 - 1.
inline int key_span(const Column* c, uint64_t nrows, uint64_t* span) {
  if (nrows == 0 || c->maxv < c->minv) {
    *span = 1;
    return GDB_OK;
  }
  if (c->minv < 0 || c->maxv >= 0x7FFFFFFELL)
    return gdb::fail(GDB_E_UNSUPPORTED, "key column " + c->name + " outside [0, 2^31-2]");
  *span = (uint64_t)c->maxv + 1;
  return GDB_OK;
}

// Workgroups for a grid-stride scan of `groups` vector groups.
inline int scan_blocks(gdb_ctx* ctx, const void* kernel, uint64_t groups, size_t lds = 0) {
  const int res = gdb::resident_blocks(ctx, kernel, BLOCK, lds, 0);
  const uint64_t need = (groups + BLOCK - 1) / BLOCK;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)res, need));
}

__host__ __device__ inline uint32_t quads(uint64_t n) { return (uint32_t)((n + 3) / 4); }

// ---- device ------------------------------------------------------------------

__device__ __forceinline__ bool bit_test(const uint32_t* __restrict__ bm, uint32_t k) {
  return (bm[k >> 5] >> (k & 31)) & 1u;
}
// Map entries at the four keys of a quad, loaded together so their latencies
// overlap (a per-row load under a per-row branch waits once per row). Every key
// must index the map, including rows that fail earlier tests and padding rows
// (key 0): maps are sized over the whole key column.
template <class T>
__device__ __forceinline__ void gather4(const T* __restrict__ map, const int4& k, uint32_t* v) {
  v[0] = map[(uint32_t)k.x];
  v[1] = map[(uint32_t)k.y];
  v[2] = map[(uint32_t)k.z];
  v[3] = map[(uint32_t)k.w];
}
// Bit j set when bit k.j of the bitmap is set (four loads together).
__device__ __forceinline__ uint32_t bits4(const uint32_t* __restrict__ bm, const int4& k) {
  const uint32_t w0 = bm[(uint32_t)k.x >> 5], w1 = bm[(uint32_t)k.y >> 5],
                 w2 = bm[(uint32_t)k.z >> 5], w3 = bm[(uint32_t)k.w >> 5];
  return ((w0 >> (k.x & 31)) & 1u) | (((w1 >> (k.y & 31)) & 1u) << 1) |
         (((w2 >> (k.z & 31)) & 1u) << 2) | (((w3 >> (k.w & 31)) & 1u) << 3);
}
__device__ __forceinline__ void bit_set(uint32_t* bm, uint32_t k) {
  atomicOr(&bm[k >> 5], 1u << (k & 31));
}
__device__ __forceinline__ void add_i64(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
__device__ __forceinline__ void lds_add_i64(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
__device__ __forceinline__ uint8_t byte_of(uint32_t w, int j) { return (uint8_t)(w >> (8 * j)); }
__device__ __forceinline__ int lane_of(const int4& v, int j) {
  return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}
__device__ __forceinline__ int64_t lane_of(const longlong2& a, const longlong2& b, int j) {
  return j == 0 ? a.x : j == 1 ? a.y : j == 2 ? b.x : b.y;
}

// Day number (1992-01-01 = 0) -> year - 1992, for days in [0, 2922).
__device__ __forceinline__ int year_index(int32_t day) {
  // 1992 and 1996 are leap years; cumulative year starts from day 0.
  const int starts[8] = {0, 366, 731, 1096, 1461, 1827, 2192, 2557};
  int y = 0;
#pragma unroll
  for (int i = 1; i < 8; ++i) y += day >= starts[i];
  return y;
}

// Grid-stride walk over nq vector groups, two groups per trip: the first-stage
// loads of both groups (`first(g)`, the streamed columns every row needs) are
// issued before either group is processed (`rest(g, loaded)`, which may load
// more for surviving rows), so two independent loads are in flight per lane.
template <class First, class Rest>
__device__ __forceinline__ void walk2(uint32_t nq, First first, Rest rest) {
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  for (; g + stride < nq; g += 2 * stride) {
    const auto a = first(g);
    const auto b = first(g + stride);
    rest(g, a);
    rest(g + stride, b);
  }
  if (g < nq) rest(g, first(g));
}

// Grid-stride walk over nq vector groups, QB groups per lane per trip (g,
// g + stride, ...): `body(g, live)` runs each load phase for all QB groups
// before using any of them, so a lane has QB groups' loads in flight per phase.
// Groups past nq are not live (their loads must be skipped).
template <int QB, class Body>
__device__ __forceinline__ void walk_batch(uint32_t nq, Body body) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t g0 = blockIdx.x * blockDim.x + threadIdx.x; g0 < nq; g0 += QB * stride) {
    uint32_t g[QB];
    bool live[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      g[b] = g0 + (uint32_t)b * stride;
      live[b] = g[b] < nq;
    }
    body(g, live);
  }
}

// Spread accumulators. Atomics from every workgroup on one address serialise
// (measured on MI355X: 2048 workgroups x 5 int64 counters on one line cost about
// 120 us). A kernel that accumulates into a small result therefore adds into copy
// blockIdx.x % ACC_COPIES of it: the copies lie ACC_STRIDE int64 (256 bytes) apart,
// so at most 32 workgroups share an address. acc_fold then combines the copies into
// copy 0 (sum, OR or unsigned max per value). A result of nv values needs a buffer
// of ACC_COPIES * ACC_STRIDE int64 cleared first, and nv <= ACC_STRIDE.
constexpr int ACC_COPIES = 64, ACC_STRIDE = 32;
constexpr size_t ACC_BYTES = (size_t)ACC_COPIES * ACC_STRIDE * 8;

template <class T>
__device__ __forceinline__ T* acc_copy(T* out) {
  return out + (size_t)(blockIdx.x % ACC_COPIES) * ACC_STRIDE * (8 / sizeof(T));
}

// Values [lo, hi) of the fold of all copies go to dst[lo..hi) (dst = acc: into
// copy 0); bit i of or_mask / max_mask selects OR / unsigned max for value i (sum
// otherwise). One wave per value: lane c reads copy c, then a shuffle reduction.
static_assert(ACC_COPIES == 64, "acc_fold reads one copy per lane of a wave");
namespace {
__global__ __launch_bounds__(64) void acc_fold(int64_t* __restrict__ acc, int lo,
                                               uint32_t or_mask, uint32_t max_mask,
                                               int64_t* __restrict__ dst) {
  const int i = lo + (int)blockIdx.x;
  const bool is_or = (or_mask >> i) & 1, is_max = (max_mask >> i) & 1;
  uint64_t v = (uint64_t)acc[(size_t)threadIdx.x * ACC_STRIDE + i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t x = (uint64_t)__shfl_xor((long long)v, off, 64);
    v = is_or ? (v | x) : is_max ? (x > v ? x : v) : v + x;
  }
  if (threadIdx.x == 0) dst[i] = (int64_t)v;
}
}  // namespace

inline int acc_finish(gdb_ctx* ctx, int64_t* acc, int lo, int hi, uint32_t or_mask = 0,
                      uint32_t max_mask = 0, int64_t* dst = nullptr) {
  if (hi <= lo) return GDB_OK;
  acc_fold<<<hi - lo, 64, 0, ctx->stream>>>(acc, lo, or_mask, max_mask, dst ? dst : acc);
  GDB_HIP(hipGetLastError());
  return GDB_OK;
}

// Block sum of `v` added to *out (one atomic per workgroup).
__device__ __forceinline__ void block_add(int64_t v, int64_t* red, int64_t* out) {
  const int64_t s = gdbk::block_sum(v, red);
  if (threadIdx.x == 0 && s) add_i64(out, s);
}

// Top-k sources are functors `bool (uint64_t i, gdbk::TopEntry& e)`: candidate i's
// entry and whether it is valid. They load unconditionally (every index below n
// is valid) and return the verdict, with no early exit.

// Threshold top-k: a histogram of the candidates' value buckets (gdbk::value_bucket,
// monotone in the revenue), the bucket b the k-th best falls in, then only the
// candidates in buckets >= b are collected and ranked by one workgroup. The
// histogram's total is the number of valid candidates. If heavy ties in bucket b
// overflow the candidate list, every workgroup of the ranking grid ranks its
// share of the candidates and the last one to finish ranks their winners.
constexpr uint32_t THK_BUCKETS = 4096, THK_CAP = 8192;
struct ThkState {   // device memory, cleared before thk_hist
  uint32_t hist[THK_BUCKETS];
  uint32_t thr, ncand;
  unsigned long long total;
  uint32_t done;    // ranking workgroups finished (overflow path)
};

template <class Src>
__global__ __launch_bounds__(256) void thk_hist(Src src, uint64_t n, const uint32_t* __restrict__ dn,
                                                ThkState* __restrict__ st) {
  __shared__ uint32_t lh[THK_BUCKETS];
  if (dn) n = n < *dn ? n : *dn;
  for (uint32_t i = threadIdx.x; i < THK_BUCKETS; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    gdbk::TopEntry e;
    if (src(i, e)) atomicAdd(&lh[gdbk::value_bucket(e.rev)], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < THK_BUCKETS; i += blockDim.x)
    if (lh[i]) atomicAdd(&st->hist[i], lh[i]);
}

namespace {
// One workgroup of 256: the highest bucket b with at least k candidates in
// buckets >= b (0 when there are fewer than k), and the candidate total.
// Thread t owns buckets [16t, 16t + 16); a block scan gives each thread the
// candidates above its range, so one thread finds b without a serial walk.
__global__ __launch_bounds__(256) void thk_threshold(int k, ThkState* __restrict__ st,
                                                     unsigned long long* __restrict__ count) {
  __shared__ uint32_t scratch[17];
  uint32_t c[16];
  uint32_t part = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    c[i] = st->hist[threadIdx.x * 16 + i];
    part += c[i];
  }
  uint32_t total;
  const uint32_t pre = gdbk::block_excl_scan(part, scratch, &total);
  const uint32_t at = total - pre, above = at - part;  // in buckets >= 16t, >= 16(t+1)
  if (at >= (uint32_t)k && above < (uint32_t)k) {
    uint32_t acc = above, b = threadIdx.x * 16;
    for (int i = 15; i >= 0; --i) {
      acc += c[i];
      if (acc >= (uint32_t)k) {
        b = threadIdx.x * 16 + i;
        break;
      }
    }
    st->thr = b;
  }
  if (threadIdx.x == 0) {
    if (total < (uint32_t)k) st->thr = 0u;
    st->ncand = 0u;
    st->total = total;
    if (count) *count = total;
  }
}
}  // namespace

// Candidates in buckets >= thr -> cand (one atomic per wave and trip).
template <class Src>
__global__ __launch_bounds__(256) void thk_collect(Src src, uint64_t n,
                                                   const uint32_t* __restrict__ dn,
                                                   ThkState* __restrict__ st,
                                                   gdbk::TopEntry* __restrict__ cand) {
  if (dn) n = n < *dn ? n : *dn;
  const uint32_t thr = st->thr;
  const int lane = threadIdx.x & 63;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; w0 < n; w0 += stride) {
    const uint64_t i = w0 + lane;
    gdbk::TopEntry e;
    const bool q = i < n && src(i, e) && gdbk::value_bucket(e.rev) >= thr;
    const unsigned long long m = __ballot(q);
    if (!m) continue;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&st->ncand, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    const uint32_t pos = base + __popcll(m & ((1ull << lane) - 1));
    if (q && pos < THK_CAP) cand[pos] = e;
  }
}

// Normally workgroup 0 ranks the candidate list into out[0, k) and the
// others return. After an overflow each workgroup ranks its share of the
// candidates in buckets >= thr, leaves its k best in cand[b*k, (b+1)*k) (the
// grid is at most THK_CAP / k workgroups), and the workgroup that finishes
// last ranks those into out[0, k).
template <class Src>
__global__ __launch_bounds__(256) void thk_final(Src src, uint64_t n, const uint32_t* __restrict__ dn,
                                                 int k, ThkState* __restrict__ st,
                                                 gdbk::TopEntry* __restrict__ cand,
                                                 gdbk::TopEntry* __restrict__ out) {
  __shared__ gdbk::TopEntry buf[1024];
  __shared__ gdbk::TopEntry thr;
  __shared__ int nbuf, have;
  gdbk::BlockTopK<1024> tk{buf, &nbuf, k, &thr, &have};
  if (dn) n = n < *dn ? n : *dn;
  const uint32_t nc = st->ncand;
  if (nc <= THK_CAP) {
    if (blockIdx.x != 0) return;
    tk.init();
    for (uint32_t b0 = 0; b0 < nc; b0 += blockDim.x) {
      const uint32_t i = b0 + threadIdx.x;
      const bool v = i < nc;
      tk.push(v, v ? cand[i] : gdbk::worst_entry());
    }
    tk.finish(out);
    return;
  }
  tk.init();
  const uint32_t tb = st->thr;
  const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  for (uint64_t b0 = lo; b0 < hi; b0 += blockDim.x) {
    const uint64_t i = b0 + threadIdx.x;
    gdbk::TopEntry e = gdbk::worst_entry();
    const bool v = i < hi && src(i, e) && gdbk::value_bucket(e.rev) >= tb;
    tk.push(v, e);
  }
  tk.finish(cand + (size_t)blockIdx.x * k);
  if (!gdbk::last_block(&st->done)) return;
  tk.init();
  const uint32_t m = gridDim.x * (uint32_t)k;
  for (uint32_t b0 = 0; b0 < m; b0 += blockDim.x) {
    const uint32_t i = b0 + threadIdx.x;
    gdbk::TopEntry e = gdbk::worst_entry();
    if (i < m) e = cand[i];
    // a workgroup's list is padded with worst entries when it had fewer than k
    tk.push(i < m && !(e.rev == INT64_MIN && e.key == INT32_MAX), e);
  }
  tk.finish(out);
}

// The state and the THK_CAP candidate entries, in the arena slot `slot`.
inline int thk_buffers(gdb_ctx* ctx, int slot, ThkState** state, gdbk::TopEntry** cand) {
  constexpr size_t head = (sizeof(ThkState) + 255) & ~size_t(255);
  char* p;
  GDB_TRY(gdb::arena(ctx, slot, head + sizeof(gdbk::TopEntry) * (size_t)THK_CAP, (void**)&p));
  *state = (ThkState*)p;
  *cand = (gdbk::TopEntry*)(p + head);
  return GDB_OK;
}

// Runs the four steps; `state` points to a ThkState and `cand` to THK_CAP
// entries (device memory). *count receives the number of valid candidates. With
// `dn`, n is cut to *dn (a count another kernel wrote). The histogram takes one
// workgroup per 4096 candidates, at most four per CU (its LDS merge adds only
// the buckets a workgroup touched).
template <class Src>
int thk_run(gdb_ctx* ctx, Src src, uint64_t n, int k, ThkState* state, gdbk::TopEntry* cand,
            gdbk::TopEntry* final_out, unsigned long long* count, const uint32_t* dn = nullptr) {
  hipStream_t s = ctx->stream;
  GDB_TRY(gdb::fill(state, 0, sizeof(ThkState), s));
  const int hb = (int)std::max<uint64_t>(
      1, std::min<uint64_t>(4ull * ctx->info.cu_count, (n + 4095) / 4096));
  thk_hist<Src><<<hb, BLOCK, 0, s>>>(src, n, dn, state);
  thk_threshold<<<1, 256, 0, s>>>(std::max(k, 1), state, count);
  if (k > 0) {
    thk_collect<Src><<<hb * 4, BLOCK, 0, s>>>(src, n, dn, state, cand);
    const int fb = std::max(1, std::min<int>(ctx->info.cu_count, (int)(THK_CAP / (uint32_t)k)));
    thk_final<Src><<<fb, BLOCK, 0, s>>>(src, n, dn, k, state, cand, final_out);
  }
  GDB_HIP(hipGetLastError());
  return GDB_OK;
}

// Folded filter: bit k of a key bitmap is ORed into bit (k & 31) of word
// (k >> 5) mod FOLD_WORDS (2^18 bits, 32 KiB). Kept in LDS by a scan, a clear
// folded bit rules a key out without a memory access; a set one is confirmed
// against the full bitmap. For a selective key set (a few % of parts) the
// folded bits stay mostly clear.
constexpr uint32_t FOLD_WORDS = 8192;
constexpr int FOLD_BLOCK = 512;  // threads of a scan with the 32 KiB LDS filter

__device__ __forceinline__ void fold_set(uint32_t* fold, uint32_t k) {
  atomicOr(&fold[(k >> 5) & (FOLD_WORDS - 1)], 1u << (k & 31));
}
__device__ __forceinline__ bool fold_test(const uint32_t* lfold, uint32_t k) {
  return (lfold[(k >> 5) & (FOLD_WORDS - 1)] >> (k & 31)) & 1u;
}
// Copies the folded filter into LDS; every thread of the block must call it.
__device__ __forceinline__ void fold_load(uint32_t* lfold, const uint32_t* __restrict__ fold) {
  for (uint32_t i = threadIdx.x; i < FOLD_WORDS / 4; i += blockDim.x)
    reinterpret_cast<uint4*>(lfold)[i] = reinterpret_cast<const uint4*>(fold)[i];
  __syncthreads();
}

// Workgroups of FOLD_BLOCK threads for a scan with the LDS filter.
inline int fold_blocks(gdb_ctx* ctx, const void* kernel, uint64_t groups) {
  const int res = gdb::resident_blocks(ctx, kernel, FOLD_BLOCK, 0, 0);
  const uint64_t need = (groups + FOLD_BLOCK - 1) / FOLD_BLOCK;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)res, need));
}

// Launch helper: grid-stride over n items, at most the resident workgroups.
template <typename K>
int blocks_for(gdb_ctx* ctx, K kernel, uint64_t items) {
  return scan_blocks(ctx, (const void*)kernel, items);
}

inline int args_ok(gdb_ctx* ctx, const void* db, const void* p, const void* out) {
  if (!ctx || !db || !p || !out) return gdb::fail(GDB_E_INVALID, "null argument");
  if (hipSetDevice(ctx->device) != hipSuccess) return gdb::fail(GDB_E_HIP, "hipSetDevice failed");
  return GDB_OK;
}

// A set of dictionary codes 0..255 as a bit mask.
struct Mask4 {
  uint64_t w[4];
  __device__ __forceinline__ bool has(uint32_t c) const { return (w[c >> 6] >> (c & 63)) & 1u; }
};
inline Mask4 mask4(const uint64_t* m3) { return Mask4{{m3[0], m3[1], m3[2], 0}}; }

}  // namespace gdbt

namespace {

// bits[key[i]] = mask contains code[i] (part type / container filters).
// With fold != nullptr the folded filter is filled too.
__global__ __launch_bounds__(256) void code_bits(const int32_t* __restrict__ key,
                                                 const uint8_t* __restrict__ code, uint32_t n,
                                                 gdbt::Mask4 mask, uint32_t* __restrict__ bits,
                                                 uint32_t* __restrict__ fold = nullptr) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (mask.has(code[i])) {
      gdbt::bit_set(bits, (uint32_t)key[i]);
      if (fold) gdbt::fold_set(fold, (uint32_t)key[i]);
    }
}

// nreg[n_nationkey] = n_regionkey (255 for keys without a row); nkeys < 256.
__global__ void nation_regions(const int32_t* __restrict__ nk, const int32_t* __restrict__ rk,
                               uint32_t n, uint8_t* __restrict__ nreg) {
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) nreg[i] = 255;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    if (nk[i] >= 0 && nk[i] < 256) nreg[nk[i]] = (uint8_t)(rk[i] >= 0 && rk[i] < 255 ? rk[i] : 254);
}

// map[key[i]] = val[i] < 255 ? val[i] : 255 (a nation per customer / supplier).
__global__ __launch_bounds__(256) void nation_bytes(const int32_t* __restrict__ key,
                                                    const int32_t* __restrict__ val, uint32_t n,
                                                    uint8_t* __restrict__ map) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    map[key[i]] = (uint8_t)(val[i] >= 0 && val[i] < 255 ? val[i] : 255);
}

}  // namespace
