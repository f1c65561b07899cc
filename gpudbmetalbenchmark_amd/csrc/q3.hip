// Q3 three-table join + group by orderkey + top-k (reference kernels/q3.metal:43-340,
// host src/q3.cpp:8-173, CPU top-10 src/infra.h:538-564, SQL sql/q3.sql).
//
// Plan (all on the engine stream, no host round trip until the k results):
//   clear     zero the customer bitmap, the counters and the revenue histogram
//   customer  bit[custkey] = (segment == s)                        (188 KB bitmap)
//   orders    qualifying = orderdate < D && bit[custkey]: append (key, date,
//             prio) to a list in row order, tag[orderkey] = this run's epoch,
//             slot[orderkey] = list position (plain stores, no atomics)
//   acc_clear zero the per-list-entry revenue and hit flags
//   probe     fact rows with shipdate > D and tag[orderkey] == epoch add
//             extprice*(100-discount) into their order's list entry; discount
//             and price are loaded only for lanes with a surviving row, and a
//             lane's consecutive rows of one order are added once
//   topk      a 4096-bucket log-linear histogram of the groups' revenues finds
//             the bucket holding the k-th best; only groups at or above it are
//             collected, and one block ranks them (revenue desc, orderdate asc,
//             orderkey asc) with the LDS top-k of topk.h
// The reference's direct map orders_map[orderkey] (kernels/q3.metal:59-78) and
// its reset before every run become the byte tags: a run only trusts entries
// tagged with its own epoch, so neither map is cleared between runs (the
// byte map is zeroed once every 255 runs). Device-scope global atomics run at
// ~10 G/s on MI355X, so the build uses none; only the probe's revenue adds are
// atomic (one per run of rows of one order, ~10^5 per run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "internal.h"
#include "kcommon.h"
#include "topk.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int TOPK_BUF = 2048;
constexpr uint32_t CAND_CAP = 65536;  // candidate list entries
constexpr int OTILE_PER_THREAD = 16;

struct Q3Scratch {
  uint32_t* cbm;      // customer bitmap
  uint32_t* counters; // [0] qualifying orders, [1] groups, [2] threshold, [3] candidates
  uint32_t* hist;     // 4096 revenue buckets
  int32_t* qkey;      // list of qualifying orders
  int32_t* qdate;
  int32_t* qprio;
  int64_t* rev;       // per list entry
  uint8_t* hit;
  uint8_t* tag;       // [maxok+1] epoch of the run that listed the key
  int32_t* slot;      // [maxok+1] list position (valid where tag == epoch)
  TopEntry* cand;     // CAND_CAP candidates
  TopEntry* out;      // k final
};

__global__ __launch_bounds__(256) void q3_clear(uint32_t* __restrict__ a, uint32_t na,
                                                uint32_t* __restrict__ b, uint32_t nb) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += stride) {
    if (i < na) a[i] = 0u;
    else b[i - na] = 0u;
  }
}

__global__ __launch_bounds__(256) void q3_fill_bytes(uint32_t* __restrict__ p, uint32_t nwords) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) p[i] = 0u;
}

__global__ __launch_bounds__(256) void q3_customer(const int32_t* __restrict__ custkey,
                                                   const uint8_t* __restrict__ seg, uint32_t n,
                                                   int32_t segment, uint32_t maxck,
                                                   uint32_t* __restrict__ cbm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t ck = (uint32_t)custkey[i];
  if (seg[i] == (uint8_t)segment && ck <= maxck) atomicOr(&cbm[ck >> 5], 1u << (ck & 31));
}

// Tiles of 16 x 256 rows: thread t reads rows 4t..4t+3 of each of the tile's
// four 1024-row quarters (coalesced dwordx4 loads, all twelve issued before
// the first use), then the 16 customer-bitmap lookups together. Qualifying
// rows are appended with one atomic per tile and a block-wide scan for the
// slots, written from the values still held in registers. The list order
// (tile, thread, quarter, row) is deterministic; nothing downstream depends on
// it.
__global__ __launch_bounds__(256) void q3_orders(const int32_t* __restrict__ okey,
                                                 const int32_t* __restrict__ ocust,
                                                 const int32_t* __restrict__ odate,
                                                 const int32_t* __restrict__ oprio, uint32_t n,
                                                 int32_t date, uint32_t maxck, uint32_t maxok,
                                                 uint8_t epoch, Q3Scratch s) {
  constexpr int R = OTILE_PER_THREAD;
  constexpr int V = R / 4;
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t base_sh;
  const uint32_t tile_rows = R * blockDim.x;
  const uint32_t ntiles = (n + tile_rows - 1) / tile_rows;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int32_t ks[R], cs[R], ds[R];
    uint32_t rows[V];
#pragma unroll
    for (int v = 0; v < V; ++v) rows[v] = t * tile_rows + v * 4 * blockDim.x + 4 * threadIdx.x;
    if (t * tile_rows + tile_rows <= n) {
      int4 k4[V], c4[V], d4[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        k4[v] = ld4(okey, rows[v] / 4);
        c4[v] = ld4(ocust, rows[v] / 4);
        d4[v] = ld4(odate, rows[v] / 4);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        ks[4 * v] = k4[v].x; ks[4 * v + 1] = k4[v].y; ks[4 * v + 2] = k4[v].z; ks[4 * v + 3] = k4[v].w;
        cs[4 * v] = c4[v].x; cs[4 * v + 1] = c4[v].y; cs[4 * v + 2] = c4[v].z; cs[4 * v + 3] = c4[v].w;
        ds[4 * v] = d4[v].x; ds[4 * v + 1] = d4[v].y; ds[4 * v + 2] = d4[v].z; ds[4 * v + 3] = d4[v].w;
      }
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t r = rows[v] + j;
          const bool in = r < n;
          ks[4 * v + j] = in ? okey[r] : 0;
          cs[4 * v + j] = in ? ocust[r] : 0;
          ds[4 * v + j] = in ? odate[r] : date;  // fails the date test
        }
      }
    }
    // Customer bits only for orders that pass the date test.
    uint32_t words[R];
#pragma unroll
    for (int j = 0; j < R; ++j) words[j] = ds[j] < date ? s.cbm[min((uint32_t)cs[j], maxck) >> 5] : 0u;
    uint32_t mask = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint32_t ck = (uint32_t)cs[j];
      const bool q = ds[j] < date && ck <= maxck && (uint32_t)ks[j] <= maxok &&
                     ((words[j] >> (ck & 31)) & 1u);
      mask |= (q ? 1u : 0u) << j;
    }
    uint32_t total;
    uint32_t pos = block_excl_scan((uint32_t)__popc(mask), scratch, &total);
    if (total == 0) continue;
    if (threadIdx.x == 0) base_sh = atomicAdd(&s.counters[0], total);
    __syncthreads();
    pos += base_sh;
    if (mask) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if ((mask >> j) & 1u) {
          const uint32_t k = (uint32_t)ks[j];
          s.qkey[pos] = ks[j];
          s.qdate[pos] = ds[j];
          s.qprio[pos] = oprio[rows[j / 4] + (j & 3)];
          s.tag[k] = epoch;
          s.slot[k] = (int32_t)pos;
          ++pos;
        }
      }
    }
    __syncthreads();  // base_sh is rewritten by the next tile
  }
}

__global__ __launch_bounds__(256) void q3_acc_clear(const uint32_t* __restrict__ counters,
                                                    int64_t* __restrict__ rev,
                                                    uint8_t* __restrict__ hit) {
  const uint32_t count = counters[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
    rev[j] = 0;
    hit[j] = 0;
  }
}

struct Q3Fact {
  const int32_t *okey, *ship, *disc;
  const int64_t* price;
};

__device__ __forceinline__ void q3_flush(int32_t key, int64_t r, const Q3Scratch& s) {
  const int32_t j = s.slot[(uint32_t)key];
  atomicAdd((unsigned long long*)&s.rev[j], (unsigned long long)r);
  s.hit[j] = 1;
}

// Rows of one 4-row group whose keys and ship dates are loaded and whose
// order tags have been fetched: discount and price are loaded only if a row
// survives, and a lane's consecutive rows of one order are added once.
__device__ __forceinline__ void q3_finish(const Q3Fact& f, uint32_t g, int4 k4, int4 s4,
                                          const uint8_t* tg, int32_t date, uint32_t maxok,
                                          uint8_t epoch, const Q3Scratch& s) {
  const int32_t ks[4] = {k4.x, k4.y, k4.z, k4.w};
  const int32_t ss[4] = {s4.x, s4.y, s4.z, s4.w};
  bool pass[4];
  bool any = false;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    pass[j] = ss[j] > date && (uint32_t)ks[j] <= maxok && tg[j] == epoch;
    any |= pass[j];
  }
  if (!any) return;
  const int4 d4 = ld4(f.disc, g);
  const longlong2 pa = ld2(f.price, 2 * g), pb = ld2(f.price, 2 * g + 1);
  const int32_t ds[4] = {d4.x, d4.y, d4.z, d4.w};
  const int64_t ps[4] = {pa.x, pa.y, pb.x, pb.y};
  int32_t run_key = 0;
  int64_t run_rev = 0;
  bool open = false;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (!pass[j]) continue;
    const int64_t r = ps[j] * (int64_t)(100 - ds[j]);
    if (open && ks[j] == run_key) {
      run_rev += r;
    } else {
      if (open) q3_flush(run_key, run_rev, s);
      run_key = ks[j];
      run_rev = r;
      open = true;
    }
  }
  if (open) q3_flush(run_key, run_rev, s);
}

// U groups of 4 rows per thread and step: all key and ship-date loads are
// issued first, then all order-tag gathers (keys are clamped into the map; a
// clamped key fails the key test), then the groups are finished.
template <int U>
__global__ __launch_bounds__(256) void q3_probe(Q3Fact f, uint32_t n, int32_t date, uint32_t maxok,
                                                uint8_t epoch, Q3Scratch s) {
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  for (; g + (U - 1) * nthr < nvec; g += U * nthr) {
    int4 k4[U], s4[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      k4[u] = ld4nt(f.okey, g + u * nthr);
      s4[u] = ld4nt(f.ship, g + u * nthr);
    }
    uint8_t tg[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      tg[u][0] = s.tag[min((uint32_t)k4[u].x, maxok)];
      tg[u][1] = s.tag[min((uint32_t)k4[u].y, maxok)];
      tg[u][2] = s.tag[min((uint32_t)k4[u].z, maxok)];
      tg[u][3] = s.tag[min((uint32_t)k4[u].w, maxok)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) q3_finish(f, g + u * nthr, k4[u], s4[u], tg[u], date, maxok, epoch, s);
  }
  for (; g < nvec; g += nthr) {
    const int4 k4 = ld4nt(f.okey, g), s4 = ld4nt(f.ship, g);
    const uint8_t tg[4] = {s.tag[min((uint32_t)k4.x, maxok)], s.tag[min((uint32_t)k4.y, maxok)],
                           s.tag[min((uint32_t)k4.z, maxok)], s.tag[min((uint32_t)k4.w, maxok)]};
    q3_finish(f, g, k4, s4, tg, date, maxok, epoch, s);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const uint32_t r = nvec * 4 + threadIdx.x;
    const int32_t k = f.okey[r];
    if (f.ship[r] > date && (uint32_t)k <= maxok && s.tag[(uint32_t)k] == epoch)
      q3_flush(k, f.price[r] * (int64_t)(100 - f.disc[r]), s);
  }
}

// Top-k, step 1: histogram of the groups' revenue buckets (LDS per block,
// merged by global atomics). Its total is the group count (step 2 records it).
__global__ __launch_bounds__(256) void q3_topk_hist(Q3Scratch s) {
  __shared__ uint32_t lh[4096];
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  const uint32_t count = s.counters[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride)
    if (s.hit[j]) atomicAdd(&lh[value_bucket(s.rev[j])], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < 4096; i += blockDim.x)
    if (lh[i]) atomicAdd(&s.hist[i], lh[i]);
}

// Step 2 (one block): the highest bucket b such that buckets >= b hold at
// least k groups (0 when there are fewer than k groups), and the group count
// (every group with revenue is in exactly one bucket).
__global__ __launch_bounds__(256) void q3_topk_threshold(int k, Q3Scratch s) {
  __shared__ uint32_t part[256];
  uint32_t v = 0;
  for (int i = 0; i < 16; ++i) v += s.hist[threadIdx.x * 16 + i];
  part[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0, b = 0;
    for (int t = 255; t >= 0; --t) {
      if (acc + part[t] >= (uint32_t)k) {
        for (int i = 15; i >= 0; --i) {
          acc += s.hist[t * 16 + i];
          if (acc >= (uint32_t)k) {
            b = (uint32_t)(t * 16 + i);
            break;
          }
        }
        break;
      }
      acc += part[t];
    }
    uint32_t groups = 0;
    for (int t = 0; t < 256; ++t) groups += part[t];
    s.counters[1] = groups;
    s.counters[2] = b;
    s.counters[3] = 0u;
  }
}

// Step 3: groups in buckets >= threshold -> candidate list (tile-aggregated append).
__global__ __launch_bounds__(256) void q3_topk_collect(Q3Scratch s) {
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t base_sh;
  const uint32_t count = s.counters[0];
  const uint32_t thr = s.counters[2];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t span = (count + stride - 1) / stride * stride;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < span; j += stride) {
    const bool q = j < count && s.hit[j] && value_bucket(s.rev[j]) >= thr;
    uint32_t total;
    const uint32_t pre = block_excl_scan(q ? 1u : 0u, scratch, &total);
    if (total == 0) continue;
    if (threadIdx.x == 0) base_sh = atomicAdd(&s.counters[3], total);
    __syncthreads();
    const uint32_t pos = base_sh + pre;
    if (q && pos < CAND_CAP) s.cand[pos] = TopEntry{s.rev[j], s.qdate[j], s.qkey[j], s.qprio[j], 0};
    __syncthreads();
  }
}

// Step 4 (one block): rank the candidates; if more than fit the list (heavy
// ties in the threshold bucket), rank the qualifying list entries directly.
__global__ __launch_bounds__(256) void q3_topk_final(int k, Q3Scratch s) {
  __shared__ TopEntry buf[TOPK_BUF];
  __shared__ int nbuf;
  BlockTopK<TOPK_BUF> tk{buf, &nbuf, k};
  tk.init();
  const uint32_t nc = s.counters[3];
  if (nc <= CAND_CAP) {
    for (uint32_t base = 0; base < nc; base += blockDim.x) {
      const uint32_t i = base + threadIdx.x;
      TopEntry e;
      const bool valid = i < nc;
      if (valid) e = s.cand[i];
      tk.push(valid, e);
    }
  } else {
    const uint32_t count = s.counters[0], thr = s.counters[2];
    for (uint32_t base = 0; base < count; base += blockDim.x) {
      const uint32_t j = base + threadIdx.x;
      TopEntry e;
      bool valid = false;
      if (j < count && s.hit[j] && value_bucket(s.rev[j]) >= thr) {
        e = TopEntry{s.rev[j], s.qdate[j], s.qkey[j], s.qprio[j], 0};
        valid = true;
      }
      if (__syncthreads_or(valid)) tk.push(valid, e);
    }
  }
  tk.finish(s.out);
}

}  // namespace

extern "C" int gdb_q3(gdb_ctx* ctx, const gdb_table* customer, const gdb_table* orders,
                      const gdb_table* lineitem, const gdb_q3_params* p, gdb_q3_result* out,
                      gdb_timing* timing) {
  if (!ctx || !customer || !orders || !lineitem || !p || !out)
    return fail(GDB_E_INVALID, "null argument");
  if (p->top_k < 1 || p->top_k > 100) return fail(GDB_E_INVALID, "top_k must be in 1..100");
  const Column *cck, *cseg, *cok, *coc, *cod, *cop, *clk, *cls, *cld, *clp;
  GDB_TRY(column(customer, "c_custkey", GDB_I32, &cck));
  GDB_TRY(column(customer, "c_mktsegment", GDB_U8, &cseg));
  GDB_TRY(column(orders, "o_orderkey", GDB_I32, &cok));
  GDB_TRY(column(orders, "o_custkey", GDB_I32, &coc));
  GDB_TRY(column(orders, "o_orderdate", GDB_I32, &cod));
  GDB_TRY(column(orders, "o_shippriority", GDB_I32, &cop));
  GDB_TRY(column(lineitem, "l_orderkey", GDB_I32, &clk));
  GDB_TRY(column(lineitem, "l_shipdate", GDB_I32, &cls));
  GDB_TRY(column(lineitem, "l_discount", GDB_I32, &cld));
  GDB_TRY(column(lineitem, "l_extendedprice", GDB_I64, &clp));
  GDB_TRY(require_resident(customer, "Q3 customer"));
  GDB_TRY(require_resident(orders, "Q3 orders"));
  if ((customer->nrows && cck->minv < 0) || (orders->nrows && cok->minv < 0))
    return fail(GDB_E_UNSUPPORTED, "Q3 needs non-negative custkey and orderkey values");
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t nc = (uint32_t)customer->nrows, no = (uint32_t)orders->nrows;
  const uint64_t nl = lineitem->nrows;
  const uint32_t maxck = (uint32_t)std::max<int64_t>(0, cck->maxv);
  const uint32_t maxok = (uint32_t)std::max<int64_t>(0, cok->maxv);
  const uint32_t ncw = maxck / 32 + 1;
  const int k = p->top_k;
  const uint32_t nlist = std::max<uint32_t>(no, 1);

  Q3Scratch s{};
  {
    char* b;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) & ~size_t(255);
      return o;
    };
    const size_t o_cbm = take(ncw * 4ull), o_cnt = take((16 + 4096) * 4), o_qk = take(nlist * 4ull),
                 o_qd = take(nlist * 4ull), o_qp = take(nlist * 4ull), o_rev = take(nlist * 8ull),
                 o_hit = take(nlist);
    GDB_TRY(arena(ctx, A_BITMAP, off, (void**)&b));
    s.cbm = (uint32_t*)(b + o_cbm);
    s.counters = (uint32_t*)(b + o_cnt);
    s.hist = s.counters + 16;
    s.qkey = (int32_t*)(b + o_qk);
    s.qdate = (int32_t*)(b + o_qd);
    s.qprio = (int32_t*)(b + o_qp);
    s.rev = (int64_t*)(b + o_rev);
    s.hit = (uint8_t*)(b + o_hit);
    const size_t tag_bytes = ((size_t)maxok + 1 + 255) & ~size_t(255);
    GDB_TRY(arena(ctx, A_Q3BYTE, tag_bytes, (void**)&s.tag));
    GDB_TRY(arena(ctx, A_Q3MAP, ((size_t)maxok + 1) * 4, (void**)&s.slot));
    char* a;
    GDB_TRY(arena(ctx, A_AUX, sizeof(TopEntry) * ((size_t)CAND_CAP + k) + 64, (void**)&a));
    s.cand = (TopEntry*)a;
    s.out = s.cand + CAND_CAP;
  }
  TopEntry* host;
  GDB_TRY(pinned(ctx, sizeof(TopEntry) * k + 64, (void**)&host));
  uint32_t* host_cnt = (uint32_t*)(host + k);

  hipStream_t st = ctx->stream;
  const int B = 256;
  const int fill_blocks = ctx->info.cu_count * 4;
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  // Epoch of this run in the tag map; a new or reused-after-wrap map is zeroed.
  const size_t tag_bytes = ((size_t)maxok + 1 + 255) & ~size_t(255);
  if (ctx->q3_map != (void*)s.tag || ctx->q3_map_bytes < tag_bytes || ctx->q3_epoch == 255) {
    const size_t alloc_bytes = ctx->arenas[A_Q3BYTE].bytes;
    q3_fill_bytes<<<fill_blocks, B, 0, st>>>((uint32_t*)s.tag, (uint32_t)(alloc_bytes / 4));
    ctx->q3_map = s.tag;
    ctx->q3_map_bytes = alloc_bytes;
    ctx->q3_epoch = 0;
  }
  const uint8_t epoch = ++ctx->q3_epoch;
  q3_clear<<<std::min<uint32_t>(2048, (ncw + 16 + 4096 + B - 1) / B + 1), B, 0, st>>>(
      s.cbm, ncw, s.counters, 16 + 4096);
  if (nc)
    q3_customer<<<(nc + B - 1) / B, B, 0, st>>>((const int32_t*)cck->dev, (const uint8_t*)cseg->dev,
                                                 nc, p->segment, maxck, s.cbm);
  if (no)
    q3_orders<<<std::min<uint32_t>(ctx->info.cu_count * 8, (no + 16 * B - 1) / (16 * B)), B, 0, st>>>(
        (const int32_t*)cok->dev, (const int32_t*)coc->dev, (const int32_t*)cod->dev,
        (const int32_t*)cop->dev, no, p->date, maxck, maxok, epoch, s);
  q3_acc_clear<<<fill_blocks, B, 0, st>>>(s.counters, s.rev, s.hit);
  GDB_HIP(hipGetLastError());
  const int president =
      resident_blocks(ctx, (const void*)q3_probe<4>, B, 0, ctx->tune[T_Q3].blocks_per_cu);
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const Q3Fact f{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                   (const int64_t*)d[3]};
    const int pblocks = std::max(1, std::min<int>(president, (int)((rows / 4 + B - 1) / B)));
    q3_probe<4><<<pblocks, B, 0, st>>>(f, rows, p->date, maxok, epoch, s);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  const Column* fc[4] = {clk, cls, cld, clp};
  RunStats rs;
  if (nl) {
    GDB_TRY(run_over(ctx, lineitem, fc, 4, consume, &rs));
  } else {
    GDB_HIP(hipEventRecord(ctx->ev[1], st));
    GDB_HIP(hipEventRecord(ctx->ev[2], st));
  }
  q3_topk_hist<<<ctx->info.cu_count, B, 0, st>>>(s);  // one per CU: fewer bin merges
  q3_topk_threshold<<<1, B, 0, st>>>(k, s);
  q3_topk_collect<<<fill_blocks, B, 0, st>>>(s);
  q3_topk_final<<<1, B, 0, st>>>(k, s);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, s.out, sizeof(TopEntry) * k, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipMemcpyAsync(host_cnt, s.counters, 8, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  // Pipeline bytes: every column the query names, read once (customer 5 B,
  // orders 16 B, fact 20 B per row). Probe bytes: what the probe must read
  // whatever the data, l_orderkey and l_shipdate of every row plus one pass
  // over the orderkey tag map; discount and price are read only for 4-row
  // groups holding a surviving row and are not counted.
  const uint64_t bytes = 5ull * nc + 16ull * no + 20ull * nl;
  const uint64_t probe_bytes = 8ull * nl + (nl ? (uint64_t)maxok + 1 : 0);
  GDB_TRY(timing_finish(ctx, timing, bytes, probe_bytes, &rs));
  memset(out, 0, sizeof(*out));
  out->groups = host_cnt[1];
  for (int i = 0; i < k; ++i) {
    if (host[i].rev == INT64_MIN) break;
    gdb_q3_row& r = out->rows[out->nrows++];
    r.orderkey = host[i].key;
    r.orderdate = host[i].date;
    r.shippriority = host[i].prio;
    r.revenue = host[i].rev;
  }
  return GDB_OK;
}
