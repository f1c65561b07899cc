// Q3 three-table join + group by orderkey + top-k (reference kernels/q3.metal:43-340,
// host src/q3.cpp:8-173, CPU top-10 src/infra.h:538-564, SQL sql/q3.sql).
//
// Plan (all on the engine stream, no host round trip until the k results):
//   clear    zero the two bitmaps and the counters
//   customer bit[custkey] = (segment == s)                         (188 KB bitmap)
//   orders   qualifying = orderdate < D && bit[custkey]:
//            set obit[orderkey], append (key, date, prio) to a compact list
//   ht_clear / ht_insert
//            open-addressing table over the qualifying orders, capacity
//            nextpow2(2*count) computed on the device from the list count
//   probe    fact rows with shipdate > D and obit[orderkey] add
//            extprice*(100-discount) into their order's slot; discount and
//            price are loaded only for lanes with a surviving row
//   topk     a 4096-bucket log-linear histogram of the groups' revenues finds
//            the bucket holding the k-th best; only groups at or above it are
//            collected, and one block ranks them (revenue desc, orderdate asc,
//            orderkey asc) with the LDS top-k of topk.h
// The orders bitmap (maxkey/8 bytes) is the reference's bitmap pre-filter
// (kernels/q3.metal:72-74) moved onto the probe side, where it rejects ~99%
// of fact rows before any hash-table access; the table holds only qualifying
// orders, so it stays resident in the 256 MB Infinity Cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "internal.h"
#include "kcommon.h"
#include "topk.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int32_t EMPTY = INT32_MIN;
constexpr int TOPK_BUF = 2048;
constexpr uint32_t CAND_CAP = 65536;  // candidate list entries

__device__ __forceinline__ uint32_t ht_cap(uint32_t count) {
  uint32_t c = 1024;
  while (c < 2 * count) c <<= 1;
  return c;
}
// Folds the key's high bits onto its low bits: keys below the capacity map to
// themselves, so orders appended and probed in key order touch neighbouring
// slots (the reference hashes by key modulo size, kernels/q3.metal:189).
__device__ __forceinline__ uint32_t ht_hash(int32_t key, uint32_t cap) {
  const uint32_t k = (uint32_t)key;
  const int lg = __ffs((int)cap) - 1;
  return (k ^ (k >> lg) ^ (k >> (2 * lg < 32 ? 2 * lg : 31))) & (cap - 1);
}
__device__ __forceinline__ bool bit_test(const uint32_t* __restrict__ bm, uint32_t k) {
  return (bm[k >> 5] >> (k & 31)) & 1u;
}

struct Q3Scratch {
  uint32_t* cbm;      // customer bitmap
  uint32_t* obm;      // qualifying-orders bitmap
  uint32_t* counters; // [0] qualifying orders, [1] groups
  int32_t* qkey;      // compact list of qualifying orders
  int32_t* qdate;
  int32_t* qprio;
  int32_t* ht_key;
  int32_t* ht_q;      // slot -> list index
  int64_t* ht_rev;
  uint32_t* ht_hit;
  uint32_t* hist;     // 4096 revenue buckets
  TopEntry* cand;     // CAND_CAP candidates
  TopEntry* out;      // k final
};

__global__ __launch_bounds__(256) void q3_clear(uint32_t* __restrict__ a, uint32_t na,
                                                uint32_t* __restrict__ b, uint32_t nb,
                                                uint32_t* __restrict__ counters) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += stride) {
    if (i < na) a[i] = 0u;
    else b[i - na] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) counters[threadIdx.x] = 0u;
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

// Tiles of 16 x 256 rows, thread t owning rows [16t, 16t+16) of the tile
// (four dwordx4 loads per column). The qualifying rows of a tile are appended
// with one atomic per tile and a block-wide scan for the slots, so the list
// comes out in row order: for orders stored in key order the table inserts and
// bitmap updates that follow hit neighbouring addresses and coalesce.
constexpr int OTILE_PER_THREAD = 16;
__global__ __launch_bounds__(256) void q3_orders(const int32_t* __restrict__ okey,
                                                 const int32_t* __restrict__ ocust,
                                                 const int32_t* __restrict__ odate,
                                                 const int32_t* __restrict__ oprio, uint32_t n,
                                                 int32_t date, uint32_t maxck, uint32_t maxok,
                                                 const uint32_t* __restrict__ cbm,
                                                 uint32_t* __restrict__ obm,
                                                 uint32_t* __restrict__ counters,
                                                 int32_t* __restrict__ qkey,
                                                 int32_t* __restrict__ qdate,
                                                 int32_t* __restrict__ qprio) {
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t base_sh;
  const uint32_t tile_rows = OTILE_PER_THREAD * blockDim.x;
  const uint32_t ntiles = (n + tile_rows - 1) / tile_rows;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t r0 = t * tile_rows + threadIdx.x * OTILE_PER_THREAD;
    uint32_t mask = 0;
    if (r0 + OTILE_PER_THREAD <= n) {
#pragma unroll
      for (int v = 0; v < OTILE_PER_THREAD / 4; ++v) {
        const int4 d4 = ld4(odate, r0 / 4 + v), c4 = ld4(ocust, r0 / 4 + v), k4 = ld4(okey, r0 / 4 + v);
        const int32_t ds[4] = {d4.x, d4.y, d4.z, d4.w}, cs[4] = {c4.x, c4.y, c4.z, c4.w},
                      ks[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t ck = (uint32_t)cs[j];
          if (ds[j] < date && ck <= maxck && (uint32_t)ks[j] <= maxok && bit_test(cbm, ck))
            mask |= 1u << (v * 4 + j);
        }
      }
    } else {
      for (int j = 0; j < OTILE_PER_THREAD; ++j) {
        const uint32_t r = r0 + j;
        if (r < n) {
          const uint32_t ck = (uint32_t)ocust[r];
          if (odate[r] < date && ck <= maxck && (uint32_t)okey[r] <= maxok && bit_test(cbm, ck))
            mask |= 1u << j;
        }
      }
    }
    uint32_t total;
    uint32_t pos = block_excl_scan((uint32_t)__popc(mask), scratch, &total);
    if (total == 0) continue;
    if (threadIdx.x == 0) base_sh = atomicAdd(&counters[0], total);
    __syncthreads();
    pos += base_sh;
    while (mask) {
      const int j = __ffs((int)mask) - 1;
      mask &= mask - 1;
      const uint32_t r = r0 + j;
      const int32_t k = okey[r];
      qkey[pos] = k;
      qdate[pos] = odate[r];
      qprio[pos] = oprio[r];
      ++pos;
      atomicOr(&obm[(uint32_t)k >> 5], 1u << ((uint32_t)k & 31));
    }
    __syncthreads();  // base_sh is rewritten by the next tile
  }
}

__global__ __launch_bounds__(256) void q3_ht_clear(const uint32_t* __restrict__ counters,
                                                   int32_t* __restrict__ ht_key) {
  const uint32_t cap = ht_cap(counters[0]);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += stride) ht_key[i] = EMPTY;
}

__global__ __launch_bounds__(256) void q3_ht_insert(const uint32_t* __restrict__ counters,
                                                    const int32_t* __restrict__ qkey,
                                                    int32_t* __restrict__ ht_key,
                                                    int32_t* __restrict__ ht_q,
                                                    int64_t* __restrict__ ht_rev,
                                                    uint32_t* __restrict__ ht_hit) {
  const uint32_t count = counters[0];
  const uint32_t cap = ht_cap(count);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
    const int32_t key = qkey[j];
    uint32_t h = ht_hash(key, cap);
    for (uint32_t probe = 0; probe < cap; ++probe) {
      const int32_t prev = atomicCAS(&ht_key[h], EMPTY, key);
      if (prev == EMPTY) {
        ht_q[h] = (int32_t)j;
        ht_rev[h] = 0;
        ht_hit[h] = 0u;
        break;
      }
      if (prev == key) break;  // duplicate orderkey: the first insert keeps the slot
      h = (h + 1) & (cap - 1);
    }
  }
}

__device__ __forceinline__ void q3_flush(int32_t key, int64_t rev, uint32_t cap,
                                         const int32_t* __restrict__ ht_key,
                                         int64_t* __restrict__ ht_rev,
                                         uint32_t* __restrict__ ht_hit) {
  uint32_t h = ht_hash(key, cap);
  for (uint32_t probe = 0; probe < cap; ++probe) {
    const int32_t k = ht_key[h];
    if (k == key) {
      atomicAdd((unsigned long long*)&ht_rev[h], (unsigned long long)rev);
      ht_hit[h] = 1u;
      return;
    }
    if (k == EMPTY) return;
    h = (h + 1) & (cap - 1);
  }
}

struct Q3Fact {
  const int32_t *okey, *ship, *disc;
  const int64_t* price;
};

__device__ __forceinline__ void q3_group(const Q3Fact& f, uint32_t g, int32_t date,
                                         uint32_t maxok, const uint32_t* __restrict__ obm,
                                         uint32_t cap, const int32_t* __restrict__ ht_key,
                                         int64_t* __restrict__ ht_rev,
                                         uint32_t* __restrict__ ht_hit) {
  const int4 k4 = ld4(f.okey, g), s4 = ld4(f.ship, g);
  const int32_t ks[4] = {k4.x, k4.y, k4.z, k4.w};
  const int32_t ss[4] = {s4.x, s4.y, s4.z, s4.w};
  bool pass[4];
  bool any = false;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    pass[j] = ss[j] > date && (uint32_t)ks[j] <= maxok && bit_test(obm, (uint32_t)ks[j]);
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
      if (open) q3_flush(run_key, run_rev, cap, ht_key, ht_rev, ht_hit);
      run_key = ks[j];
      run_rev = r;
      open = true;
    }
  }
  if (open) q3_flush(run_key, run_rev, cap, ht_key, ht_rev, ht_hit);
}

template <int U>
__global__ __launch_bounds__(256) void q3_probe(Q3Fact f, uint32_t n, int32_t date, uint32_t maxok,
                                                const uint32_t* __restrict__ obm,
                                                const uint32_t* __restrict__ counters,
                                                const int32_t* __restrict__ ht_key,
                                                int64_t* __restrict__ ht_rev,
                                                uint32_t* __restrict__ ht_hit) {
  const uint32_t cap = ht_cap(counters[0]);
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  for (; g + (U - 1) * nthr < nvec; g += U * nthr) {
#pragma unroll
    for (int u = 0; u < U; ++u) q3_group(f, g + u * nthr, date, maxok, obm, cap, ht_key, ht_rev, ht_hit);
  }
  for (; g < nvec; g += nthr) q3_group(f, g, date, maxok, obm, cap, ht_key, ht_rev, ht_hit);
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const uint32_t r = nvec * 4 + threadIdx.x;
    const int32_t k = f.okey[r];
    if (f.ship[r] > date && (uint32_t)k <= maxok && bit_test(obm, (uint32_t)k))
      q3_flush(k, f.price[r] * (int64_t)(100 - f.disc[r]), cap, ht_key, ht_rev, ht_hit);
  }
}

// Top-k, step 1: histogram of the groups' revenue buckets (LDS per block,
// merged by global atomics) and the group count.
__global__ __launch_bounds__(256) void q3_topk_hist(const uint32_t* __restrict__ counters_in,
                                                    const int32_t* __restrict__ ht_key,
                                                    const int64_t* __restrict__ ht_rev,
                                                    const uint32_t* __restrict__ ht_hit,
                                                    uint32_t* __restrict__ hist,
                                                    uint32_t* __restrict__ counters) {
  __shared__ uint32_t lh[4096];
  __shared__ int64_t red[4];
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) lh[i] = 0u;
  __syncthreads();
  const uint32_t cap = ht_cap(counters_in[0]);
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t hits = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += stride) {
    if (ht_key[i] != EMPTY && ht_hit[i]) {
      atomicAdd(&lh[value_bucket(ht_rev[i])], 1u);
      ++hits;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4096; i += blockDim.x)
    if (lh[i]) atomicAdd(&hist[i], lh[i]);
  const int64_t bh = block_sum((int64_t)hits, red);
  if (threadIdx.x == 0 && bh) atomicAdd(&counters[1], (uint32_t)bh);
}

// Step 2 (one block): the highest bucket b such that buckets >= b hold at
// least k groups (0 when there are fewer than k groups).
__global__ __launch_bounds__(256) void q3_topk_threshold(const uint32_t* __restrict__ hist, int k,
                                                         uint32_t* __restrict__ counters) {
  __shared__ uint32_t part[256];
  // Thread t sums buckets [16t, 16t+16); then a serial suffix walk on thread 0.
  uint32_t s = 0;
  for (int i = 0; i < 16; ++i) s += hist[threadIdx.x * 16 + i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0, b = 0;
    for (int t = 255; t >= 0; --t) {
      if (acc + part[t] >= (uint32_t)k) {
        for (int i = 15; i >= 0; --i) {
          acc += hist[t * 16 + i];
          if (acc >= (uint32_t)k) {
            b = (uint32_t)(t * 16 + i);
            break;
          }
        }
        break;
      }
      acc += part[t];
    }
    counters[2] = b;
    counters[3] = 0u;  // candidate count
  }
}

// Step 3: groups in buckets >= threshold -> candidate list (tile-aggregated append).
__global__ __launch_bounds__(256) void q3_topk_collect(const int32_t* __restrict__ ht_key,
                                                       const int32_t* __restrict__ ht_q,
                                                       const int64_t* __restrict__ ht_rev,
                                                       const uint32_t* __restrict__ ht_hit,
                                                       const int32_t* __restrict__ qdate,
                                                       const int32_t* __restrict__ qprio,
                                                       uint32_t* __restrict__ counters,
                                                       uint32_t cand_cap,
                                                       TopEntry* __restrict__ cand) {
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t base_sh;
  const uint32_t cap = ht_cap(counters[0]);
  const uint32_t thr = counters[2];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t span = (cap + stride - 1) / stride * stride;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < span; i += stride) {
    bool q = false;
    if (i < cap && ht_key[i] != EMPTY && ht_hit[i]) q = value_bucket(ht_rev[i]) >= thr;
    uint32_t total;
    const uint32_t pre = block_excl_scan(q ? 1u : 0u, scratch, &total);
    if (total == 0) continue;
    if (threadIdx.x == 0) base_sh = atomicAdd(&counters[3], total);
    __syncthreads();
    const uint32_t pos = base_sh + pre;
    if (q && pos < cand_cap) {
      const int32_t j = ht_q[i];
      cand[pos] = TopEntry{ht_rev[i], qdate[j], ht_key[i], qprio[j], 0};
    }
    __syncthreads();
  }
}

// Step 4 (one block): rank the candidates; if more than fit the list (heavy
// ties in the threshold bucket), rank the qualifying table slots directly.
__global__ __launch_bounds__(256) void q3_topk_final(const TopEntry* __restrict__ cand,
                                                     uint32_t cand_cap,
                                                     const uint32_t* __restrict__ counters,
                                                     const int32_t* __restrict__ ht_key,
                                                     const int32_t* __restrict__ ht_q,
                                                     const int64_t* __restrict__ ht_rev,
                                                     const uint32_t* __restrict__ ht_hit,
                                                     const int32_t* __restrict__ qdate,
                                                     const int32_t* __restrict__ qprio, int k,
                                                     TopEntry* __restrict__ out) {
  __shared__ TopEntry buf[TOPK_BUF];
  __shared__ int nbuf;
  BlockTopK<TOPK_BUF> tk{buf, &nbuf, k};
  tk.init();
  const uint32_t nc = counters[3];
  if (nc <= cand_cap) {
    for (uint32_t base = 0; base < nc; base += blockDim.x) {
      const uint32_t i = base + threadIdx.x;
      TopEntry e;
      const bool valid = i < nc;
      if (valid) e = cand[i];
      tk.push(valid, e);
    }
  } else {
    const uint32_t cap = ht_cap(counters[0]);
    const uint32_t thr = counters[2];
    for (uint32_t base = 0; base < cap; base += blockDim.x) {
      const uint32_t i = base + threadIdx.x;
      TopEntry e;
      bool valid = false;
      if (i < cap && ht_key[i] != EMPTY && ht_hit[i] && value_bucket(ht_rev[i]) >= thr) {
        const int32_t j = ht_q[i];
        e = TopEntry{ht_rev[i], qdate[j], ht_key[i], qprio[j], 0};
        valid = true;
      }
      tk.push(valid, e);
    }
  }
  tk.finish(out);
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
  const uint32_t ncw = maxck / 32 + 1, now_ = maxok / 32 + 1;
  uint64_t capmax = 1024;
  while (capmax < 2ull * no) capmax <<= 1;
  const int k = p->top_k;

  // Scratch: bitmaps + counters + list (A_BITMAP), table (A_HT), top-k (A_AUX).
  Q3Scratch s{};
  {
    char* b;
    const size_t bm_bytes = ((size_t)ncw + now_ + 16 + 4096) * 4;
    const size_t list_bytes = (size_t)std::max<uint32_t>(no, 1) * 12;
    GDB_TRY(arena(ctx, A_BITMAP, bm_bytes + list_bytes + 256, (void**)&b));
    s.cbm = (uint32_t*)b;
    s.obm = s.cbm + ncw;
    s.counters = s.obm + now_;
    s.hist = s.counters + 16;
    s.qkey = (int32_t*)(b + ((bm_bytes + 255) & ~size_t(255)));
    s.qdate = s.qkey + std::max<uint32_t>(no, 1);
    s.qprio = s.qdate + std::max<uint32_t>(no, 1);
    char* h;
    GDB_TRY(arena(ctx, A_HT, (size_t)capmax * 20 + 256, (void**)&h));
    s.ht_rev = (int64_t*)h;
    s.ht_key = (int32_t*)(h + capmax * 8);
    s.ht_q = s.ht_key + capmax;
    s.ht_hit = (uint32_t*)(s.ht_q + capmax);
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
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  q3_clear<<<std::min<uint32_t>(2048, (ncw + now_ + 16 + 4096 + B - 1) / B + 1), B, 0, st>>>(
      s.cbm, ncw, s.obm, now_ + 16 + 4096, s.counters);
  if (nc)
    q3_customer<<<(nc + B - 1) / B, B, 0, st>>>((const int32_t*)cck->dev, (const uint8_t*)cseg->dev,
                                                 nc, p->segment, maxck, s.cbm);
  if (no)
    q3_orders<<<std::min<uint32_t>(ctx->info.cu_count * 8, (no + 16 * B - 1) / (16 * B)), B, 0, st>>>(
        (const int32_t*)cok->dev, (const int32_t*)coc->dev, (const int32_t*)cod->dev,
        (const int32_t*)cop->dev, no, p->date, maxck, maxok, s.cbm, s.obm, s.counters, s.qkey,
        s.qdate, s.qprio);
  const int fill_blocks = ctx->info.cu_count * 4;
  q3_ht_clear<<<fill_blocks, B, 0, st>>>(s.counters, s.ht_key);
  q3_ht_insert<<<fill_blocks, B, 0, st>>>(s.counters, s.qkey, s.ht_key, s.ht_q, s.ht_rev, s.ht_hit);
  GDB_HIP(hipGetLastError());
  const int president =
      resident_blocks(ctx, (const void*)q3_probe<2>, B, 0, ctx->tune[T_Q3].blocks_per_cu);
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const Q3Fact f{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                   (const int64_t*)d[3]};
    const int pblocks = std::max(1, std::min<int>(president, (int)((rows / 4 + B - 1) / B)));
    q3_probe<2><<<pblocks, B, 0, st>>>(f, rows, p->date, maxok, s.obm, s.counters, s.ht_key,
                                       s.ht_rev, s.ht_hit);
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
  q3_topk_hist<<<fill_blocks, B, 0, st>>>(s.counters, s.ht_key, s.ht_rev, s.ht_hit, s.hist,
                                          s.counters);
  q3_topk_threshold<<<1, B, 0, st>>>(s.hist, k, s.counters);
  q3_topk_collect<<<fill_blocks, B, 0, st>>>(s.ht_key, s.ht_q, s.ht_rev, s.ht_hit, s.qdate, s.qprio,
                                             s.counters, CAND_CAP, s.cand);
  q3_topk_final<<<1, B, 0, st>>>(s.cand, CAND_CAP, s.counters, s.ht_key, s.ht_q, s.ht_rev,
                                 s.ht_hit, s.qdate, s.qprio, k, s.out);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, s.out, sizeof(TopEntry) * k, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipMemcpyAsync(host_cnt, s.counters, 8, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  const uint64_t bytes = 5ull * nc + 16ull * no + 20ull * nl;
  GDB_TRY(timing_finish(ctx, timing, bytes, 20ull * nl, &rs));
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
