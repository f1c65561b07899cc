// Q9 multi-way join with a substring filter, grouped by (nation, year)
// (reference kernels/q9.metal:44-394, host src/q9.cpp:8-241, post-processing
// src/infra.h:733-776, SQL sql/q9.sql).
//
//   clear      part bitmap, counters, supplier->nation map (-1), orderkey->year map
//   part       tiles of 256 names staged in LDS; bit[partkey] if the name contains
//              the pattern (strmatch.h)                      (the reference's SWAR
//              'g' scan, kernels/q9.metal:58-83, generalised to any pattern)
//   supplier   nation[suppkey]
//   partsupp   rows whose part passed -> compact list (partkey, suppkey, cost)
//   ps_clear / ps_insert
//              open-addressing table keyed by (partkey, suppkey), capacity
//              nextpow2(2*count) from the device-side count
//   orders     year[orderkey] (uint8 offset from 1970)
//   probe      (dominant) fact rows whose part bit is set: supplier nation,
//              partsupp cost, order year; amount = extprice*(100-disc) -
//              cost*qty*100 added into an LDS-private [nation][year] table
//              (64 x 64 int64 + hit flags), merged by global atomics per block
// Columns other than l_partkey are loaded only for lanes with a surviving row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "kcommon.h"
#include "strmatch.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int TILE = 256;
constexpr int NW = 55;          // p_name width
constexpr int SLOT = 60;        // LDS bytes per name: 56 (zero pad) + 4 lookahead
constexpr int NNAT = 64, NYEAR = 64, YEAR0 = 1970;
constexpr int64_t EMPTY64 = -1;
constexpr uint8_t NO_YEAR = 0xFF;

__device__ __forceinline__ uint32_t ps_cap(uint32_t count) {
  uint32_t c = 1024;
  while (c < 2 * count) c <<= 1;
  return c;
}
__device__ __forceinline__ uint64_t ps_key(int32_t pk, int32_t sk) {
  return ((uint64_t)(uint32_t)pk << 32) | (uint32_t)sk;
}
// Neighbouring partkeys map to neighbouring slots (the list is inserted in
// partkey order, so the CAS inserts coalesce); the suppkey and the key's high
// bits are folded in to spread a part's suppliers and large keys.
__device__ __forceinline__ uint32_t ps_hash(uint64_t k, uint32_t cap) {
  const uint32_t pk = (uint32_t)(k >> 32), sk = (uint32_t)k;
  const int lg = __ffs((int)cap) - 1;
  const uint32_t x = pk * 4u + (sk & 3u);
  return (x ^ (x >> lg) ^ (sk * 0x9E3779B1u >> (32 - lg))) & (cap - 1);
}

// Year of a day number (day 0 = 1992-01-01) by the days-from-civil inverse.
__device__ __forceinline__ int year_of_day(int32_t day) {
  int64_t z = (int64_t)day + 8035 + 719468;
  int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  int64_t doe = z - era * 146097;
  int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  int64_t y = yoe + era * 400;
  int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  int64_t mp = (5 * doy + 2) / 153;
  return (int)(y + (mp >= 10));
}

struct Q9Scratch {
  uint32_t* pbm;       // part bitmap
  uint32_t* counters;  // [0] partsupp list count
  int32_t* snat;       // suppkey -> nation or -1
  uint8_t* oyear;      // orderkey -> year - 1970 or 0xFF
  int32_t* lpk;        // compact partsupp list
  int32_t* lsk;
  int64_t* lcost;
  int64_t* ht_key;     // (partkey, suppkey) or EMPTY64
  int64_t* ht_cost;
  int64_t* agg;        // [NNAT*NYEAR] profit
  uint32_t* agg_hit;   // [NNAT*NYEAR]
};

__global__ __launch_bounds__(256) void q9_clear(Q9Scratch s, uint32_t npw, uint32_t nsupp,
                                                uint32_t nord) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t a = npw, b = a + nsupp, c = b + (nord + 3) / 4, d = c + NNAT * NYEAR;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d; i += stride) {
    if (i < a) s.pbm[i] = 0u;
    else if (i < b) s.snat[i - a] = -1;
    else if (i < c) reinterpret_cast<uint32_t*>(s.oyear)[i - b] = 0xFFFFFFFFu;
    else {
      s.agg[i - c] = 0;
      s.agg_hit[i - c] = 0u;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) s.counters[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(TILE) void q9_part(const uint8_t* __restrict__ name,
                                                const int32_t* __restrict__ partkey, uint32_t n,
                                                Pattern pat, uint32_t maxpk,
                                                uint32_t* __restrict__ pbm) {
  __shared__ __align__(16) uint8_t tile[TILE * SLOT + 16];
  uint8_t* mine = tile + threadIdx.x * SLOT;
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t row0 = t * TILE;
    const uint32_t rows = min((uint32_t)TILE, n - row0);
    const uint8_t* src = name + (size_t)row0 * NW;
    __syncthreads();
    // Names are 55 bytes: copy byte-exact with dword loads where aligned.
    const uint32_t bytes = rows * NW;
    for (uint32_t b = threadIdx.x * 4; b < bytes; b += TILE * 4) {
      uint32_t v;
      if (b + 4 <= bytes && (((uintptr_t)(src + b)) & 3) == 0) {
        v = *reinterpret_cast<const uint32_t*>(src + b);
      } else {
        v = 0;
        for (int q = 0; q < 4 && b + q < bytes; ++q) v |= (uint32_t)src[b + q] << (8 * q);
      }
      for (int q = 0; q < 4 && b + q < bytes; ++q) {
        const uint32_t r = (b + q) / NW, off = (b + q) % NW;
        tile[r * SLOT + off] = (uint8_t)(v >> (8 * q));
      }
    }
    for (int q = NW; q < SLOT; ++q) mine[q] = 0;
    __syncthreads();
    if (threadIdx.x < rows) {
      const int len = field_len(mine, 56);
      if (find_from(mine, len, 0, pat) >= 0) {
        const uint32_t pk = (uint32_t)partkey[row0 + threadIdx.x];
        if (pk <= maxpk) atomicOr(&pbm[pk >> 5], 1u << (pk & 31));
      }
    }
  }
}

__global__ __launch_bounds__(256) void q9_supplier(const int32_t* __restrict__ sk,
                                                   const int32_t* __restrict__ nat, uint32_t n,
                                                   uint32_t maxsk, int32_t* __restrict__ snat) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (uint32_t)sk[i] <= maxsk) snat[sk[i]] = nat[i];
}

__global__ __launch_bounds__(256) void q9_orders(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ od, uint32_t n,
                                                 uint32_t maxok, uint8_t* __restrict__ oyear) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (uint32_t)ok[i] <= maxok) oyear[ok[i]] = (uint8_t)(year_of_day(od[i]) - YEAR0);
}

// Partsupp rows whose part passed -> compact list in row order, one atomic per
// tile of 16 x 256 rows (thread t owns rows [16t, 16t+16) of the tile).
__global__ __launch_bounds__(256) void q9_partsupp(const int32_t* __restrict__ pk,
                                                   const int32_t* __restrict__ sk,
                                                   const int64_t* __restrict__ cost, uint32_t n,
                                                   uint32_t maxpk, const uint32_t* __restrict__ pbm,
                                                   uint32_t* __restrict__ counters,
                                                   int32_t* __restrict__ lpk,
                                                   int32_t* __restrict__ lsk,
                                                   int64_t* __restrict__ lcost) {
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t base_sh;
  const uint32_t tile_rows = 16 * blockDim.x;
  const uint32_t ntiles = (n + tile_rows - 1) / tile_rows;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint32_t r0 = t * tile_rows + threadIdx.x * 16;
    uint32_t mask = 0;
    for (int j = 0; j < 16; ++j) {
      const uint32_t r = r0 + j;
      if (r < n) {
        const uint32_t p = (uint32_t)pk[r];
        if (p <= maxpk && ((pbm[p >> 5] >> (p & 31)) & 1u)) mask |= 1u << j;
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
      lpk[pos] = pk[r];
      lsk[pos] = sk[r];
      lcost[pos] = cost[r];
      ++pos;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void q9_ps_clear(const uint32_t* __restrict__ counters,
                                                   int64_t* __restrict__ ht_key) {
  const uint32_t cap = ps_cap(counters[0]);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += stride) ht_key[i] = EMPTY64;
}

__global__ __launch_bounds__(256) void q9_ps_insert(const uint32_t* __restrict__ counters,
                                                    const int32_t* __restrict__ lpk,
                                                    const int32_t* __restrict__ lsk,
                                                    const int64_t* __restrict__ lcost,
                                                    int64_t* __restrict__ ht_key,
                                                    int64_t* __restrict__ ht_cost) {
  const uint32_t count = counters[0];
  const uint32_t cap = ps_cap(count);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
    const uint64_t key = ps_key(lpk[j], lsk[j]);
    uint32_t h = ps_hash(key, cap);
    for (uint32_t probe = 0; probe < cap; ++probe) {
      const unsigned long long prev =
          atomicCAS((unsigned long long*)&ht_key[h], (unsigned long long)EMPTY64, key);
      if (prev == (unsigned long long)EMPTY64) {
        ht_cost[h] = lcost[j];
        break;
      }
      if (prev == key) break;  // duplicate (partkey, suppkey): first insert keeps the slot
      h = (h + 1) & (cap - 1);
    }
  }
}

struct Q9Fact {
  const int32_t *pk, *sk, *ok, *qty, *disc;
  const int64_t* price;
};

__device__ __forceinline__ void q9_row(int32_t pk, int32_t sk, int32_t ok, int32_t qty,
                                       int32_t disc, int64_t price, uint32_t cap,
                                       uint32_t maxsk, uint32_t maxok, const Q9Scratch& s,
                                       int64_t* lagg, uint32_t* lhit) {
  if ((uint32_t)sk > maxsk || (uint32_t)ok > maxok) return;
  const int32_t nat = s.snat[sk];
  if (nat < 0) return;
  const uint8_t y = s.oyear[ok];
  if (y == NO_YEAR) return;
  const uint64_t key = ps_key(pk, sk);
  uint32_t h = ps_hash(key, cap);
  int64_t cost = -1;
  for (uint32_t probe = 0; probe < cap; ++probe) {
    const int64_t k = s.ht_key[h];
    if (k == (int64_t)key) {
      cost = s.ht_cost[h];
      break;
    }
    if (k == EMPTY64) break;
    h = (h + 1) & (cap - 1);
  }
  if (cost < 0) return;
  const int64_t amount = price * (int64_t)(100 - disc) - cost * (int64_t)qty * 100;
  const int slot = nat * NYEAR + y;
  atomicAdd((unsigned long long*)&lagg[slot], (unsigned long long)amount);
  lhit[slot] = 1u;
}

// Orders LDS traffic within one wave: writes by some lanes become visible to
// reads by other lanes of the same wave after this.
constexpr int RING = 512;  // per-wave queue of surviving row ids

// Fact rows are tested 4 per lane on l_partkey alone; the survivors (about 5%)
// are queued per wave in LDS and processed 64 at a time, one row per lane, so
// the dependent chain (nation map, partsupp table, year map, LDS add) runs
// with full waves instead of once per row slot with most lanes idle.
__global__ __launch_bounds__(256) void q9_probe(Q9Fact f, uint32_t n, uint32_t maxpk,
                                                uint32_t maxsk, uint32_t maxok, Q9Scratch s) {
  __shared__ int64_t lagg[NNAT * NYEAR];
  __shared__ uint32_t lhit[NNAT * NYEAR];
  __shared__ uint32_t ring[4][RING];
  for (int i = threadIdx.x; i < NNAT * NYEAR; i += blockDim.x) {
    lagg[i] = 0;
    lhit[i] = 0u;
  }
  __syncthreads();
  const uint32_t cap = ps_cap(s.counters[0]);
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t* q = ring[w];
  uint32_t head = 0, tail = 0;  // wave-uniform
  auto drain = [&](bool all) {
    while (tail - head >= 64 || (all && tail != head)) {
      const uint32_t avail = min(64u, tail - head);
      if (lane < avail) {
        const uint32_t r = q[(head + lane) & (RING - 1)];
        q9_row(f.pk[r], f.sk[r], f.ok[r], f.qty[r], f.disc[r], f.price[r], cap, maxsk, maxok, s,
               lagg, lhit);
      }
      head += avail;
    }
  };
  const uint32_t nvec = n >> 2;
  const uint32_t wstride = gridDim.x * blockDim.x;           // vector groups per grid step
  const uint32_t wbase = blockIdx.x * blockDim.x + w * 64;   // first group of this wave
  for (uint32_t g0 = wbase; g0 < nvec; g0 += wstride) {
    const uint32_t g = g0 + lane;
    uint32_t m = 0;
    if (g < nvec) {
      const int4 p4 = ld4nt(f.pk, g);
      const int32_t ps[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if ((uint32_t)ps[j] <= maxpk && ((s.pbm[(uint32_t)ps[j] >> 5] >> (ps[j] & 31)) & 1u))
          m |= 1u << j;
    }
    const uint32_t c = (uint32_t)__popc(m);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t pos = tail + incl - c;
    while (m) {
      const int j = __ffs((int)m) - 1;
      m &= m - 1;
      q[(pos++) & (RING - 1)] = g * 4 + (uint32_t)j;
    }
    tail += total;
    wave_lds_sync();
    drain(false);
    wave_lds_sync();
  }
  if (blockIdx.x == 0 && w == 0) {  // wave-uniform: the last n % 4 rows
    const uint32_t r = nvec * 4 + lane;
    bool pass = false;
    if (lane < (n & 3)) {
      const uint32_t p = (uint32_t)f.pk[r];
      pass = p <= maxpk && ((s.pbm[p >> 5] >> (p & 31)) & 1u);
    }
    const uint64_t b = __ballot(pass);
    if (pass) q[(tail + (uint32_t)__popcll(b & ((1ull << lane) - 1))) & (RING - 1)] = r;
    tail += (uint32_t)__popcll(b);
  }
  wave_lds_sync();
  drain(true);
  __syncthreads();
  for (int i = threadIdx.x; i < NNAT * NYEAR; i += blockDim.x) {
    if (lhit[i]) {
      atomicAdd((unsigned long long*)&s.agg[i], (unsigned long long)lagg[i]);
      s.agg_hit[i] = 1u;
    }
  }
}

}  // namespace

extern "C" int gdb_q9(gdb_ctx* ctx, const gdb_table* part, const gdb_table* supplier,
                      const gdb_table* partsupp, const gdb_table* orders,
                      const gdb_table* lineitem, const gdb_q9_params* prm, gdb_q9_result* out,
                      gdb_timing* timing) {
  if (!ctx || !part || !supplier || !partsupp || !orders || !lineitem || !prm || !out)
    return fail(GDB_E_INVALID, "null argument");
  Pattern pat;
  {
    const size_t n = strnlen(prm->pattern, 32);
    if (n == 0 || n > 31) return fail(GDB_E_INVALID, "Q9 pattern must be 1..31 bytes");
    memset(&pat, 0, sizeof(pat));
    memcpy(pat.b, prm->pattern, n);
    pat.len = (int)n;
  }
  const Column *ppk, *pnm, *ssk, *snk, *pspk, *pssk, *psc, *ook, *ood, *lpk, *lsk, *lok, *lq,
      *ld, *lp;
  GDB_TRY(column(part, "p_partkey", GDB_I32, &ppk));
  GDB_TRY(column(part, "p_name", GDB_BYTES, &pnm));
  GDB_TRY(column(supplier, "s_suppkey", GDB_I32, &ssk));
  GDB_TRY(column(supplier, "s_nationkey", GDB_I32, &snk));
  GDB_TRY(column(partsupp, "ps_partkey", GDB_I32, &pspk));
  GDB_TRY(column(partsupp, "ps_suppkey", GDB_I32, &pssk));
  GDB_TRY(column(partsupp, "ps_supplycost", GDB_I64, &psc));
  GDB_TRY(column(orders, "o_orderkey", GDB_I32, &ook));
  GDB_TRY(column(orders, "o_orderdate", GDB_I32, &ood));
  GDB_TRY(column(lineitem, "l_partkey", GDB_I32, &lpk));
  GDB_TRY(column(lineitem, "l_suppkey", GDB_I32, &lsk));
  GDB_TRY(column(lineitem, "l_orderkey", GDB_I32, &lok));
  GDB_TRY(column(lineitem, "l_quantity", GDB_I32, &lq));
  GDB_TRY(column(lineitem, "l_discount", GDB_I32, &ld));
  GDB_TRY(column(lineitem, "l_extendedprice", GDB_I64, &lp));
  if (pnm->width != NW) return fail(GDB_E_UNSUPPORTED, "p_name must be a 55-byte field");
  GDB_TRY(require_resident(part, "Q9 part"));
  GDB_TRY(require_resident(supplier, "Q9 supplier"));
  GDB_TRY(require_resident(partsupp, "Q9 partsupp"));
  GDB_TRY(require_resident(orders, "Q9 orders"));
  if ((part->nrows && ppk->minv < 0) || (supplier->nrows && ssk->minv < 0) ||
      (orders->nrows && ook->minv < 0))
    return fail(GDB_E_UNSUPPORTED, "Q9 needs non-negative partkey, suppkey and orderkey values");
  if (supplier->nrows && (snk->minv < 0 || snk->maxv >= NNAT))
    return fail(GDB_E_UNSUPPORTED, "Q9 nation keys must be in 0..63");
  if (orders->nrows) {
    const int64_t d0 = ood->minv, d1 = ood->maxv;
    // days 1970-01-01 .. 2033-12-31 relative to 1992-01-01
    if (d0 < -8035 || d1 > 15340) return fail(GDB_E_UNSUPPORTED, "Q9 order years must be in 1970..2033");
  }
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t np = (uint32_t)part->nrows, ns = (uint32_t)supplier->nrows,
                 nps = (uint32_t)partsupp->nrows, no = (uint32_t)orders->nrows;
  const uint64_t nl = lineitem->nrows;
  const uint32_t maxpk = (uint32_t)std::max<int64_t>(0, ppk->maxv);
  const uint32_t maxsk = (uint32_t)std::max<int64_t>(0, ssk->maxv);
  const uint32_t maxok = (uint32_t)std::max<int64_t>(0, ook->maxv);
  const uint32_t npw = maxpk / 32 + 1, nsupp = maxsk + 1, nord = maxok + 1;
  uint64_t capmax = 1024;
  while (capmax < 2ull * nps) capmax <<= 1;
  Q9Scratch s{};
  {
    char* b;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) & ~size_t(255);
      return o;
    };
    const size_t o_pbm = take(npw * 4ull), o_cnt = take(16), o_snat = take(nsupp * 4ull),
                 o_oy = take((nord + 3) / 4 * 4ull), o_agg = take(NNAT * NYEAR * 8),
                 o_hit = take(NNAT * NYEAR * 4), o_lpk = take(std::max<uint32_t>(nps, 1) * 4ull),
                 o_lsk = take(std::max<uint32_t>(nps, 1) * 4ull),
                 o_lc = take(std::max<uint32_t>(nps, 1) * 8ull);
    GDB_TRY(arena(ctx, A_MAP, off, (void**)&b));
    s.pbm = (uint32_t*)(b + o_pbm);
    s.counters = (uint32_t*)(b + o_cnt);
    s.snat = (int32_t*)(b + o_snat);
    s.oyear = (uint8_t*)(b + o_oy);
    s.agg = (int64_t*)(b + o_agg);
    s.agg_hit = (uint32_t*)(b + o_hit);
    s.lpk = (int32_t*)(b + o_lpk);
    s.lsk = (int32_t*)(b + o_lsk);
    s.lcost = (int64_t*)(b + o_lc);
    char* h;
    GDB_TRY(arena(ctx, A_HT, capmax * 16 + 256, (void**)&h));
    s.ht_key = (int64_t*)h;
    s.ht_cost = s.ht_key + capmax;
  }
  char* host;
  GDB_TRY(pinned(ctx, NNAT * NYEAR * 12 + 64, (void**)&host));
  hipStream_t st = ctx->stream;
  const int B = 256;
  const uint32_t clear_items = npw + nsupp + (nord + 3) / 4 + NNAT * NYEAR;
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  q9_clear<<<std::min<uint32_t>(8192, (clear_items + B - 1) / B), B, 0, st>>>(s, npw, nsupp, nord);
  if (np) {
    int pb = std::max(1, std::min<int>(resident_blocks(ctx, (const void*)q9_part, TILE, 0, 0),
                                       (int)((np + TILE - 1) / TILE)));
    q9_part<<<pb, TILE, 0, st>>>((const uint8_t*)pnm->dev, (const int32_t*)ppk->dev, np, pat,
                                 maxpk, s.pbm);
  }
  if (ns)
    q9_supplier<<<(ns + B - 1) / B, B, 0, st>>>((const int32_t*)ssk->dev, (const int32_t*)snk->dev,
                                                 ns, maxsk, s.snat);
  if (no)
    q9_orders<<<(no + B - 1) / B, B, 0, st>>>((const int32_t*)ook->dev, (const int32_t*)ood->dev, no,
                                               maxok, s.oyear);
  if (nps)
    q9_partsupp<<<std::min<uint32_t>(ctx->info.cu_count * 8, (nps + 16 * B - 1) / (16 * B)), B, 0, st>>>(
        (const int32_t*)pspk->dev, (const int32_t*)pssk->dev, (const int64_t*)psc->dev, nps, maxpk,
        s.pbm, s.counters, s.lpk, s.lsk, s.lcost);
  const int fill_blocks = ctx->info.cu_count * 4;
  q9_ps_clear<<<fill_blocks, B, 0, st>>>(s.counters, s.ht_key);
  q9_ps_insert<<<fill_blocks, B, 0, st>>>(s.counters, s.lpk, s.lsk, s.lcost, s.ht_key, s.ht_cost);
  GDB_HIP(hipGetLastError());
  const int president =
      resident_blocks(ctx, (const void*)q9_probe, B, 0, ctx->tune[T_Q9].blocks_per_cu);
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const Q9Fact f{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                   (const int32_t*)d[3], (const int32_t*)d[4], (const int64_t*)d[5]};
    const int pblocks = std::max(1, std::min<int>(president, (int)((rows / 4 + B - 1) / B)));
    q9_probe<<<pblocks, B, 0, st>>>(f, rows, maxpk, maxsk, maxok, s);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  const Column* fc[6] = {lpk, lsk, lok, lq, ld, lp};
  RunStats rs;
  if (nl) {
    GDB_TRY(run_over(ctx, lineitem, fc, 6, consume, &rs));
  } else {
    GDB_HIP(hipEventRecord(ctx->ev[1], st));
    GDB_HIP(hipEventRecord(ctx->ev[2], st));
  }
  GDB_HIP(hipMemcpyAsync(host, s.agg, NNAT * NYEAR * 8, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipMemcpyAsync(host + NNAT * NYEAR * 8, s.agg_hit, NNAT * NYEAR * 4,
                         hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  const uint64_t bytes = 28ull * nl + 59ull * np + 16ull * nps + 8ull * ns + 8ull * no;
  GDB_TRY(timing_finish(ctx, timing, bytes, 28ull * nl, &rs));
  const int64_t* agg = (const int64_t*)host;
  const uint32_t* hit = (const uint32_t*)(host + NNAT * NYEAR * 8);
  std::vector<gdb_q9_row> rows;
  for (int i = 0; i < NNAT * NYEAR; ++i)
    if (hit[i]) rows.push_back(gdb_q9_row{i / NYEAR, YEAR0 + i % NYEAR, agg[i]});
  std::sort(rows.begin(), rows.end(), [](const gdb_q9_row& a, const gdb_q9_row& b) {
    const int c = strcmp(gdb_nation_name(a.nationkey), gdb_nation_name(b.nationkey));
    if (c != 0) return c < 0;
    if (a.nationkey != b.nationkey) return a.nationkey < b.nationkey;
    return a.year > b.year;
  });
  if (rows.size() > sizeof(out->rows) / sizeof(out->rows[0]))
    return fail(GDB_E_UNSUPPORTED, "more (nation, year) groups than the result holds");
  memset(out, 0, sizeof(*out));
  out->nrows = (int32_t)rows.size();
  std::copy(rows.begin(), rows.end(), out->rows);
  return GDB_OK;
}
