// Q9 multi-way join with a substring filter, grouped by (nation, year)
// (reference kernels/q9.metal:44-394, host src/q9.cpp:8-241, post-processing
// src/infra.h:733-776, SQL sql/q9.sql).
//
//   clear      part bitmap, counters, supplier->nation map (-1), orderkey->year map
//   part       tiles of 256 names copied contiguously into LDS with dwordx4
//              loads; bit[partkey] if the name contains the pattern
//              (strmatch.h; the reference's SWAR 'g' scan, kernels/q9.metal:58-83,
//              generalised to any pattern)
//   rank       each bitmap word gets the number of passing parts before it
//              (its high half), so a passing part's rank is that plus a
//              popcount; the bitmap is ORed into the probe's folded filter
//   supplier   nation[suppkey]
//   partsupp   direct table: PS_WAYS {suppkey, cost} slots per passing part,
//              indexed by the part's rank (about 3.5 MB at 2M parts, so the
//              probe's row reads stay in L2); else rows whose part passed ->
//              compact list (partkey, suppkey, cost) for the hash table:
//   ps_clear / ps_insert
//              open-addressing table of 16-byte {key, cost} slots keyed by
//              (partkey, suppkey), capacity nextpow2(2*count) from the
//              device-side count, so a probe touches one cache line
//   orders     year[orderkey] as an index into the query's year range, found
//              from the Jan-1 day numbers of those years (a kernel argument)
//   probe      (dominant) every fact column is streamed once with nontemporal
//              dwordx4 loads; rows whose part bit is set are queued per wave in
//              LDS with their values, and full waves of 64 queued rows then
//              run the dependent chain (nation map, partsupp table, year map)
//              and add amount = extprice*(100-disc) - cost*qty*100 into an
//              LDS table sized to the data's (nation, year) ranges, merged into
//              the global 64 x 64 table by atomics once per block
// Streaming all six columns reads each byte once; gathering the columns of
// the ~5% surviving rows touched over 80% of the 128-byte lines anyway and
// fetched 1.5x the column bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "internal.h"
#include "kcommon.h"
#include "strmatch.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int TILE = 256;
constexpr int NW = 55;          // p_name width
constexpr int NNAT = 64, NYEAR = 64, YEAR0 = 1970;
constexpr int64_t EMPTY64 = -1;
constexpr uint8_t NO_YEAR = 0xFF;
constexpr int QCAP = 128;       // per-wave survivor queue: < 64 left over + 64 pushed
constexpr int PS_WAYS = 4;      // direct partsupp table: supplier slots per part (TPC-H has 4)
// Device result block: profit and hit flags per (nation, year) slot, then
// four counters (q9_publish copies the query's slots to the host).
constexpr size_t RESULT_BYTES = (size_t)NNAT * NYEAR * 12 + 16;

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

// Jan-1 day numbers of the years an orders table spans: year index of `day`
// = number of year starts after the first that are <= day.
struct Years {
  int32_t jan1[NYEAR];
  int32_t ny;        // years spanned, 1..NYEAR
  int32_t first;     // first year - YEAR0
};

__device__ __forceinline__ uint32_t year_index(int32_t day, const Years& y) {
  uint32_t k = 0;
  for (int i = 1; i < y.ny; ++i) k += day >= y.jan1[i] ? 1u : 0u;
  return k;
}

// The year map, orderkey -> year index. NIB: 4 bits per key value, eight to
// a 32-bit word, 15 = no order (for at most 15 years and unique keys: half
// the footprint of the probe's gathers); else one byte per key value, 0xFF
// = no order. Both start as all ones.
template <bool NIB>
__device__ __forceinline__ uint8_t year_at(const uint8_t* oyear, uint32_t ok) {
  if (NIB) {
    const uint32_t v = (reinterpret_cast<const uint32_t*>(oyear)[ok >> 3] >> ((ok & 7u) * 4u)) & 15u;
    return v == 15u ? NO_YEAR : (uint8_t)v;
  }
  return oyear[ok];
}

// Part bitmap word: bits of 32 consecutive partkeys, and the number of set
// bits in all earlier words (the rank of the word's first passing part).
struct alignas(8) PartWord {
  uint32_t bits, rank;
};

__device__ __forceinline__ uint32_t part_rank(PartWord w, uint32_t pk) {
  return w.rank + (uint32_t)__popc(w.bits & ((1u << (pk & 31)) - 1u));
}

struct Q9Scratch {
  PartWord* pbw;       // part bitmap with ranks
  uint32_t* pfold;     // part bitmap folded to the probe's LDS filter size
  uint32_t* rsum;      // passing parts per rank block
  uint32_t* counters;  // [0] partsupp list count, [1] a part with more than PS_WAYS suppliers
  int32_t* snat;       // suppkey -> nation or -1
  uint8_t* oyear;      // orderkey -> year index (bytes, or 4-bit codes)
  int32_t* lpk;        // compact partsupp list
  int32_t* lsk;
  int64_t* lcost;
  longlong2* ht;       // {(partkey, suppkey) or EMPTY64, cost}
  int2* pst;           // direct table: PS_WAYS {suppkey or -1, cost} per passing part's rank
  uint32_t* pcnt;      // direct table: slots taken per rank
  int64_t* agg;        // [NNAT*NYEAR] profit, then agg_hit and counters (one readback)
  uint32_t* agg_hit;   // [NNAT*NYEAR]
};

// Clears the part bitmap, the first nsupp supplier-map entries, the first
// oywords words of the year map and the result table (the map entries past
// the cleared prefixes are all written by the supplier map and q9_orders).
__global__ __launch_bounds__(256) void q9_clear(Q9Scratch s, uint32_t npw, uint32_t nsupp,
                                                uint32_t oywords) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t a = npw, b = a + nsupp, c = b + oywords, d = c + NNAT * NYEAR;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d; i += stride) {
    if (i < a) s.pbw[i] = PartWord{0u, 0u};
    else if (i < b) s.snat[i - a] = -1;
    else if (i < c) reinterpret_cast<uint32_t*>(s.oyear)[i - b] = 0xFFFFFFFFu;
    else {
      s.agg[i - c] = 0;
      s.agg_hit[i - c] = 0u;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) s.counters[threadIdx.x] = 0u;
}

// Arguments of the dimension-side kernels (part filter, year map, nation map).
struct Q9DimArgs {
  const uint8_t* name;
  const int32_t* partkey;
  uint32_t np;
  PatternW pat;
  uint32_t maxpk;
  PartWord* pbw;
  const int32_t* ok;
  const int32_t* od;
  uint32_t no;
  uint32_t maxok;
  Years years;
  uint8_t* oyear;
  const int32_t* sk;
  const int32_t* nat;
  uint32_t ns;
  uint32_t maxsk;
  int32_t* snat;
};

// A tile of 256 names is 14080 contiguous bytes (16-byte aligned, as every
// tile starts at a multiple of 256 rows): it is copied into LDS with dwordx4
// loads (the next tile's chunks are loaded into registers while the current
// one is searched), then lane t assembles name t (55 bytes at LDS offset 55t)
// into 14 aligned dwords plus two zero pad dwords and searches it in
// registers. Names with the pattern set their part's bit.
__device__ __forceinline__ void q9_part_tiles(const Q9DimArgs& a, uint32_t b0, uint32_t nb,
                                              int4* tile) {
  constexpr int TBYTES = TILE * NW;          // 14080
  constexpr int TCHUNKS = TBYTES / 16;       // 880
  constexpr int PER = (TCHUNKS + TILE - 1) / TILE;  // 4 chunks per thread
  const uint32_t* tw = reinterpret_cast<const uint32_t*>(tile);
  const uint64_t col_chunks = ((uint64_t)a.np * NW + 15) / 16;
  const uint32_t ntiles = (a.np + TILE - 1) / TILE;
  int4 nxt[PER];
  auto fetch = [&](uint32_t t) {
    const uint64_t c0 = (uint64_t)t * TCHUNKS;
    const int4* src = reinterpret_cast<const int4*>(a.name) + c0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t c = threadIdx.x + k * TILE;
      nxt[k] = c < TCHUNKS && c0 + c < col_chunks ? src[c] : int4{0, 0, 0, 0};
    }
  };
  if (b0 < ntiles) fetch(b0);
  for (uint32_t t = b0; t < ntiles; t += nb) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t c = threadIdx.x + k * TILE;
      if (c < TCHUNKS) tile[c] = nxt[k];
    }
    if (threadIdx.x == 0) tile[TCHUNKS] = int4{0, 0, 0, 0};
    __syncthreads();
    if (t + nb < ntiles) fetch(t + nb);
    const uint32_t r = t * TILE + threadIdx.x;
    if (r >= a.np) continue;
    const uint32_t off = threadIdx.x * NW, w0 = off >> 2, sh = off & 3;
    uint32_t d[16];
#pragma unroll
    for (int i = 0; i < 14; ++i) {
      const uint32_t lo = tw[w0 + i], hi = tw[w0 + i + 1];
      d[i] = __builtin_amdgcn_alignbyte(hi, lo, sh);
    }
    d[13] &= 0x00FFFFFFu;  // byte 55 belongs to the next name
    d[14] = 0u;
    d[15] = 0u;
    const int len = reg_len<14>(d);
    if (reg_find<14>(d, len, 0, a.pat) >= 0) {
      const uint32_t pk = (uint32_t)a.partkey[r];
      if (pk <= a.maxpk) atomicOr(&a.pbw[pk >> 5].bits, 1u << (pk & 31));
    }
  }
}

__global__ __launch_bounds__(TILE) void q9_part(Q9DimArgs a) {
  __shared__ int4 tile[TILE * NW / 16 + 1];
  q9_part_tiles(a, blockIdx.x, gridDim.x, tile);
}

// Year index by orderkey: one vector group of 4 orders per thread (dwordx4
// loads of key and date). (Running the part filter, this and the supplier map
// as one grid, workgroup ranges per task, took longer than the three
// launches: each is latency-bound and got fewer workgroups.)
// NIB: a key's 4-bit code is set by ANDing its word with the code in place
// of the key's 1111 (one non-returning atomicAnd per word the thread's four
// orders touch); exact because the keys are unique.
template <bool NIB>
__global__ __launch_bounds__(256) void q9_orders(Q9DimArgs a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (a.no + 3) / 4) return;
  const int4 k = ld4nt(a.ok, g), d = ld4nt(a.od, g);
  const int32_t ks[4] = {k.x, k.y, k.z, k.w}, ds[4] = {d.x, d.y, d.z, d.w};
  if (!NIB) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (4 * g + j < a.no && (uint32_t)ks[j] <= a.maxok)
        a.oyear[ks[j]] = (uint8_t)year_index(ds[j], a.years);
    return;
  }
  uint32_t* w32 = reinterpret_cast<uint32_t*>(a.oyear);
  uint32_t cur = 0xFFFFFFFFu, mask = 0xFFFFFFFFu;  // keys are < 2^31: word < 2^28
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (4 * g + j < a.no && (uint32_t)ks[j] <= a.maxok) {
      const uint32_t key = (uint32_t)ks[j], wi = key >> 3;
      const uint32_t m = ~((15u ^ year_index(ds[j], a.years)) << ((key & 7u) * 4u));
      if (wi != cur) {
        if (cur != 0xFFFFFFFFu) atomicAnd(&w32[cur], mask);
        cur = wi;
        mask = m;
      } else {
        mask &= m;
      }
    }
  }
  if (cur != 0xFFFFFFFFu) atomicAnd(&w32[cur], mask);
}

// 4-bit year codes with every map word written by one workgroup and no
// atomics or clear (keys strictly ascending). Workgroup g takes orders
// [g * OY_SLICE, +OY_SLICE) and owns the map words from its first key's word
// (word 0 for g = 0) up to the next slice's first key's word (the map's end
// for the last slice); the at most seven orders before its slice whose keys
// share its first word are its too, and its own last orders whose word is
// the next slice's first are left to that slice. The owned words are built in
// LDS chunks (all ones, then an LDS AND per order) and stored whole, by
// non-returning atomic exchanges: the probe's gathers from the map ran at
// 335 us after such writes against 368 us after plain stores of the same
// words. (Writing the bitmap words, partsupp rows and supplier nations the
// same way gained 2 us in the probe and cost 4 us in their kernels.)
constexpr uint32_t OY_SLICE = 4096;  // orders per workgroup
constexpr uint32_t OY_CHUNK = 2048;  // map words per LDS pass
__global__ __launch_bounds__(256) void q9_orders_owned(Q9DimArgs a, uint32_t nwords) {
  __shared__ uint32_t kw[OY_SLICE + 8];  // the slice's keys, then the earlier orders'
  __shared__ uint8_t code[OY_SLICE + 8];
  __shared__ uint32_t chunk[OY_CHUNK];
  const uint32_t o0 = blockIdx.x * OY_SLICE, o1 = min(a.no, o0 + OY_SLICE);
  const uint32_t w0 = blockIdx.x == 0 ? 0u : (uint32_t)a.ok[o0] >> 3;
  const uint32_t w1 = o1 < a.no ? (uint32_t)a.ok[o1] >> 3 : nwords;
  const uint32_t nq = (o1 - o0 + 3) / 4;  // 4-order groups (the columns are padded)
  // All of the thread's key and date loads are issued before any is used
  // (clamped addresses; groups past the slice are masked afterwards).
  constexpr uint32_t QPT = OY_SLICE / 4 / 256;
  int4 kk[QPT], dd[QPT];
#pragma unroll
  for (uint32_t u = 0; u < QPT; ++u) {
    const uint32_t q = min(threadIdx.x + u * 256, nq - 1);
    kk[u] = ld4nt(a.ok, o0 / 4 + q);
    dd[u] = ld4nt(a.od, o0 / 4 + q);
  }
#pragma unroll
  for (uint32_t u = 0; u < QPT; ++u) {
    const uint32_t q = threadIdx.x + u * 256;
    if (q >= nq) continue;
    const int32_t ks[4] = {kk[u].x, kk[u].y, kk[u].z, kk[u].w};
    const int32_t ds[4] = {dd[u].x, dd[u].y, dd[u].z, dd[u].w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = 4 * q + j;
      const bool in = o0 + i < o1;
      kw[i] = in ? (uint32_t)ks[j] : 0xFFFFFFFFu;
      code[i] = in ? (uint8_t)year_index(ds[j], a.years) : (uint8_t)15;
    }
  }
  if (threadIdx.x < 8) {
    const uint32_t t = threadIdx.x;
    uint32_t key = 0xFFFFFFFFu;
    uint8_t c = 15;
    if (t < 7 && o0 >= t + 1) {
      const uint32_t k = (uint32_t)a.ok[o0 - 1 - t];
      if ((k >> 3) == w0) {
        key = k;
        c = (uint8_t)year_index(a.od[o0 - 1 - t], a.years);
      }
    }
    kw[OY_SLICE + t] = key;
    code[OY_SLICE + t] = c;
  }
  const uint32_t nk = 4 * nq;
  uint32_t* map = reinterpret_cast<uint32_t*>(a.oyear);
  for (uint32_t c0 = w0; c0 < w1; c0 += OY_CHUNK) {
    const uint32_t cw = min(OY_CHUNK, w1 - c0);
    for (uint32_t j = threadIdx.x; j < OY_CHUNK; j += 256) chunk[j] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nk + 8; i += 256) {
      const uint32_t e = i < nk ? i : OY_SLICE + (i - nk);
      const uint32_t key = kw[e];
      const uint32_t w = key >> 3;  // 0x1FFFFFFF for an empty entry: never in range
      if (key != 0xFFFFFFFFu && w >= c0 && w < c0 + cw)
        atomicAnd(&chunk[w - c0], ~((15u ^ code[e]) << ((key & 7u) * 4u)));
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < cw; j += 256) atomicExch(&map[c0 + j], chunk[j]);
    __syncthreads();
  }
}

// Supplier map: suppkey -> nation, 256 suppliers per workgroup `blk`.
__device__ __forceinline__ void q9_supplier_block(const Q9DimArgs& a, uint32_t blk) {
  const uint32_t i = blk * 256u + threadIdx.x;
  if (i < a.ns && (uint32_t)a.sk[i] <= a.maxsk) a.snat[a.sk[i]] = a.nat[i];
}

// The part bitmap folded onto fwords words: bit k of the fold is the OR of
// the bits of every key congruent to k modulo 32 * fwords (256 fold words
// per workgroup `blk`).
__device__ __forceinline__ void q9_fold_block(const PartWord* __restrict__ pbw, uint32_t npw,
                                              uint32_t fwords, uint32_t* __restrict__ fold,
                                              uint32_t blk) {
  const uint32_t w = blk * 256u + threadIdx.x;
  if (w >= fwords) return;
  uint32_t v = 0;
  for (uint32_t i = w; i < npw; i += fwords) v |= pbw[i].bits;
  fold[w] = v;
}

// Rank, step 1: passing parts in rank block `blk` of RANK_WORDS bitmap words.
constexpr uint32_t RANK_WORDS = 512;  // two words per thread of a 256-thread block
__device__ __forceinline__ void q9_rank_sums_block(const PartWord* __restrict__ pbw, uint32_t npw,
                                                   uint32_t* __restrict__ rsum, uint32_t blk) {
  __shared__ int64_t red[4];
  const uint32_t w = blk * RANK_WORDS + 2 * threadIdx.x;
  uint32_t a = 0, b = 0;
  if (w + 1 < npw) {
    const uint4 v = reinterpret_cast<const uint4*>(pbw)[w / 2];
    a = v.x;
    b = v.z;
  } else if (w < npw) {
    a = pbw[w].bits;
  }
  const int64_t sum = block_sum(__popc(a) + __popc(b), red);
  if (threadIdx.x == 0) rsum[blk] = (uint32_t)sum;
}

// The rank sums, the fold and the supplier map in one launch (each alone is
// mostly launch and drain time, about 4.6 us): workgroups [0, rb) count rank
// blocks, [rb, rb + fb) fold the bitmap, the rest fill the supplier map.
// The role is uniform per workgroup, so the rank count's barriers hold.
__global__ __launch_bounds__(256) void q9_after_part(const PartWord* __restrict__ pbw,
                                                     uint32_t npw, uint32_t fwords,
                                                     uint32_t* __restrict__ fold,
                                                     uint32_t* __restrict__ rsum, uint32_t rb,
                                                     uint32_t fb, Q9DimArgs a) {
  const uint32_t b = blockIdx.x;
  if (b < rb) q9_rank_sums_block(pbw, npw, rsum, b);
  else if (b < rb + fb) q9_fold_block(pbw, npw, fwords, fold, b - rb);
  else q9_supplier_block(a, b - rb - fb);
}

// Rank, step 2: the block's base is the sum of the earlier blocks' counts;
// each word gets the number of passing parts before it. With the direct
// partsupp table, the block empties the supplier slots and slot counters of
// the ranks it covers.
__global__ __launch_bounds__(256) void q9_rank_write(PartWord* __restrict__ pbw, uint32_t npw,
                                                     const uint32_t* __restrict__ rsum,
                                                     int2* __restrict__ pst,
                                                     uint32_t* __restrict__ pcnt) {
  __shared__ uint32_t scratch[17];
  __shared__ int64_t red[4];
  __shared__ uint32_t base_sh;
  int64_t pre = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += blockDim.x) pre += rsum[b];
  pre = block_sum(pre, red);
  if (threadIdx.x == 0) base_sh = (uint32_t)pre;
  __syncthreads();
  const uint32_t base = base_sh;
  const uint32_t w = blockIdx.x * RANK_WORDS + 2 * threadIdx.x;
  uint32_t a = 0, b = 0;
  const bool pair = w + 1 < npw;
  if (pair) {
    const uint4 v = reinterpret_cast<const uint4*>(pbw)[w / 2];
    a = v.x;
    b = v.z;
  } else if (w < npw) {
    a = pbw[w].bits;
  }
  uint32_t total;
  const uint32_t r = base + block_excl_scan(__popc(a) + __popc(b), scratch, &total);
  if (pair) reinterpret_cast<uint4*>(pbw)[w / 2] = uint4{a, r, b, r + __popc(a)};
  else if (w < npw) pbw[w] = PartWord{a, r};
  if (!pst) return;
  int4* rows = reinterpret_cast<int4*>(pst);
  for (uint32_t j = threadIdx.x; j < 2 * total; j += blockDim.x)
    rows[2 * (size_t)base + j] = int4{-1, 0, -1, 0};
  for (uint32_t j = threadIdx.x; j < total; j += blockDim.x) pcnt[base + j] = 0u;
}

// Partsupp rows whose part passed -> compact list in row order, one atomic per
// tile of 16 x 256 rows: thread t tests rows 4t..4t+3 of each of the tile's
// four 1024-row quarters with dwordx4 loads of ps_partkey.
__global__ __launch_bounds__(256) void q9_partsupp(const int32_t* __restrict__ pk,
                                                   const int32_t* __restrict__ sk,
                                                   const int64_t* __restrict__ cost, uint32_t n,
                                                   uint32_t maxpk, const PartWord* __restrict__ pbw,
                                                   uint32_t* __restrict__ counters,
                                                   int32_t* __restrict__ lpk,
                                                   int32_t* __restrict__ lsk,
                                                   int64_t* __restrict__ lcost) {
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t base_sh;
  const uint32_t tile_rows = 16 * blockDim.x;
  const uint32_t ntiles = (n + tile_rows - 1) / tile_rows;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    uint32_t rows[4];
    int32_t ks[16];
#pragma unroll
    for (int v = 0; v < 4; ++v) rows[v] = t * tile_rows + v * 4 * blockDim.x + 4 * threadIdx.x;
    if (t * tile_rows + tile_rows <= n) {
      int4 k4[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) k4[v] = ld4(pk, rows[v] / 4);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ks[4 * v] = k4[v].x; ks[4 * v + 1] = k4[v].y; ks[4 * v + 2] = k4[v].z; ks[4 * v + 3] = k4[v].w;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int j = 0; j < 4; ++j) ks[4 * v + j] = rows[v] + j < n ? pk[rows[v] + j] : -1;
    }
    uint32_t words[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) words[j] = pbw[min((uint32_t)ks[j], maxpk) >> 5].bits;
    uint32_t mask = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t p = (uint32_t)ks[j];
      mask |= (p <= maxpk && ((words[j] >> (p & 31)) & 1u) ? 1u : 0u) << j;
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
      const uint32_t r = rows[j >> 2] + (j & 3);
      lpk[pos] = ks[j];
      lsk[pos] = sk[r];
      lcost[pos] = cost[r];
      ++pos;
    }
    __syncthreads();
  }
}

// Direct table: each partsupp row whose part passed takes the next slot of
// its part's rank (a per-rank counter: the adds of one thread's rows are
// independent, so they are in flight together); a fifth supplier of a part
// sets counters[1] and the query runs again through the hash table.
// (Claiming slots by compare-and-swap of the suppkey took up to ten dependent
// round trips per thread, a part's rows being adjacent.)
__global__ __launch_bounds__(256) void q9_partsupp_direct(const int32_t* __restrict__ pk,
                                                          const int32_t* __restrict__ sk,
                                                          const int64_t* __restrict__ cost,
                                                          uint32_t n, uint32_t maxpk,
                                                          const PartWord* __restrict__ pbw,
                                                          int2* __restrict__ pst,
                                                          uint32_t* __restrict__ pcnt,
                                                          uint32_t* __restrict__ counters) {
  const uint32_t nvec = (n + 3) / 4;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nvec; g += gridDim.x * blockDim.x) {
    int32_t ks[4];
    if (4 * g + 4 <= n) {
      const int4 k4 = ld4nt(pk, g);
      ks[0] = k4.x; ks[1] = k4.y; ks[2] = k4.z; ks[3] = k4.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) ks[j] = 4 * g + j < n ? pk[4 * g + j] : -1;
    }
    PartWord words[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) words[j] = pbw[min((uint32_t)ks[j], maxpk) >> 5];
    bool green[4];
    uint32_t rank[4], slot[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t p = (uint32_t)ks[j];
      green[j] = p <= maxpk && ((words[j].bits >> (p & 31)) & 1u);
      rank[j] = part_rank(words[j], p);
      slot[j] = green[j] ? atomicAdd(&pcnt[rank[j]], 1u) : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!green[j]) continue;
      const uint32_t r = 4 * g + j;
      if (slot[j] < PS_WAYS)
        pst[(size_t)rank[j] * PS_WAYS + slot[j]] = int2{sk[r], (int32_t)cost[r]};
      else
        atomicOr(&counters[1], 1u);
    }
  }
}

__global__ __launch_bounds__(256) void q9_ps_clear(const uint32_t* __restrict__ counters,
                                                   longlong2* __restrict__ ht) {
  const uint32_t cap = ps_cap(counters[0]);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += stride) {
    longlong2 e;
    e.x = EMPTY64;
    e.y = 0;
    ht[i] = e;
  }
}

__global__ __launch_bounds__(256) void q9_ps_insert(const uint32_t* __restrict__ counters,
                                                    const int32_t* __restrict__ lpk,
                                                    const int32_t* __restrict__ lsk,
                                                    const int64_t* __restrict__ lcost,
                                                    longlong2* __restrict__ ht) {
  const uint32_t count = counters[0];
  const uint32_t cap = ps_cap(count);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
    const uint64_t key = ps_key(lpk[j], lsk[j]);
    uint32_t h = ps_hash(key, cap);
    for (uint32_t probe = 0; probe < cap; ++probe) {
      const unsigned long long prev =
          atomicCAS((unsigned long long*)&ht[h].x, (unsigned long long)EMPTY64, key);
      if (prev == (unsigned long long)EMPTY64) {
        ht[h].y = lcost[j];
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

struct Q9Dims {
  uint32_t maxpk, maxsk, maxok;
  int32_t ny;        // years in the LDS table
  int32_t first;     // global year slot of LDS year 0
  int32_t nslots;    // nations x ny
  uint32_t fwords;   // words of the folded part filter in LDS (a power of two)
  int32_t direct;    // 1: partsupp costs in the direct table, else the hash table
};

// The result into mapped host memory, compacted to the nnat x ny slots the
// query can fill (2.1 KB for 25 nations over 7 years, against 48 KB for the
// whole 64 x 64 table, which a workgroup's stores to host memory take 13 us
// to write): profit[nslots], hit[nslots], then the four counters.
__global__ __launch_bounds__(256) void q9_publish(Q9Scratch s, int32_t ny, int32_t first,
                                                  int32_t nslots, char* out) {
  int64_t* profit = reinterpret_cast<int64_t*>(out);
  uint32_t* hit = reinterpret_cast<uint32_t*>(out + (size_t)nslots * 8);
  for (int i = threadIdx.x; i < nslots; i += blockDim.x) {
    const int g = (i / ny) * NYEAR + first + i % ny;
    profit[i] = s.agg[g];
    hit[i] = s.agg_hit[g];
  }
  if (threadIdx.x < 4) hit[nslots + threadIdx.x] = s.counters[threadIdx.x];
}

// One queued row: nation, year and the first partsupp slot are loaded
// together (none depends on another), then the rare further probe steps. pk
// is the part's rank with the direct table, else its partkey.
template <bool NIB>
__device__ __forceinline__ void q9_entry(int32_t pk, int32_t sk, int32_t ok, int32_t qty,
                                         int64_t amt, uint32_t cap, const Q9Dims& dm,
                                         const Q9Scratch& s, int64_t* ltab, uint32_t* lhit) {
  const bool inr = (uint32_t)sk <= dm.maxsk && (uint32_t)ok <= dm.maxok;
  const int32_t nat = s.snat[inr ? (uint32_t)sk : 0u];
  const uint8_t y = year_at<NIB>(s.oyear, inr ? (uint32_t)ok : 0u);
  int64_t cost = -1;
  if (dm.direct) {  // the part passed the bitmap, so its rank's slots were set up
    const int4* row = reinterpret_cast<const int4*>(s.pst + (size_t)(uint32_t)pk * PS_WAYS);
    const int4 a = row[0], b = row[1];
    cost = a.x == sk ? a.y : a.z == sk ? a.w : b.x == sk ? b.y : b.z == sk ? b.w : -1;
    if (sk < 0) cost = -1;
  } else {
    const uint64_t key = ps_key(pk, sk);
    uint32_t h = ps_hash(key, cap);
    longlong2 e = s.ht[h];
    for (uint32_t probe = 0; probe < cap; ++probe) {
      if (e.x == (int64_t)key) {
        cost = e.y;
        break;
      }
      if (e.x == EMPTY64) break;
      h = (h + 1) & (cap - 1);
      e = s.ht[h];
    }
  }
  if (!inr || nat < 0 || y == NO_YEAR || cost < 0) return;
  const int slot = nat * dm.ny + y;
  atomicAdd((unsigned long long*)&ltab[slot], (unsigned long long)(amt - cost * (int64_t)qty * 100));
  lhit[slot] = 1u;
}

// A wave's queue of rows whose part passed: {key, suppkey, orderkey,
// quantity} and the amount, one 16-byte and one 8-byte LDS access per row.
struct Q9Queue {
  int4 e[QCAP];
  int64_t amt[QCAP];
};

// The probe is software-pipelined over trips of U groups of 4 rows per lane.
// A wave's vector loads complete in issue order, so a wait for any load also
// waits for every load issued before it. Each trip therefore issues its
// dependent loads before the next trip's column loads, and uses their results
// one trip later, after the wait for its own columns:
//   A  push the previous trip's rows whose part bit is set (their bitmap
//      words arrived with this trip's columns) onto the wave's LDS queue
//   B  finish the previous trip's queue batch: cost from the partsupp row,
//      amount into the LDS (nation, year) table
//   C  wait for this trip's columns; test the keys against the folded filter
//      in LDS and load the global bitmap word of every fold hit
//   D  with 64 rows queued, take them off the queue and load their nation,
//      year and partsupp row (direct table)
//   E  load the next trip's six columns
// The wait in C is the only wait on HBM: its columns were issued one trip
// earlier, and the dependent loads in flight meanwhile were issued before
// them. (With the dependent loads issued after the prefetch, each of their
// waits also waited for the next trip's columns: 400 us at 60M fact rows.)
// The part test reads the folded filter in LDS: random 4-byte gathers from
// the 250 KB global bitmap cost a vector-cache line per lane and took as long
// as streaming all six columns. Only keys whose folded bit is set read the
// global bitmap word, which also gives a passing part's rank: the queue
// carries the rank to the direct table's row. One workgroup of 512 threads
// per CU shares one fold (2^20 bits next to eight wave queues); two trips'
// columns take 167 VGPRs per lane, which eight waves per CU allow (sixteen
// waves with one group per trip spilled). With the hash table (no direct
// table) queued rows are processed in place, 64 at a time, by q9_entry.
constexpr int PROBE_THREADS = 512;
constexpr int PROBE_U = 2;  // 4-row groups per lane and trip
template <int U, bool DIRECT, bool NIB>
__global__ __launch_bounds__(PROBE_THREADS) void q9_probe(Q9Fact f, uint32_t n, Q9Dims dm,
                                                          Q9Scratch s) {
  extern __shared__ uint32_t dyn[];  // folded filter, then nslots sums and nslots hit flags
  __shared__ Q9Queue queues[PROBE_THREADS / 64];
  uint32_t* lf = dyn;
  int64_t* ltab = reinterpret_cast<int64_t*>(dyn + dm.fwords);
  uint32_t* lhit = reinterpret_cast<uint32_t*>(ltab + dm.nslots);
  for (uint32_t i = threadIdx.x; i < dm.fwords; i += blockDim.x) lf[i] = s.pfold[i];
  for (int i = threadIdx.x; i < dm.nslots; i += blockDim.x) {
    ltab[i] = 0;
    lhit[i] = 0u;
  }
  __syncthreads();
  const uint32_t fmask = dm.fwords * 32u - 1u;
  const uint32_t cap = ps_cap(s.counters[0]);
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1;
  Q9Queue& q = queues[w];
  uint32_t head = 0, tail = 0;  // wave-uniform
  // Processes queued rows now (64 per pass, waiting for each) until at most
  // `keep` are left.
  auto drain_now = [&](uint32_t keep) {
    wave_lds_sync();
    while (tail - head > keep) {
      const uint32_t avail = min(64u, tail - head);
      if (lane < avail) {
        const uint32_t i = (head + lane) & (QCAP - 1);
        const int4 e = q.e[i];
        q9_entry<NIB>(e.x, e.y, e.z, e.w, q.amt[i], cap, dm, s, ltab, lhit);
      }
      head += avail;
    }
    wave_lds_sync();
  };
  // Queues the rows with pass set; a full queue is processed first.
  auto push = [&](bool pass, int32_t key, int32_t sk, int32_t ok, int32_t qty, int64_t amt) {
    const uint64_t b = __ballot(pass);
    const uint32_t cnt = (uint32_t)__popcll(b);
    if (cnt == 0) return;
    if (tail - head + cnt > QCAP) drain_now(0);
    if (pass) {
      const uint32_t i = (tail + (uint32_t)__popcll(b & below)) & (QCAP - 1);
      q.e[i] = int4{key, sk, ok, qty};
      q.amt[i] = amt;
    }
    tail += cnt;
  };
  // Batch of 64 queued rows (direct table): taken off the queue and its loads
  // issued in D, finished in B of the next trip. The loads are issued on
  // every trip, from clamped addresses when no batch is due, so that no value
  // merges with an older one while the loads are in flight (a merge is a
  // register copy, and the copy waits for the load).
  bool gpend = false;  // wave-uniform
  bool g_in = false;
  int32_t g_sk = 0, g_qty = 0, g_nat = -1;
  int64_t g_amt = 0;
  uint8_t g_y = NO_YEAR;
  int4 g_ra{-1, 0, -1, 0}, g_rb{-1, 0, -1, 0};
  auto batch_issue = [&]() {
    const bool due = tail - head >= 64;
    wave_lds_sync();
    const uint32_t i = (head + lane) & (QCAP - 1);
    const int4 e = q.e[i];
    const int32_t key = due ? e.x : 0, ok = e.z;
    g_sk = e.y;
    g_qty = e.w;
    g_amt = q.amt[i];
    g_in = due && (uint32_t)g_sk <= dm.maxsk && (uint32_t)ok <= dm.maxok;
    g_nat = s.snat[g_in ? (uint32_t)g_sk : 0u];
    g_y = year_at<NIB>(s.oyear, g_in ? (uint32_t)ok : 0u);
    const int4* row = reinterpret_cast<const int4*>(s.pst + (size_t)(uint32_t)key * PS_WAYS);
    g_ra = row[0];
    g_rb = row[1];
    head += due ? 64u : 0u;
    gpend = due;
  };
  auto batch_finish = [&]() {
    int64_t cost = g_ra.x == g_sk ? g_ra.y : g_ra.z == g_sk ? g_ra.w
                 : g_rb.x == g_sk ? g_rb.y : g_rb.z == g_sk ? g_rb.w : -1;
    if (g_sk < 0) cost = -1;
    if (g_in && g_nat >= 0 && g_y != NO_YEAR && cost >= 0) {
      const int slot = g_nat * dm.ny + g_y;
      atomicAdd((unsigned long long*)&ltab[slot],
                (unsigned long long)(g_amt - cost * (int64_t)g_qty * 100));
      lhit[slot] = 1u;
    }
    gpend = false;
  };
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  // Two register buffers of U groups each (this trip's and the other one);
  // the bitmap words loaded for the last tested trip (word 0 where the folded
  // filter missed) and its fold hits, one bit per row.
  // The column loads are issued unconditionally (lanes past the end reload
  // the last vector group and are masked by their buffer's base gbuf[b] in
  // the test): a load under a branch leaves a path where it was not issued,
  // and every wait after the join then has to count for that path too.
  int4 p4[2][U], s4[2][U], o4[2][U], q4[2][U], d4[2][U];
  longlong2 pa[2][U], pb[2][U];
  uint32_t gbuf[2] = {0u, 0u};
  PartWord pw[U][4];
  uint32_t hits = 0;
  auto load = [&](uint32_t gb, auto B) {
    constexpr int b = decltype(B)::value;
    gbuf[b] = gb;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = min(gb + u * nthr + lane, nvec - 1);
      p4[b][u] = ld4nt(f.pk, g);
      s4[b][u] = ld4nt(f.sk, g);
      o4[b][u] = ld4nt(f.ok, g);
      q4[b][u] = ld4nt(f.qty, g);
      d4[b][u] = ld4nt(f.disc, g);
      pa[b][u] = ld2nt(f.price, 2 * g);
      pb[b][u] = ld2nt(f.price, 2 * g + 1);
    }
  };
  // C: fold test of buffer b's keys; bitmap words for the fold hits.
  auto test = [&](auto B) {
    constexpr int b = decltype(B)::value;
    hits = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t ks[4] = {(uint32_t)p4[b][u].x, (uint32_t)p4[b][u].y, (uint32_t)p4[b][u].z,
                              (uint32_t)p4[b][u].w};
      const bool valid = gbuf[b] + u * nthr + lane < nvec;
      // the four folded words are read before any is tested (kf >> 5 is
      // always inside the fold), so the LDS reads overlap
      uint32_t fw[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) fw[j] = lf[(ks[j] & fmask) >> 5];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t bit = (fw[j] >> (ks[j] & 31)) & 1u;
        const uint32_t hit = (valid ? 1u : 0u) & (ks[j] <= dm.maxpk ? 1u : 0u) & bit;
        hits |= hit << (4 * u + j);
        pw[u][j] = s.pbw[hit ? ks[j] >> 5 : 0u];
      }
    }
  };
  // A: queue buffer b's rows whose part bit is set. Each lane's place in the
  // queue is the wave's exclusive prefix of its pass count, formed from one
  // ballot per bit of the count; then row slot by row slot the passing lanes
  // write their row. When the trip's rows might not fit, they are pushed slot
  // by slot, processing the queue as it fills.
  auto push_trip = [&](auto B) {
    constexpr int b = decltype(B)::value;
    uint32_t pmask = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t ks[4] = {(uint32_t)p4[b][u].x, (uint32_t)p4[b][u].y, (uint32_t)p4[b][u].z,
                              (uint32_t)p4[b][u].w};
#pragma unroll
      for (int j = 0; j < 4; ++j)
        pmask |= (((hits >> (4 * u + j)) & (pw[u][j].bits >> (ks[j] & 31))) & 1u) << (4 * u + j);
    }
    const uint32_t c = (uint32_t)__popc(pmask);
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int bit = 0; (1 << bit) <= 4 * U; ++bit) {
      const uint64_t m = __ballot((c >> bit) & 1u);
      pre += (uint32_t)__popcll(m & below) << bit;
      tot += (uint32_t)__popcll(m) << bit;
    }
    if (tot == 0) return;
    const bool fits = tail - head + tot <= QCAP;
    uint32_t pos = tail + pre;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t ps[4] = {p4[b][u].x, p4[b][u].y, p4[b][u].z, p4[b][u].w};
      const int32_t ss[4] = {s4[b][u].x, s4[b][u].y, s4[b][u].z, s4[b][u].w};
      const int32_t os[4] = {o4[b][u].x, o4[b][u].y, o4[b][u].z, o4[b][u].w};
      const int32_t qs[4] = {q4[b][u].x, q4[b][u].y, q4[b][u].z, q4[b][u].w};
      const int32_t ds[4] = {d4[b][u].x, d4[b][u].y, d4[b][u].z, d4[b][u].w};
      const int64_t pr[4] = {pa[b][u].x, pa[b][u].y, pb[b][u].x, pb[b][u].y};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool pass = (pmask >> (4 * u + j)) & 1u;
        const int32_t key =
            DIRECT ? (int32_t)part_rank(pw[u][j], (uint32_t)ps[j]) : ps[j];
        const int64_t amt = pr[j] * (int64_t)(100 - ds[j]);
        if (fits) {
          if (pass) {
            const uint32_t i = pos & (QCAP - 1);
            q.e[i] = int4{key, ss[j], os[j], qs[j]};
            q.amt[i] = amt;
            ++pos;
          }
        } else {
          push(pass, key, ss[j], os[j], qs[j], amt);
        }
      }
    }
    if (fits) tail += tot;
  };
  // One trip: buffer b holds this trip's columns (in flight), buffer o the
  // previous trip's (tested, not yet queued) if `prev`.
  bool prev = false;  // wave-uniform
  auto trip = [&](uint32_t gb, auto B, auto O) {
    if (DIRECT && gpend) batch_finish();             // B
    if (prev) push_trip(O);                          // A
    test(B);                                         // C
    prev = true;
    if (DIRECT) batch_issue();                       // D
    else if (tail - head >= 64) drain_now(63);
    __builtin_amdgcn_sched_barrier(0);
    load(gb + U * nthr, O);                          // E (past the end: masked)
  };
  constexpr std::integral_constant<int, 0> B0{};
  constexpr std::integral_constant<int, 1> B1{};
  const uint32_t g0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
  const uint32_t step = U * nthr;
  bool last1 = false;  // the last tested trip is in buffer 1
  if (g0 < nvec) load(g0, B0);
  for (uint32_t gb = g0; gb < nvec; gb += 2 * step) {
    trip(gb, B0, B1);
    if (gb + step >= nvec) break;
    trip(gb + step, B1, B0);
    last1 = true;
    if (gb + 2 * step >= nvec) break;
    last1 = false;
  }
  if (DIRECT && gpend) batch_finish();
  if (prev) {
    if (last1) push_trip(B1);
    else push_trip(B0);
  }
  if (blockIdx.x == 0 && w == 0) {  // wave-uniform: the last n % 4 rows
    const uint32_t r = nvec * 4 + lane;
    const bool in = lane < (n & 3);
    const uint32_t k = in ? (uint32_t)f.pk[r] : 0xFFFFFFFFu;
    const uint32_t kf = k & fmask;
    PartWord wd{0u, 0u};
    if (k <= dm.maxpk && ((lf[kf >> 5] >> (kf & 31)) & 1u)) wd = s.pbw[k >> 5];
    const bool pass = (wd.bits >> (k & 31)) & 1u;
    push(pass, DIRECT ? (int32_t)part_rank(wd, k) : (int32_t)k, pass ? f.sk[r] : 0,
         pass ? f.ok[r] : 0, pass ? f.qty[r] : 0,
         pass ? f.price[r] * (int64_t)(100 - f.disc[r]) : 0);
  }
  drain_now(0);
  __syncthreads();
  for (int i = threadIdx.x; i < dm.nslots; i += blockDim.x) {
    if (lhit[i]) {
      const int gslot = (i / dm.ny) * NYEAR + dm.first + i % dm.ny;
      atomicAdd((unsigned long long*)&s.agg[gslot], (unsigned long long)ltab[i]);
      s.agg_hit[gslot] = 1u;
    }
  }
}

// Days from 1970-01-01 to y-01-01 in the proleptic Gregorian calendar.
int64_t days_to_jan1(int64_t y) {
  const int64_t yp = y - 1;  // March-based year of 1 January
  const int64_t era = (yp >= 0 ? yp : yp - 399) / 400;
  const int64_t yoe = yp - era * 400;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + 306;  // 306 = Mar 1 .. Jan 1
  return era * 146097 + doe - 719468;
}
constexpr int64_t DAY0_1992 = 8035;  // 1992-01-01 - 1970-01-01

// Year of a day number (day 0 = 1992-01-01).
int64_t year_of(int64_t day) {
  int64_t y = 1992 + day / 366;
  while (days_to_jan1(y + 1) - DAY0_1992 <= day) ++y;
  while (days_to_jan1(y) - DAY0_1992 > day) --y;
  return y;
}

}  // namespace

// One run of Q9. With try_direct the partsupp costs go into the direct
// per-part table when the data allows (costs fit int32, the table fits);
// *redo is set when a part turned out to have more suppliers than the table
// holds, and the caller runs the query again through the hash table.
static int q9_run(gdb_ctx* ctx, const gdb_table* part, const gdb_table* supplier,
                  const gdb_table* partsupp, const gdb_table* orders, const gdb_table* lineitem,
                  const gdb_q9_params* prm, gdb_q9_result* out, gdb_timing* timing,
                  bool try_direct, bool* redo) {
  *redo = false;
  if (!ctx || !part || !supplier || !partsupp || !orders || !lineitem || !prm || !out)
    return fail(GDB_E_INVALID, "null argument");
  PatternW pat;
  if (!make_pattern(prm->pattern, &pat)) return fail(GDB_E_INVALID, "Q9 pattern must be 1..31 bytes");
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
  GDB_TRY(sum_fits(lineitem->nrows,
                   (unsigned __int128)magnitude(lp) * (100u + magnitude(ld)) +
                       (unsigned __int128)magnitude(psc) * magnitude(lq) * 100u,
                   "Q9 SUM(amount)"));
  GDB_TRY(require_resident(part, "Q9 part"));
  GDB_TRY(require_resident(supplier, "Q9 supplier"));
  GDB_TRY(require_resident(partsupp, "Q9 partsupp"));
  GDB_TRY(require_resident(orders, "Q9 orders"));
  if ((part->nrows && ppk->minv < 0) || (supplier->nrows && ssk->minv < 0) ||
      (orders->nrows && ook->minv < 0))
    return fail(GDB_E_UNSUPPORTED, "Q9 needs non-negative partkey, suppkey and orderkey values");
  if (supplier->nrows && (snk->minv < 0 || snk->maxv >= NNAT))
    return fail(GDB_E_UNSUPPORTED, "Q9 nation keys must be in 0..63");
  // The order dates' years, 1970..2033, index the LDS table; their Jan-1 day
  // numbers let the orders kernel find a date's year with compares only.
  Years years{};
  years.ny = 1;
  years.first = 0;
  if (orders->nrows) {
    const int64_t y0 = year_of(ood->minv), y1 = year_of(ood->maxv);
    if (y0 < YEAR0 || y1 >= YEAR0 + NYEAR)
      return fail(GDB_E_UNSUPPORTED, "Q9 order years must be in 1970..2033");
    years.ny = (int32_t)(y1 - y0 + 1);
    years.first = (int32_t)(y0 - YEAR0);
    for (int i = 0; i < years.ny; ++i) years.jan1[i] = (int32_t)(days_to_jan1(y0 + i) - DAY0_1992);
  }
  const int nnat = supplier->nrows ? (int)snk->maxv + 1 : 1;
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t np = (uint32_t)part->nrows, ns = (uint32_t)supplier->nrows,
                 nps = (uint32_t)partsupp->nrows, no = (uint32_t)orders->nrows;
  const uint64_t nl = lineitem->nrows;
  const uint32_t maxpk = (uint32_t)std::max<int64_t>(0, ppk->maxv);
  const uint32_t maxsk = (uint32_t)std::max<int64_t>(0, ssk->maxv);
  const uint32_t maxok = (uint32_t)std::max<int64_t>(0, ook->maxv);
  const uint32_t npw = maxpk / 32 + 1, nsupp = maxsk + 1, nord = maxok + 1;
  // LDS part filter: the bitmap itself when it fits 2^20 bits, else folded;
  // halved until it fits next to the (nation, year) table and the queues.
  uint32_t fwords = 32;
  while (fwords < npw && fwords < (1u << 15)) fwords <<= 1;
  {
    const size_t lds_max = (size_t)std::max(ctx->info.lds_bytes_per_cu, 65536) -
                           sizeof(Q9Queue) * (PROBE_THREADS / 64) - 1024;
    while (fwords > 32 && (size_t)fwords * 4 + (size_t)nnat * years.ny * 12 > lds_max) fwords >>= 1;
  }
  // Passing parts (ranks) are at most the part rows and the partkey values.
  const uint64_t nrank = std::max<uint64_t>(1, std::min<uint64_t>(np, maxpk + 1ull));
  const uint32_t rblocks = (npw + RANK_WORDS - 1) / RANK_WORDS;
  const bool direct = try_direct && nps > 0 && psc->minv >= INT32_MIN && psc->maxv <= INT32_MAX;
  uint64_t capmax = 1024;
  while (!direct && capmax < 2ull * nps) capmax <<= 1;
  const uint32_t nlist = direct ? 1u : std::max<uint32_t>(nps, 1);
  // Supplier-map prefix to clear: every entry from the first suppkey on is
  // written when suppkeys are unique and gapless. (The year map is cleared
  // whole: marking the gaps between orderkeys in q9_orders instead took
  // longer than the clear.)
  const bool sup_dense = ns && ssk->strict && (uint64_t)(ssk->maxv - ssk->minv) + 1 == ns;
  const uint32_t snat_clear = sup_dense ? (uint32_t)ssk->minv : nsupp;
  // 4-bit year codes for at most 15 years and unique (strictly ascending)
  // orderkeys; else a byte per key value.
  const bool nib = years.ny <= 15 && (no == 0 || ook->strict);
  const uint32_t oy_words = nib ? (nord + 7) / 8 : (nord + 3) / 4;
  // Codes written by q9_orders_owned (no clear, no atomics) unless the keys
  // are sparser than 1 in 16, where a slice's word span gets long.
  const bool owned = nib && no > 0 && (uint64_t)nord <= 16ull * no;
  const uint32_t oy_clear = owned ? 0u : oy_words;
  Q9Scratch s{};
  {
    char* b;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) & ~size_t(255);
      return o;
    };
    // agg, agg_hit and counters are adjacent: the result is one readback.
    const size_t o_pbw = take(npw * 8ull), o_fold = take(fwords * 4ull),
                 o_rsum = take(rblocks * 4ull), o_snat = take(nsupp * 4ull),
                 o_oy = take(oy_words * 4ull), o_agg = take(RESULT_BYTES),
                 o_lpk = take(nlist * 4ull), o_lsk = take(nlist * 4ull), o_lc = take(nlist * 8ull);
    GDB_TRY(arena(ctx, A_MAP, off, (void**)&b));
    s.pbw = (PartWord*)(b + o_pbw);
    s.pfold = (uint32_t*)(b + o_fold);
    s.rsum = (uint32_t*)(b + o_rsum);
    s.snat = (int32_t*)(b + o_snat);
    s.oyear = (uint8_t*)(b + o_oy);
    s.agg = (int64_t*)(b + o_agg);
    s.agg_hit = (uint32_t*)(b + o_agg + NNAT * NYEAR * 8);
    s.counters = (uint32_t*)(b + o_agg + NNAT * NYEAR * 12);
    s.lpk = (int32_t*)(b + o_lpk);
    s.lsk = (int32_t*)(b + o_lsk);
    s.lcost = (int64_t*)(b + o_lc);
    if (direct) {
      const size_t tab = (nrank * PS_WAYS * 8 + 255) & ~size_t(255);
      char* t;
      GDB_TRY(arena(ctx, A_HT, tab + nrank * 4 + 256, (void**)&t));
      s.pst = (int2*)t;
      s.pcnt = (uint32_t*)(t + tab);
    } else {
      GDB_TRY(arena(ctx, A_HT, capmax * 16 + 256, (void**)&s.ht));
    }
  }
  char *host, *hdev;
  const int nslots = nnat * years.ny;
  GDB_TRY(result_buf(ctx, (size_t)nslots * 12 + 16, (void**)&host, (void**)&hdev));
  hipStream_t st = ctx->stream;
  const int B = 256;
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  q9_clear<<<std::min<uint32_t>(8192, (npw + snat_clear + oy_clear + NNAT * NYEAR + B - 1) / B), B,
             0, st>>>(s, npw, snat_clear, oy_clear);
  const Q9DimArgs da{(const uint8_t*)pnm->dev, (const int32_t*)ppk->dev, np, pat, maxpk, s.pbw,
                     (const int32_t*)ook->dev,
                     (const int32_t*)ood->dev, no, maxok, years, s.oyear,
                     (const int32_t*)ssk->dev, (const int32_t*)snk->dev, ns, maxsk, s.snat};
  if (np) {
    const int pb = std::max(1, std::min<int>(resident_blocks(ctx, (const void*)q9_part, TILE, 0, 0),
                                             (int)((np + TILE - 1) / TILE)));
    q9_part<<<pb, TILE, 0, st>>>(da);
  }
  if (owned) {
    q9_orders_owned<<<(no + OY_SLICE - 1) / OY_SLICE, 256, 0, st>>>(da, oy_words);
  } else if (no) {
    if (nib) q9_orders<true><<<((no + 3) / 4 + B - 1) / B, B, 0, st>>>(da);
    else q9_orders<false><<<((no + 3) / 4 + B - 1) / B, B, 0, st>>>(da);
  }
  {
    const uint32_t fb = (fwords + B - 1) / B, sb = (ns + B - 1) / B;
    q9_after_part<<<rblocks + fb + sb, B, 0, st>>>(s.pbw, npw, fwords, s.pfold, s.rsum, rblocks,
                                                   fb, da);
  }
  q9_rank_write<<<rblocks, B, 0, st>>>(s.pbw, npw, s.rsum, direct ? s.pst : nullptr, s.pcnt);
  if (direct) {
    q9_partsupp_direct<<<std::min<uint32_t>(ctx->info.cu_count * 8, ((nps + 3) / 4 + B - 1) / B), B, 0,
                         st>>>((const int32_t*)pspk->dev, (const int32_t*)pssk->dev,
                               (const int64_t*)psc->dev, nps, maxpk, s.pbw, s.pst, s.pcnt,
                               s.counters);
  } else {
    if (nps)
      q9_partsupp<<<std::min<uint32_t>(ctx->info.cu_count * 8, (nps + 16 * B - 1) / (16 * B)), B, 0,
                    st>>>((const int32_t*)pspk->dev, (const int32_t*)pssk->dev,
                          (const int64_t*)psc->dev, nps, maxpk, s.pbw, s.counters, s.lpk, s.lsk,
                          s.lcost);
    const int fill_blocks = ctx->info.cu_count * 4;
    q9_ps_clear<<<fill_blocks, B, 0, st>>>(s.counters, s.ht);
    q9_ps_insert<<<fill_blocks, B, 0, st>>>(s.counters, s.lpk, s.lsk, s.lcost, s.ht);
  }
  GDB_HIP(hipGetLastError());
  const Q9Dims dm{maxpk, maxsk, maxok, years.ny, years.first, nnat * years.ny, fwords,
                  direct ? 1 : 0};
  const size_t tab_lds = (size_t)fwords * 4 + (size_t)dm.nslots * 12;
  // One 4-row group per lane and trip: the pipelined probe holds two trips'
  // columns plus the dependent loads in flight in registers (116 VGPRs at
  // one group; two groups spill).
  const void* kfn = direct ? (nib ? (const void*)q9_probe<PROBE_U, true, true>
                                  : (const void*)q9_probe<PROBE_U, true, false>)
                           : (nib ? (const void*)q9_probe<PROBE_U, false, true>
                                  : (const void*)q9_probe<PROBE_U, false, false>);
  GDB_TRY(allow_dynamic_lds(ctx, kfn, tab_lds));
  const int PB = PROBE_THREADS;
  const int president = resident_blocks(ctx, kfn, PB, tab_lds, ctx->tune[T_Q9].blocks_per_cu);
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const Q9Fact f{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                   (const int32_t*)d[3], (const int32_t*)d[4], (const int64_t*)d[5]};
    const int pblocks = std::max(1, std::min<int>(president, (int)((rows / 4 + PB - 1) / PB)));
    if (direct && nib) q9_probe<PROBE_U, true, true><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
    else if (direct) q9_probe<PROBE_U, true, false><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
    else if (nib) q9_probe<PROBE_U, false, true><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
    else q9_probe<PROBE_U, false, false><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  const Column* fc[6] = {lpk, lsk, lok, lq, ld, lp};
  RunStats rs;
  if (nl) {
    GDB_TRY(run_over(ctx, lineitem, fc, 6, consume, &rs));
  } else {
    GDB_HIP(kernel_mark(ctx, 1, st));
    GDB_HIP(kernel_mark(ctx, 2, st));
  }
  q9_publish<<<1, 256, 0, st>>>(s, years.ny, years.first, nslots, hdev);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  const uint64_t bytes = 28ull * nl + 59ull * np + 16ull * nps + 8ull * ns + 8ull * no;
  GDB_TRY(timing_finish(ctx, timing, bytes, 28ull * nl, &rs));
  const int64_t* agg = (const int64_t*)host;
  const uint32_t* hit = (const uint32_t*)(host + (size_t)nslots * 8);
  if (direct && hit[nslots + 1]) {
    *redo = true;  // a part with more than PS_WAYS suppliers
    return GDB_OK;
  }
  // Output order: nation name, then nationkey, then year descending. Every
  // row is a (nation, year) slot, so the nations are walked in name order
  // (ties by key) and each nation's years downward: no comparison sort.
  static const std::array<uint8_t, NNAT> by_name = [] {
    std::array<uint8_t, NNAT> o{};
    for (int k = 0; k < NNAT; ++k) o[k] = (uint8_t)k;
    std::stable_sort(o.begin(), o.end(), [](uint8_t a, uint8_t b) {
      return strcmp(gdb_nation_name(a), gdb_nation_name(b)) < 0;
    });
    return o;
  }();
  size_t nrows = 0;
  for (int i = 0; i < nslots; ++i) nrows += hit[i] != 0u;
  if (nrows > sizeof(out->rows) / sizeof(out->rows[0]))
    return fail(GDB_E_UNSUPPORTED, "more (nation, year) groups than the result holds");
  memset(out, 0, sizeof(*out));
  for (const uint8_t k : by_name) {
    if (k >= nnat) continue;
    for (int y = years.ny - 1; y >= 0; --y) {
      const int i = k * years.ny + y;
      if (hit[i]) out->rows[out->nrows++] = gdb_q9_row{k, YEAR0 + years.first + y, agg[i]};
    }
  }
  return GDB_OK;
}

extern "C" int gdb_q9(gdb_ctx* ctx, const gdb_table* part, const gdb_table* supplier,
                      const gdb_table* partsupp, const gdb_table* orders,
                      const gdb_table* lineitem, const gdb_q9_params* prm, gdb_q9_result* out,
                      gdb_timing* timing) {
  bool redo = false;
  GDB_TRY(q9_run(ctx, part, supplier, partsupp, orders, lineitem, prm, out, timing, true, &redo));
  if (redo)
    GDB_TRY(q9_run(ctx, part, supplier, partsupp, orders, lineitem, prm, out, timing, false, &redo));
  return GDB_OK;
}
