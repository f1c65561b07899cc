// Q9 multi-way join with a substring filter, grouped by (nation, year)
// (reference kernels/q9.metal:44-394, host src/q9.cpp:8-241, post-processing
// src/infra.h:733-776, SQL sql/q9.sql).
//
//   clear      part bitmap, counters, supplier->nation map (-1), orderkey->year map
//   part       tiles of 256 names copied contiguously into LDS with dwordx4
//              loads; bit[partkey] if the name contains the pattern
//              (strmatch.h; the reference's SWAR 'g' scan, kernels/q9.metal:58-83,
//              generalised to any pattern)
//   supplier   nation[suppkey]
//   partsupp   rows whose part passed -> compact list (partkey, suppkey, cost)
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

struct Q9Scratch {
  uint32_t* pbm;       // part bitmap
  uint32_t* pfold;     // part bitmap folded to the probe's LDS filter size
  uint32_t* counters;  // [0] partsupp list count, [1] a part with more than PS_WAYS suppliers
  int32_t* snat;       // suppkey -> nation or -1
  uint8_t* oyear;      // orderkey -> year index or 0xFF
  int32_t* lpk;        // compact partsupp list
  int32_t* lsk;
  int64_t* lcost;
  longlong2* ht;       // {(partkey, suppkey) or EMPTY64, cost}
  int2* pst;           // direct table: PS_WAYS {suppkey or -1, cost} per partkey
  uint32_t* pcnt;      // direct table: slots taken per partkey
  int64_t* agg;        // [NNAT*NYEAR] profit
  uint32_t* agg_hit;   // [NNAT*NYEAR]
};

// Clears the part bitmap, the first nsupp supplier-map entries, the first
// oywords words of the year map and the result table (the map entries past
// the cleared prefixes are all written by q9_supplier / q9_orders).
__global__ __launch_bounds__(256) void q9_clear(Q9Scratch s, uint32_t npw, uint32_t nsupp,
                                                uint32_t oywords) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t a = npw, b = a + nsupp, c = b + oywords, d = c + NNAT * NYEAR;
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

// Arguments of the dimension-side kernels (part filter, year map, nation map).
struct Q9DimArgs {
  const uint8_t* name;
  const int32_t* partkey;
  uint32_t np;
  PatternW pat;
  uint32_t maxpk;
  uint32_t* pbm;
  int2* pst;       // direct partsupp table to set up, or null
  uint32_t* pcnt;
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
// registers. Names with the pattern set their part's bit and, for the direct
// partsupp table, empty their part's supplier slots.
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
      if (pk <= a.maxpk) {
        atomicOr(&a.pbm[pk >> 5], 1u << (pk & 31));
        if (a.pst) {  // the part's supplier slots start empty
          int4* row = reinterpret_cast<int4*>(a.pst + (size_t)pk * PS_WAYS);
          row[0] = int4{-1, 0, -1, 0};
          row[1] = int4{-1, 0, -1, 0};
          a.pcnt[pk] = 0u;
        }
      }
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
__global__ __launch_bounds__(256) void q9_orders(Q9DimArgs a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (a.no + 3) / 4) return;
  const int4 k = ld4nt(a.ok, g), d = ld4nt(a.od, g);
  const int32_t ks[4] = {k.x, k.y, k.z, k.w}, ds[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (4 * g + j < a.no && (uint32_t)ks[j] <= a.maxok)
      a.oyear[ks[j]] = (uint8_t)year_index(ds[j], a.years);
}

__global__ __launch_bounds__(256) void q9_supplier(Q9DimArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.ns && (uint32_t)a.sk[i] <= a.maxsk) a.snat[a.sk[i]] = a.nat[i];
}

// The part bitmap folded onto fwords words: bit k of the fold is the OR of
// the bits of every key congruent to k modulo 32 * fwords.
__global__ __launch_bounds__(256) void q9_fold(const uint32_t* __restrict__ pbm, uint32_t npw,
                                               uint32_t fwords, uint32_t* __restrict__ fold) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= fwords) return;
  uint32_t v = 0;
  for (uint32_t i = w; i < npw; i += fwords) v |= pbm[i];
  fold[w] = v;
}

// Partsupp rows whose part passed -> compact list in row order, one atomic per
// tile of 16 x 256 rows: thread t tests rows 4t..4t+3 of each of the tile's
// four 1024-row quarters with dwordx4 loads of ps_partkey.
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
    for (int j = 0; j < 16; ++j) words[j] = pbm[min((uint32_t)ks[j], maxpk) >> 5];
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
// its part (a per-part counter: the adds of one thread's rows are
// independent, so they are in flight together); a fifth supplier of a part
// sets counters[1] and the query runs again through the hash table.
// (Claiming slots by compare-and-swap of the suppkey took up to ten dependent
// round trips per thread, a part's rows being adjacent.)
__global__ __launch_bounds__(256) void q9_partsupp_direct(const int32_t* __restrict__ pk,
                                                          const int32_t* __restrict__ sk,
                                                          const int64_t* __restrict__ cost,
                                                          uint32_t n, uint32_t maxpk,
                                                          const uint32_t* __restrict__ pbm,
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
    uint32_t words[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) words[j] = pbm[min((uint32_t)ks[j], maxpk) >> 5];
    bool green[4];
    uint32_t slot[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t p = (uint32_t)ks[j];
      green[j] = p <= maxpk && ((words[j] >> (p & 31)) & 1u);
      slot[j] = green[j] ? atomicAdd(&pcnt[p], 1u) : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!green[j]) continue;
      const uint32_t r = 4 * g + j;
      if (slot[j] < PS_WAYS)
        pst[(size_t)(uint32_t)ks[j] * PS_WAYS + slot[j]] = int2{sk[r], (int32_t)cost[r]};
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
  int32_t exact;     // 1: every partkey is below 32 * fwords, the fold is the bitmap
  int32_t direct;    // 1: partsupp costs in the direct table, else the hash table
};

// One queued row: nation, year and the first partsupp slot are loaded
// together (none depends on another), then the rare further probe steps.
__device__ __forceinline__ void q9_entry(int32_t pk, int32_t sk, int32_t ok, int32_t qty,
                                         int64_t amt, uint32_t cap, const Q9Dims& dm,
                                         const Q9Scratch& s, int64_t* ltab, uint32_t* lhit) {
  const bool inr = (uint32_t)sk <= dm.maxsk && (uint32_t)ok <= dm.maxok;
  const int32_t nat = s.snat[inr ? (uint32_t)sk : 0u];
  const uint8_t y = s.oyear[inr ? (uint32_t)ok : 0u];
  int64_t cost = -1;
  if (dm.direct) {  // the part passed the exact bitmap, so its slots were set up
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

struct Q9Queue {
  int32_t pk[QCAP], sk[QCAP], ok[QCAP], qty[QCAP];
  int64_t amt[QCAP];
};

// U groups of 4 rows per lane and step: the loads of all six columns for the
// U groups are issued first; then, row slot by row slot, lanes whose row's part
// bit is set push the row's values onto the wave's queue, and whenever 64 are
// queued one full wave of entries is processed.
// The part test reads the folded filter in LDS: random 4-byte gathers from
// the 250 KB global bitmap cost a vector-cache line per lane and took as long
// as streaming all six columns. Only keys whose folded bit is set read the
// global bitmap (none when the fold is exact). One workgroup of 1024 threads
// per CU shares one fold (2^19 bits next to sixteen wave queues): sixteen
// waves per CU stream the columns, against eight with 512-thread workgroups
// and a 2^20-bit fold (466 -> 412 us at 60M fact rows; the global bitmap reads
// for the extra fold hits cost nothing measurable).
constexpr int PROBE_THREADS = 1024;
template <int U>
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
  auto push = [&](bool pass, int32_t pk, int32_t sk, int32_t ok, int32_t qty, int64_t amt) {
    const uint64_t b = __ballot(pass);
    if (pass) {
      const uint32_t i = (tail + (uint32_t)__popcll(b & below)) & (QCAP - 1);
      q.pk[i] = pk;
      q.sk[i] = sk;
      q.ok[i] = ok;
      q.qty[i] = qty;
      q.amt[i] = amt;
    }
    tail += (uint32_t)__popcll(b);
  };
  auto drain = [&](bool all) {
    wave_lds_sync();
    while (tail - head >= 64 || (all && tail != head)) {
      const uint32_t avail = min(64u, tail - head);
      if (lane < avail) {
        const uint32_t i = (head + lane) & (QCAP - 1);
        q9_entry(q.pk[i], q.sk[i], q.ok[i], q.qty[i], q.amt[i], cap, dm, s, ltab, lhit);
      }
      head += avail;
    }
    wave_lds_sync();
  };
  auto test = [&](int32_t pk) {
    const uint32_t k = (uint32_t)pk, kf = k & fmask;
    bool pass = k <= dm.maxpk && ((lf[kf >> 5] >> (kf & 31)) & 1u);
    if (!dm.exact && pass) pass = (s.pbm[k >> 5] >> (k & 31)) & 1u;
    return pass;
  };
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  // Two register buffers of U groups each: the loads of the next trip are in
  // flight while the current trip's rows are tested, queued and drained.
  // Trips are wave-uniform: every lane of a wave runs the same iterations so
  // the ballots and queue counters stay in step; lanes past the end load
  // nothing (partkey -1 fails the test).
  int4 p4[2][U], s4[2][U], o4[2][U], q4[2][U], d4[2][U];
  longlong2 pa[2][U], pb[2][U];
  auto load = [&](uint32_t gb, auto B) {
    constexpr int b = decltype(B)::value;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = gb + u * nthr + lane;
      if (g < nvec) {
        p4[b][u] = ld4nt(f.pk, g);
        s4[b][u] = ld4nt(f.sk, g);
        o4[b][u] = ld4nt(f.ok, g);
        q4[b][u] = ld4nt(f.qty, g);
        d4[b][u] = ld4nt(f.disc, g);
        pa[b][u] = ld2nt(f.price, 2 * g);
        pb[b][u] = ld2nt(f.price, 2 * g + 1);
      } else {
        p4[b][u] = int4{-1, -1, -1, -1};
        s4[b][u] = o4[b][u] = q4[b][u] = d4[b][u] = int4{0, 0, 0, 0};
        pa[b][u].x = pa[b][u].y = pb[b][u].x = pb[b][u].y = 0;
      }
    }
  };
  auto process = [&](auto B) {
    constexpr int b = decltype(B)::value;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t ps[4] = {p4[b][u].x, p4[b][u].y, p4[b][u].z, p4[b][u].w};
      const int32_t ss[4] = {s4[b][u].x, s4[b][u].y, s4[b][u].z, s4[b][u].w};
      const int32_t os[4] = {o4[b][u].x, o4[b][u].y, o4[b][u].z, o4[b][u].w};
      const int32_t qs[4] = {q4[b][u].x, q4[b][u].y, q4[b][u].z, q4[b][u].w};
      const int32_t ds[4] = {d4[b][u].x, d4[b][u].y, d4[b][u].z, d4[b][u].w};
      const int64_t pr[4] = {pa[b][u].x, pa[b][u].y, pb[b][u].x, pb[b][u].y};
      bool pass[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) pass[j] = test(ps[j]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        push(pass[j], ps[j], ss[j], os[j], qs[j], pr[j] * (int64_t)(100 - ds[j]));
        if (tail - head >= 64) drain(false);
      }
    }
  };
  constexpr std::integral_constant<int, 0> B0{};
  constexpr std::integral_constant<int, 1> B1{};
  const uint32_t g0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
  const uint32_t step = U * nthr;
  if (g0 < nvec) load(g0, B0);
  for (uint32_t gb = g0; gb < nvec; gb += 2 * step) {
    const bool more = gb + step < nvec;
    if (more) load(gb + step, B1);
    process(B0);
    if (!more) break;
    if (gb + 2 * step < nvec) load(gb + 2 * step, B0);
    process(B1);
  }
  if (blockIdx.x == 0 && w == 0) {  // wave-uniform: the last n % 4 rows
    const uint32_t r = nvec * 4 + lane;
    const bool in = lane < (n & 3);
    const int32_t pk = in ? f.pk[r] : -1;
    const bool pass = in && test(pk);
    if (tail - head > QCAP - 64) drain(false);
    push(pass, pk, pass ? f.sk[r] : 0, pass ? f.ok[r] : 0, pass ? f.qty[r] : 0,
         pass ? f.price[r] * (int64_t)(100 - f.disc[r]) : 0);
  }
  drain(true);
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
  const bool exact = npw <= fwords;
  const bool direct = try_direct && nps > 0 && psc->minv >= INT32_MIN && psc->maxv <= INT32_MAX &&
                      (uint64_t)(maxpk + 1ull) * PS_WAYS * 8 <= (16ull << 30);
  uint64_t capmax = 1024;
  while (!direct && capmax < 2ull * nps) capmax <<= 1;
  const uint32_t nlist = direct ? 1u : std::max<uint32_t>(nps, 1);
  // Supplier-map prefix to clear: every entry from the first suppkey on is
  // written when suppkeys are unique and gapless. (The year map is cleared
  // whole: marking the gaps between orderkeys in q9_orders instead took
  // longer than the clear.)
  const bool sup_dense = ns && ssk->strict && (uint64_t)(ssk->maxv - ssk->minv) + 1 == ns;
  const uint32_t snat_clear = sup_dense ? (uint32_t)ssk->minv : nsupp;
  const uint32_t oy_words = (nord + 3) / 4;
  Q9Scratch s{};
  {
    char* b;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) & ~size_t(255);
      return o;
    };
    const size_t o_pbm = take(npw * 4ull), o_fold = take(fwords * 4ull), o_cnt = take(16),
                 o_snat = take(nsupp * 4ull),
                 o_oy = take((nord + 3) / 4 * 4ull), o_agg = take(NNAT * NYEAR * 8),
                 o_hit = take(NNAT * NYEAR * 4), o_lpk = take(nlist * 4ull),
                 o_lsk = take(nlist * 4ull), o_lc = take(nlist * 8ull);
    GDB_TRY(arena(ctx, A_MAP, off, (void**)&b));
    s.pbm = (uint32_t*)(b + o_pbm);
    s.pfold = (uint32_t*)(b + o_fold);
    s.counters = (uint32_t*)(b + o_cnt);
    s.snat = (int32_t*)(b + o_snat);
    s.oyear = (uint8_t*)(b + o_oy);
    s.agg = (int64_t*)(b + o_agg);
    s.agg_hit = (uint32_t*)(b + o_hit);
    s.lpk = (int32_t*)(b + o_lpk);
    s.lsk = (int32_t*)(b + o_lsk);
    s.lcost = (int64_t*)(b + o_lc);
    if (direct) {
      const size_t tab = ((size_t)(maxpk + 1ull) * PS_WAYS * 8 + 255) & ~size_t(255);
      char* t;
      GDB_TRY(arena(ctx, A_HT, tab + (maxpk + 1ull) * 4 + 256, (void**)&t));
      s.pst = (int2*)t;
      s.pcnt = (uint32_t*)(t + tab);
    } else {
      GDB_TRY(arena(ctx, A_HT, capmax * 16 + 256, (void**)&s.ht));
    }
  }
  char* host;
  GDB_TRY(pinned(ctx, NNAT * NYEAR * 12 + 64, (void**)&host));
  hipStream_t st = ctx->stream;
  const int B = 256;
  const uint32_t clear_items = npw + snat_clear + oy_words + NNAT * NYEAR;
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  q9_clear<<<std::min<uint32_t>(8192, (clear_items + B - 1) / B), B, 0, st>>>(s, npw, snat_clear,
                                                                              oy_words);
  {
    const Q9DimArgs da{(const uint8_t*)pnm->dev, (const int32_t*)ppk->dev, np, pat, maxpk, s.pbm,
                       direct ? s.pst : nullptr, s.pcnt, (const int32_t*)ook->dev,
                       (const int32_t*)ood->dev, no, maxok, years, s.oyear,
                       (const int32_t*)ssk->dev, (const int32_t*)snk->dev, ns, maxsk, s.snat};
    if (np) {
      const int pb = std::max(1, std::min<int>(resident_blocks(ctx, (const void*)q9_part, TILE, 0, 0),
                                               (int)((np + TILE - 1) / TILE)));
      q9_part<<<pb, TILE, 0, st>>>(da);
    }
    if (ns) q9_supplier<<<(ns + B - 1) / B, B, 0, st>>>(da);
    if (no) q9_orders<<<((no + 3) / 4 + B - 1) / B, B, 0, st>>>(da);
  }
  q9_fold<<<(fwords + B - 1) / B, B, 0, st>>>(s.pbm, npw, fwords, s.pfold);
  if (direct) {
    q9_partsupp_direct<<<std::min<uint32_t>(ctx->info.cu_count * 8, ((nps + 3) / 4 + B - 1) / B), B, 0,
                         st>>>((const int32_t*)pspk->dev, (const int32_t*)pssk->dev,
                               (const int64_t*)psc->dev, nps, maxpk, s.pbm, s.pst, s.pcnt,
                               s.counters);
  } else {
    if (nps)
      q9_partsupp<<<std::min<uint32_t>(ctx->info.cu_count * 8, (nps + 16 * B - 1) / (16 * B)), B, 0,
                    st>>>((const int32_t*)pspk->dev, (const int32_t*)pssk->dev,
                          (const int64_t*)psc->dev, nps, maxpk, s.pbm, s.counters, s.lpk, s.lsk,
                          s.lcost);
    const int fill_blocks = ctx->info.cu_count * 4;
    q9_ps_clear<<<fill_blocks, B, 0, st>>>(s.counters, s.ht);
    q9_ps_insert<<<fill_blocks, B, 0, st>>>(s.counters, s.lpk, s.lsk, s.lcost, s.ht);
  }
  GDB_HIP(hipGetLastError());
  const Q9Dims dm{maxpk, maxsk, maxok, years.ny, years.first, nnat * years.ny, fwords,
                  exact ? 1 : 0, direct ? 1 : 0};
  const size_t tab_lds = (size_t)fwords * 4 + (size_t)dm.nslots * 12;
  // 4-row groups in flight per lane: the tuning entry's unroll (1, 2 or 4)
  const int U = ctx->tune[T_Q9].unroll;
  const void* kfn = U >= 4 ? (const void*)q9_probe<4> : U == 1 ? (const void*)q9_probe<1>
                                                                : (const void*)q9_probe<2>;
  GDB_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)tab_lds));
  const int PB = PROBE_THREADS;
  const int president = resident_blocks(ctx, kfn, PB, tab_lds, ctx->tune[T_Q9].blocks_per_cu);
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const Q9Fact f{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                   (const int32_t*)d[3], (const int32_t*)d[4], (const int64_t*)d[5]};
    const int pblocks = std::max(1, std::min<int>(president, (int)((rows / 4 + PB - 1) / PB)));
    if (U >= 4) q9_probe<4><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
    else if (U == 1) q9_probe<1><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
    else q9_probe<2><<<pblocks, PB, tab_lds, st>>>(f, rows, dm, s);
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
  GDB_HIP(hipMemcpyAsync(host + NNAT * NYEAR * 12, s.counters + 1, 4, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  const uint64_t bytes = 28ull * nl + 59ull * np + 16ull * nps + 8ull * ns + 8ull * no;
  GDB_TRY(timing_finish(ctx, timing, bytes, 28ull * nl, &rs));
  if (direct && *(const uint32_t*)(host + NNAT * NYEAR * 12)) {
    *redo = true;  // a part with more than PS_WAYS suppliers
    return GDB_OK;
  }
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
