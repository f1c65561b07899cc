// Equi-join pair count, the `join` primitive (reference hash_join_build /
// hash_join_probe, kernels/microbenchmarks.metal:224-293; host
// src/microbenchmarks.cpp:135-251): the number of (build row, probe row) pairs
// with equal keys. Keys may repeat on both sides.
//
// Three strategies, chosen by the build side's size or by the tuning entry:
//   broadcast  (build <= LDS_KEYS keys) every workgroup builds the whole
//              build side as an LDS hash table and probes its slice of the
//              probe keys
//   radix      both sides are partitioned by the top B bits of a 32-bit key
//              hash (two passes of B/2 and B bits when B > 7, so every
//              pass writes long runs per partition: count, exclusive scan,
//              scatter), sized so a build partition fits an LDS table; one
//              workgroup per partition then builds its table in LDS and
//              probes the partition's probe keys against it. A partition
//              larger than the table is built and probed in table-sized
//              pieces. No global atomics anywhere.
//   global     the reference's shape: one open-addressing table in HBM built
//              with global CAS (kept for comparison by the autotuner)
//   direct     when the build keys' range [min, max] is dense (span <= 8 x the
//              build rows by default, <= 64 x when asked for): a bitmap over the
//              range, 1 bit per key
//              value (7.5 MB for 15M sparse order keys, so it stays in the
//              Infinity Cache), built with one atomicOr per word per lane and
//              checked for repeated build keys on the way; with repeats, a u32
//              count per key value instead. The probe streams its keys with
//              dwordx4 loads and adds each key's bit (or count).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "kcommon.h"
#include "scan_util.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int32_t EMPTY = INT32_MIN;
constexpr int JB = 256;               // threads per workgroup of every join kernel
constexpr uint32_t LDS_SLOTS = 8192;  // LDS table slots (64 KiB: keys + counts)
constexpr uint32_t LDS_KEYS = LDS_SLOTS / 2;  // keys per table fill (load <= 1/2)
constexpr int MAX_BITS = 14;

// MurmurHash3's 32-bit finaliser (public domain): every output bit depends on
// every key bit, so partitions stay balanced for dense or strided keys. The
// top bits pick the partition, the low bits the LDS slot.
__device__ __forceinline__ uint32_t mix32(int32_t key) {
  uint32_t h = (uint32_t)key;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// ---------------------------------------------------------------- partitioning
// A pass moves keys into partitions by (mix32(key) >> shift) & (PL - 1), PL
// <= 128 per pass. Mode 0: block b owns rows [b*chunk, (b+1)*chunk) of the
// input and its counts sit at hist[l*G + b]. Mode 1 (second pass): the input
// is the first pass's output; block b = fp*K + k owns the k-th of K slices of
// first-level partition fp, and its counts sit at hist[(fp*PL + l)*K + k], so
// the scanned array orders the final partitions fp*PL + l, each split by k.
constexpr int MAX_PL = 128;
constexpr int PT = 16;               // keys per thread and tile
constexpr int PTILE = PT * JB;       // 4096 keys per tile

struct PassArgs {
  const int32_t* in;
  int32_t* out;
  uint32_t n;            // mode 0: input rows
  uint32_t chunk;        // mode 0: rows per block
  int shift;
  uint32_t lmask;        // PL - 1
  int mode;
  const uint32_t* part_offs;   // mode 1: scanned first-pass counts
  uint32_t part_G;             // mode 1: first-pass blocks
  const uint32_t* part_total;  // mode 1: first-pass total
  uint32_t K;                  // mode 1: slices per first-level partition
};

__device__ __forceinline__ uint32_t local_id(const PassArgs& a, int32_t k) {
  return (mix32(k) >> a.shift) & a.lmask;
}

// Row range of this block and where its count of local partition l lives:
// hist[hbase + l * hstride].
__device__ __forceinline__ void block_range(const PassArgs& a, uint32_t& lo, uint32_t& hi,
                                            size_t& hbase, uint32_t& hstride) {
  if (a.mode == 0) {
    lo = min(a.n, blockIdx.x * a.chunk);
    hi = min(a.n, lo + a.chunk);
    hbase = blockIdx.x;
    hstride = gridDim.x;
  } else {
    const uint32_t P1 = gridDim.x / a.K, fp = blockIdx.x / a.K, k = blockIdx.x % a.K;
    const uint32_t s = a.part_offs[(size_t)fp * a.part_G];
    const uint32_t e = fp + 1 < P1 ? a.part_offs[(size_t)(fp + 1) * a.part_G] : *a.part_total;
    const uint32_t c = (e - s + a.K - 1) / a.K;
    lo = min(e, s + k * c);
    hi = min(e, lo + c);
    hbase = (size_t)fp * (a.lmask + 1) * a.K + k;
    hstride = a.K;
  }
}

__global__ __launch_bounds__(JB) void rj_count(PassArgs a, uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[MAX_PL];
  const uint32_t PL = a.lmask + 1;
  for (uint32_t i = threadIdx.x; i < PL; i += JB) cnt[i] = 0u;
  __syncthreads();
  uint32_t lo, hi, hstride;
  size_t hbase;
  block_range(a, lo, hi, hbase, hstride);
  uint32_t i = lo + threadIdx.x;
  for (; i + 3 * JB < hi; i += 4 * JB) {
    int32_t k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) k[u] = __builtin_nontemporal_load(a.in + i + u * JB);
#pragma unroll
    for (int u = 0; u < 4; ++u) atomicAdd(&cnt[local_id(a, k[u])], 1u);
  }
  for (; i < hi; i += JB) atomicAdd(&cnt[local_id(a, a.in[i])], 1u);
  __syncthreads();
  for (uint32_t l = threadIdx.x; l < PL; l += JB) hist[hbase + (size_t)l * hstride] = cnt[l];
}

// Moves the block's keys to their partitions (offs = scanned counts). Tiles of
// 4096 keys are sorted by local partition in LDS first (ranks from LDS
// atomics, a scan of the <= 128 counts), so each partition's keys of a tile
// leave as one contiguous run: consecutive lanes store consecutive addresses.
// Scattering key by key sent every lane of a store to a different partition
// and ran at a third of the copy rate.
__global__ __launch_bounds__(JB) void rj_scatter(PassArgs a, const uint32_t* __restrict__ offs) {
  __shared__ uint32_t off[MAX_PL], cnt[MAX_PL], start[MAX_PL];
  __shared__ int32_t stage[PTILE];
  const uint32_t PL = a.lmask + 1;
  uint32_t lo, hi, hstride;
  size_t hbase;
  block_range(a, lo, hi, hbase, hstride);
  for (uint32_t l = threadIdx.x; l < PL; l += JB) off[l] = offs[hbase + (size_t)l * hstride];
  for (uint32_t t0 = lo; t0 < hi; t0 += PTILE) {
    const uint32_t m = min((uint32_t)PTILE, hi - t0);
    int32_t key[PT];
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      const uint32_t i = threadIdx.x + j * JB;
      key[j] = i < m ? __builtin_nontemporal_load(a.in + t0 + i) : 0;
    }
    for (uint32_t l = threadIdx.x; l < PL; l += JB) cnt[l] = 0u;
    __syncthreads();
    uint32_t rank[PT];
#pragma unroll
    for (int j = 0; j < PT; ++j)
      rank[j] = threadIdx.x + j * JB < m ? atomicAdd(&cnt[local_id(a, key[j])], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the counts: two per lane
      const uint32_t l0 = 2 * threadIdx.x, l1 = l0 + 1;
      const uint32_t c0 = l0 < PL ? cnt[l0] : 0u, c1 = l1 < PL ? cnt[l1] : 0u;
      const uint32_t incl = wave_incl_scan(c0 + c1);
      if (l0 < PL) start[l0] = incl - c0 - c1;
      if (l1 < PL) start[l1] = incl - c1;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PT; ++j)
      if (threadIdx.x + j * JB < m) stage[start[local_id(a, key[j])] + rank[j]] = key[j];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += JB) {
      const int32_t k = stage[i];
      const uint32_t l = local_id(a, k);
      a.out[off[l] + (i - start[l])] = k;
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < PL; l += JB) off[l] += cnt[l];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ LDS table
// A slot is one 64-bit word: count << 32 | key, so a probe step is a single
// LDS read. The empty slot holds key INT32_MIN with count 0.
constexpr uint64_t SLOT_EMPTY = (uint64_t)(uint32_t)EMPTY;
constexpr int JT = 512;  // threads per workgroup of the join kernels

struct LdsTable {
  uint64_t* slot;

  __device__ void clear() {
    for (uint32_t i = threadIdx.x; i < LDS_SLOTS; i += blockDim.x) slot[i] = SLOT_EMPTY;
  }
  // At most LDS_KEYS distinct keys per fill, so a free slot always exists.
  __device__ void insert(int32_t k) {
    const uint64_t mine = (1ull << 32) | (uint32_t)k;
    uint32_t h = mix32(k) & (LDS_SLOTS - 1);
    while (true) {
      const uint64_t prev = atomicCAS((unsigned long long*)&slot[h], (unsigned long long)SLOT_EMPTY,
                                      (unsigned long long)mine);
      if (prev == SLOT_EMPTY) return;
      if ((uint32_t)prev == (uint32_t)k) {
        atomicAdd((unsigned long long*)&slot[h], 1ull << 32);
        return;
      }
      h = (h + 1) & (LDS_SLOTS - 1);
    }
  }
  __device__ uint32_t lookup_from(int32_t k, uint32_t h, uint64_t first) const {
    uint64_t v = first;
    while (true) {
      if ((uint32_t)v == (uint32_t)k) return (uint32_t)(v >> 32);
      if (v == SLOT_EMPTY) return 0u;
      h = (h + 1) & (LDS_SLOTS - 1);
      v = slot[h];
    }
  }
  // Four lookups with their first probes issued together.
  __device__ uint32_t lookup4(int4 k) const {
    const uint32_t h0 = mix32(k.x) & (LDS_SLOTS - 1), h1 = mix32(k.y) & (LDS_SLOTS - 1),
                   h2 = mix32(k.z) & (LDS_SLOTS - 1), h3 = mix32(k.w) & (LDS_SLOTS - 1);
    const uint64_t v0 = slot[h0], v1 = slot[h1], v2 = slot[h2], v3 = slot[h3];
    return lookup_from(k.x, h0, v0) + lookup_from(k.y, h1, v1) + lookup_from(k.z, h2, v2) +
           lookup_from(k.w, h3, v3);
  }
  __device__ uint32_t lookup(int32_t k) const {
    const uint32_t h = mix32(k) & (LDS_SLOTS - 1);
    return lookup_from(k, h, slot[h]);
  }
};

__device__ __forceinline__ uint64_t probe_range(const LdsTable& t, const int32_t* __restrict__ s,
                                                uint32_t lo, uint32_t hi, uint32_t tid,
                                                uint32_t nthr) {
  uint64_t c = 0;
  // head up to a multiple of 4, then dwordx4 groups, then the tail
  const uint32_t a = min(hi, (lo + 3) & ~3u);
  for (uint32_t r = lo + tid; r < a; r += nthr) c += t.lookup(s[r]);
  const uint32_t g1 = hi / 4;
  uint32_t g = a / 4 + tid;
  for (; g + nthr < g1; g += 2 * nthr) {
    const int4 v = ld4nt(s, g), w = ld4nt(s, g + nthr);
    c += t.lookup4(v) + t.lookup4(w);
  }
  for (; g < g1; g += nthr) c += t.lookup4(ld4nt(s, g));
  for (uint32_t r = max(a, g1 * 4) + tid; r < hi; r += nthr) c += t.lookup(s[r]);
  return c;
}

// One workgroup per partition p: build side keys [rb, re) and probe side keys
// [sb, se), from the scanned partition offsets (k-major, G blocks).
__global__ __launch_bounds__(JT) void rj_join(const int32_t* __restrict__ r,
                                              const uint32_t* __restrict__ roffs, uint32_t rG,
                                              const uint32_t* __restrict__ rtotal,
                                              const int32_t* __restrict__ s,
                                              const uint32_t* __restrict__ soffs, uint32_t sG,
                                              const uint32_t* __restrict__ stotal, uint32_t P,
                                              int64_t* __restrict__ partials) {
  __shared__ uint64_t tslot[LDS_SLOTS];
  __shared__ int64_t red[JT / 64];
  const uint32_t p = blockIdx.x;
  const uint32_t rb = roffs[(size_t)p * rG], re = p + 1 < P ? roffs[(size_t)(p + 1) * rG] : *rtotal;
  const uint32_t sb = soffs[(size_t)p * sG], se = p + 1 < P ? soffs[(size_t)(p + 1) * sG] : *stotal;
  LdsTable t{tslot};
  uint64_t c = 0;
  if (re > rb && se > sb) {
    for (uint32_t fb = rb; fb < re; fb += LDS_KEYS) {  // table-sized pieces of the partition
      const uint32_t fe = min(re, fb + LDS_KEYS);
      t.clear();
      __syncthreads();
      for (uint32_t i = fb + threadIdx.x; i < fe; i += JT) t.insert(r[i]);
      __syncthreads();
      c += probe_range(t, s, sb, se, threadIdx.x, JT);
      __syncthreads();
    }
  }
  const int64_t bs = block_sum((int64_t)c, red);
  if (threadIdx.x == 0) partials[p] = bs;
}

// Every workgroup builds the whole (small) build side and probes a slice.
__global__ __launch_bounds__(JT) void rj_broadcast(const int32_t* __restrict__ r, uint32_t nr,
                                                   const int32_t* __restrict__ s, uint32_t ns,
                                                   int64_t* __restrict__ partials) {
  __shared__ uint64_t tslot[LDS_SLOTS];
  __shared__ int64_t red[JT / 64];
  LdsTable t{tslot};
  t.clear();
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nr; i += JT) t.insert(r[i]);
  __syncthreads();
  // contiguous slice of the probe keys per workgroup, a multiple of 4 long
  const uint32_t per = ((ns + gridDim.x - 1) / gridDim.x + 3) & ~3u;
  const uint32_t lo = min(ns, blockIdx.x * per), hi = min(ns, lo + per);
  const uint64_t c = probe_range(t, s, lo, hi, threadIdx.x, JT);
  const int64_t bs = block_sum((int64_t)c, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// ---------------------------------------------------------------- global table
// Keys below the capacity map to themselves and higher bits are folded in, so
// keys arriving in (near) key order touch neighbouring slots (the reference
// hashes by key modulo size, kernels/microbenchmarks.metal:240).
__device__ __forceinline__ uint32_t jhash(int32_t key, uint32_t shift) {
  const uint32_t k = (uint32_t)key;
  const uint32_t lg = 32 - shift;
  return (k ^ (k >> lg) ^ (k >> (2 * lg < 32 ? 2 * lg : 31))) & ((1u << lg) - 1);
}

__global__ __launch_bounds__(256) void join_clear(int32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ mult, uint32_t cap) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += stride) {
    keys[i] = EMPTY;
    mult[i] = 0u;
  }
}

__global__ __launch_bounds__(256) void join_build(const int32_t* __restrict__ in, uint32_t n,
                                                  int32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ mult, uint32_t cap,
                                                  uint32_t shift) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t k = in[i];
    uint32_t h = jhash(k, shift);
    for (uint32_t p = 0; p < cap; ++p) {
      const int32_t prev = atomicCAS(&keys[h], EMPTY, k);
      if (prev == EMPTY || prev == k) {
        atomicAdd(&mult[h], 1u);
        break;
      }
      h = (h + 1) & (cap - 1);
    }
  }
}

__device__ __forceinline__ uint32_t join_lookup(int32_t k, const int32_t* __restrict__ keys,
                                                const uint32_t* __restrict__ mult, uint32_t cap,
                                                uint32_t shift) {
  uint32_t h = jhash(k, shift);
  for (uint32_t p = 0; p < cap; ++p) {
    const int32_t t = keys[h];
    if (t == k) return mult[h];
    if (t == EMPTY) return 0u;
    h = (h + 1) & (cap - 1);
  }
  return 0u;
}

__global__ __launch_bounds__(256) void join_probe(const int32_t* __restrict__ probe, uint32_t n,
                                                  const int32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ mult, uint32_t cap,
                                                  uint32_t shift, int64_t* __restrict__ partials) {
  __shared__ int64_t red[4];
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nvec; g += nthr) {
    const int4 v = ld4(probe, g);
    c += join_lookup(v.x, keys, mult, cap, shift) + join_lookup(v.y, keys, mult, cap, shift) +
         join_lookup(v.z, keys, mult, cap, shift) + join_lookup(v.w, keys, mult, cap, shift);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3))
    c += join_lookup(probe[nvec * 4 + threadIdx.x], keys, mult, cap, shift);
  const int64_t s = block_sum((int64_t)c, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ---------------------------------------------------------------- direct
constexpr uint64_t DIRECT_MAX_SPAN = 1ull << 31;

// bits[k - lo] for every build key, with no returning atomics (no lane waits on
// one). A lane ORs its keys' bits per word; every run but its last is stored at
// once. The lanes' last runs are then combined across the wave: a doubling OR over
// neighbouring lanes with the same word (sorted keys give long runs), after which
// only the first lane of each run stores. Whether a key value repeats is read
// afterwards from the bitmap's population count (jd_popc).
__global__ __launch_bounds__(256) void jd_bits(const int32_t* __restrict__ in, uint32_t n,
                                               int32_t lo, uint32_t* __restrict__ bits) {
  const uint32_t nq = (n + 3) / 4, stride = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; gw < nq; gw += stride) {
    const uint32_t g = gw + lane;
    const int4 v = g < nq ? ld4nt(in, g) : make_int4(0, 0, 0, 0);
    const int32_t k[4] = {v.x, v.y, v.z, v.w};
    uint32_t word = 0xFFFFFFFFu, mask = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) break;
      const uint32_t off = (uint32_t)((int64_t)k[j] - lo);
      const uint32_t w = off >> 5, m = 1u << (off & 31);
      if (mask && w != word) {
        atomicOr(&bits[word], mask);
        mask = 0;
      }
      word = w;
      mask |= m;
    }
    const uint32_t mine = word;  // every lane takes part in the shuffles
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t w2 = (uint32_t)__shfl_down((int)mine, d, 64);
      const uint32_t m2 = (uint32_t)__shfl_down((int)mask, d, 64);
      if (lane + d < 64 && w2 == mine) mask |= m2;
    }
    const uint32_t before = (uint32_t)__shfl_up((int)mine, 1, 64);
    if (mask && (lane == 0 || before != mine)) atomicOr(&bits[mine], mask);
  }
}

// Population count of the bitmap: workgroup sums added into JD_POPC_SLOTS
// counters 256 bytes apart (a few workgroups per counter). The total is below
// the number of build keys exactly when a key value repeats.
constexpr int JD_POPC_SLOTS = 64, JD_POPC_STRIDE = 32;  // int64 per slot
__global__ __launch_bounds__(256) void jd_popc(const uint32_t* __restrict__ bits, uint64_t words,
                                               int64_t* __restrict__ sums) {
  __shared__ int64_t red[4];
  int64_t c = 0;
  for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += __popc(bits[i]);
  const int64_t s = block_sum(c, red);
  if (threadIdx.x == 0 && s)
    atomicAdd((unsigned long long*)&sums[(blockIdx.x % JD_POPC_SLOTS) * JD_POPC_STRIDE],
              (unsigned long long)s);
}

// counts[k - lo] += 1 for every build key (the repeated-keys form).
__global__ __launch_bounds__(256) void jd_counts(const int32_t* __restrict__ in, uint32_t n,
                                                 int32_t lo, uint32_t* __restrict__ counts) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    atomicAdd(&counts[(uint32_t)((int64_t)in[i] - lo)], 1u);
}

// Matches of the probe keys: bit (COUNTS = false) or count (COUNTS = true) of
// each key inside [lo, lo + span). Two vector groups of keys in flight per lane.
template <bool COUNTS>
__global__ __launch_bounds__(256) void jd_probe(const int32_t* __restrict__ probe, uint32_t n,
                                                int32_t lo, uint64_t span,
                                                const uint32_t* __restrict__ table,
                                                int64_t* __restrict__ partials) {
  __shared__ int64_t red[4];
  const uint32_t nq = (n + 3) / 4, stride = gridDim.x * blockDim.x;
  uint64_t c = 0;
  // QB vector groups per lane per trip: their keys load together, then all 4*QB
  // table reads are issued (branch-free: a key outside the range reads entry 0
  // and is masked) before any is summed, so the reads' latencies overlap.
  constexpr int QB = 4;
  for (uint32_t g0 = blockIdx.x * blockDim.x + threadIdx.x; g0 < nq; g0 += QB * stride) {
    int4 v[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const uint32_t g = g0 + (uint32_t)b * stride;
      v[b] = g < nq ? ld4nt(probe, g) : make_int4(0, 0, 0, 0);
    }
    uint32_t t[4 * QB];
    uint32_t ok = 0, sh[4 * QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const uint32_t g = g0 + (uint32_t)b * stride;
      const int32_t ks[4] = {v[b].x, v[b].y, v[b].z, v[b].w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t off = (uint64_t)((int64_t)ks[j] - lo);
        const bool in = g < nq && 4 * g + j < n && off < span;
        const uint64_t o = in ? off : 0;
        ok |= (uint32_t)in << (4 * b + j);
        sh[4 * b + j] = (uint32_t)(o & 31);
        t[4 * b + j] = COUNTS ? table[o] : table[o >> 5];
      }
    }
#pragma unroll
    for (int i = 0; i < 4 * QB; ++i)
      if ((ok >> i) & 1) c += COUNTS ? t[i] : (t[i] >> sh[i]) & 1u;
  }
  const int64_t s = block_sum((int64_t)c, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Partition bits for a build side of nb keys: partitions average at most half
// an LDS table fill.
int derived_bits(uint32_t nb) {
  int b = 1;
  while (b < MAX_BITS && ((uint64_t)nb >> b) > LDS_KEYS / 2) ++b;
  return b;
}

// Fewest bits a tuning entry may ask for: every table-sized piece of a
// partition probes the whole probe partition again, so partitions average at
// most four pieces.
int min_bits(uint32_t nb) {
  int b = 1;
  while (b < MAX_BITS && ((uint64_t)nb >> b) > 4 * LDS_KEYS) ++b;
  return b;
}

// First pass (or the only one): contiguous chunks over the input.
PassArgs first_pass(const gdb_ctx* ctx, const int32_t* in, int32_t* out, uint32_t n, int b1,
                    uint32_t* G) {
  uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(ctx->info.cu_count * 2, (n + PTILE - 1) / PTILE));
  uint32_t chunk = std::max<uint32_t>(1, (n + g - 1) / g);
  g = std::max<uint32_t>(1, (n + chunk - 1) / chunk);
  *G = g;
  PassArgs a{};
  a.in = in;
  a.out = out;
  a.n = n;
  a.chunk = chunk;
  a.shift = 32 - b1;
  a.lmask = (1u << b1) - 1;
  a.mode = 0;
  return a;
}

// Second pass over the first pass's output: final partition bits `bits`, of
// which the first b1 were the first pass's.
PassArgs second_pass(const int32_t* in, int32_t* out, int b1, int bits, const uint32_t* offs1,
                     uint32_t G1, const uint32_t* total1, uint32_t K) {
  PassArgs a{};
  a.in = in;
  a.out = out;
  a.shift = 32 - bits;
  a.lmask = (1u << (bits - b1)) - 1;
  a.mode = 1;
  a.part_offs = offs1;
  a.part_G = G1;
  a.part_total = total1;
  a.K = K;
  return a;
}

// count, exclusive scan of the counts (in place), scatter
int run_pass(gdb_ctx* ctx, const PassArgs& a, uint32_t blocks, uint32_t nhist, uint32_t* hist,
             uint32_t* total, uint32_t* tsum, hipStream_t st) {
  rj_count<<<blocks, JB, 0, st>>>(a, hist);
  GDB_HIP(hipGetLastError());
  GDB_TRY(exclusive_scan_u32(ctx, hist, nhist, hist, total, tsum, st));
  rj_scatter<<<blocks, JB, 0, st>>>(a, hist);
  GDB_HIP(hipGetLastError());
  return GDB_OK;
}

}  // namespace

namespace gdb {
int join_derived_bits(uint64_t nb) { return derived_bits((uint32_t)std::min<uint64_t>(nb, UINT32_MAX)); }
}  // namespace gdb

extern "C" int gdb_join_count(gdb_ctx* ctx, const gdb_table* build, const char* build_col,
                              const gdb_table* probe, const char* probe_col, uint64_t* matches,
                              gdb_timing* timing) {
  if (!ctx || !build || !probe || !matches) return fail(GDB_E_INVALID, "null argument");
  const Column *cb, *cp;
  GDB_TRY(column(build, build_col, GDB_I32, &cb));
  GDB_TRY(column(probe, probe_col, GDB_I32, &cp));
  GDB_TRY(require_resident(build, "join build table"));
  GDB_TRY(require_resident(probe, "join probe table"));
  if (build->nrows && cb->minv == INT32_MIN)
    return fail(GDB_E_UNSUPPORTED, "join keys must not be INT32_MIN (the empty-slot marker)");
  if (build->nrows >= (1ull << 31) || probe->nrows >= (1ull << 31))
    return fail(GDB_E_UNSUPPORTED, "join sides are limited to 2^31-1 rows");
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t nb = (uint32_t)build->nrows, np = (uint32_t)probe->nrows;
  const int32_t* rk = (const int32_t*)cb->dev;
  const int32_t* sk = (const int32_t*)cp->dev;
  const int tuned = ctx->tune[T_JOIN].radix_bits;
  const uint64_t span = nb ? (uint64_t)(cb->maxv - cb->minv) + 1 : 0;
  // the filter's memory stays within ~256 B per build row (count form) plus 4 MiB
  const bool direct_ok = nb > 0 && span <= std::min<uint64_t>(DIRECT_MAX_SPAN, 64ull * nb + (1u << 20));
  // direct: asked for (-2) when the range allows, or derived when it is dense
  const bool direct = direct_ok && (tuned == -2 || (tuned == 0 && span <= 8ull * nb));
  const bool global = !direct && tuned == -1;
  const bool broadcast = !direct && !global && tuned <= 0 && nb <= LDS_KEYS;
  const int bits = direct || global || broadcast ? 0
                   : tuned > 0 ? std::max(min_bits(nb), std::min(tuned, MAX_BITS))
                               : derived_bits(nb);
  hipStream_t st = ctx->stream;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_OUT, 64, (void**)&dout));
  GDB_TRY(pinned(ctx, 64 * 8, (void**)&host));  // result, or the bitmap's popcount partials
  int64_t* partials;
  uint32_t nparts = 0;

  if (direct) {
    const int32_t lo = (int32_t)cb->minv;
    const uint64_t words = (span + 31) / 32;
    uint32_t* table;
    const uint64_t popc_at = (words + 63) & ~uint64_t(63);  // 256-byte aligned
    const size_t popc_bytes = (size_t)JD_POPC_SLOTS * JD_POPC_STRIDE * 8;
    GDB_TRY(arena(ctx, A_HT, popc_at * 4 + popc_bytes, (void**)&table));
    int64_t* popc = (int64_t*)(table + popc_at);
    int blocks = resident_blocks(ctx, (const void*)jd_probe<false>, 256, 0,
                                 ctx->tune[T_JOIN].blocks_per_cu);
    blocks = std::max(1, std::min<int>(blocks, (int)((np / 4 + 255) / 256)));
    nparts = (uint32_t)blocks;
    GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * nparts, (void**)&partials));
    const int fill = ctx->info.cu_count * 4;
    GDB_HIP(hipEventRecord(ctx->ev[0], st));
    GDB_TRY(gdb::fill(table, 0, popc_at * 4 + popc_bytes, st));
    const uint64_t bq = (nb + 3) / 4;
    jd_bits<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>((bq + 255) / 256, 1u << 16)), 256,
              0, st>>>(rk, nb, lo, table);
    jd_popc<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>((words + 1023) / 1024, 1024)), 256,
              0, st>>>(table, words, popc);
    GDB_HIP(hipGetLastError());
    int64_t* hpop = (int64_t*)host;
    GDB_HIP(hipMemcpy2DAsync(hpop, 8, popc, JD_POPC_STRIDE * 8, 8, JD_POPC_SLOTS,
                             hipMemcpyDeviceToHost, st));
    GDB_HIP(hipStreamSynchronize(st));
    int64_t distinct = 0;
    for (int i = 0; i < JD_POPC_SLOTS; ++i) distinct += hpop[i];
    const bool counts = (uint64_t)distinct != nb;  // a key value repeats
    uint32_t* cnt = table;
    if (counts) {  // repeated build keys: a count per key value
      GDB_TRY(arena(ctx, A_JOINPART, span * 4, (void**)&cnt));
      GDB_TRY(gdb::fill(cnt, 0, span * 4, st));
      jd_counts<<<fill, 256, 0, st>>>(rk, nb, lo, cnt);
      GDB_HIP(hipGetLastError());
    }
    GDB_HIP(kernel_mark(ctx, 1, st));
    if (counts)
      jd_probe<true><<<blocks, 256, 0, st>>>(sk, np, lo, span, cnt, partials);
    else
      jd_probe<false><<<blocks, 256, 0, st>>>(sk, np, lo, span, table, partials);
    GDB_HIP(kernel_mark(ctx, 2, st));
  } else if (global) {
    uint32_t cap = 1024, lg = 10;
    while (cap < 2ull * nb) {
      cap <<= 1;
      ++lg;
    }
    const uint32_t shift = 32 - lg;
    char* h;
    GDB_TRY(arena(ctx, A_HT, (size_t)cap * 8, (void**)&h));
    int32_t* keys = (int32_t*)h;
    uint32_t* mult = (uint32_t*)(h + (size_t)cap * 4);
    int blocks = resident_blocks(ctx, (const void*)join_probe, 256, 0, ctx->tune[T_JOIN].blocks_per_cu);
    blocks = std::max(1, std::min<int>(blocks, (int)((np / 4 + 255) / 256)));
    nparts = (uint32_t)blocks;
    GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * nparts, (void**)&partials));
    const int fill = ctx->info.cu_count * 4;
    GDB_HIP(hipEventRecord(ctx->ev[0], st));
    join_clear<<<fill, 256, 0, st>>>(keys, mult, cap);
    if (nb) join_build<<<fill, 256, 0, st>>>(rk, nb, keys, mult, cap, shift);
    GDB_HIP(kernel_mark(ctx, 1, st));
    join_probe<<<blocks, 256, 0, st>>>(sk, np, keys, mult, cap, shift, partials);
    GDB_HIP(kernel_mark(ctx, 2, st));
  } else if (broadcast) {
    const int per_cu = std::max(1, ctx->info.lds_bytes_per_cu / (int)(LDS_SLOTS * 8 + 64));
    nparts = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)per_cu * ctx->info.cu_count, ((uint64_t)np + 8191) / 8192));
    GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * nparts, (void**)&partials));
    GDB_HIP(hipEventRecord(ctx->ev[0], st));
    GDB_HIP(kernel_mark(ctx, 1, st));
    rj_broadcast<<<nparts, JT, 0, st>>>(rk, nb, sk, np, partials);
    GDB_HIP(kernel_mark(ctx, 2, st));
  } else {
    const uint32_t P = 1u << bits;
    const int b1 = bits > 7 ? bits / 2 : bits;  // first pass (or the only one)
    const bool two = b1 != bits;
    const uint32_t P1 = 1u << b1;
    // second-pass slices per first-level partition: about two workgroups per CU
    const uint32_t K = std::max<uint32_t>(1, (uint32_t)ctx->info.cu_count * 2 / P1);
    auto padded = [](uint32_t n) { return ((size_t)n + 64 + 63) & ~size_t(63); };
    const size_t rn = padded(nb), sn = padded(np);
    int32_t* kbuf;
    GDB_TRY(arena(ctx, A_JOINPART, (rn + sn) * 4 * (two ? 2 : 1), (void**)&kbuf));
    int32_t* r_a = kbuf;
    int32_t* r_b = two ? r_a + rn : r_a;
    int32_t* s_a = r_b + rn;
    int32_t* s_b = two ? s_a + sn : s_a;
    uint32_t rG1 = 1, sG1 = 1;
    PassArgs r1 = first_pass(ctx, rk, r_a, nb, b1, &rG1);
    PassArgs s1 = first_pass(ctx, sk, s_a, np, b1, &sG1);
    // final offsets per side: P x K (two passes) or P x G1 (one pass)
    const uint32_t rG = two ? K : rG1, sG = two ? K : sG1;
    const size_t o_r = (size_t)P * rG, o_s = (size_t)P * sG;
    const size_t o_1 = two ? std::max<size_t>((size_t)P1 * rG1, (size_t)P1 * sG1) : 0;
    const size_t ts = scan_tile_sums_needed((uint32_t)std::max({o_r, o_s, o_1}));
    uint32_t* obuf;
    GDB_TRY(arena(ctx, A_JOINHIST, (o_r + o_s + o_1 + ts + 16) * 4, (void**)&obuf));
    uint32_t* roffs = obuf;
    uint32_t* soffs = roffs + o_r;
    uint32_t* offs1 = soffs + o_s;
    uint32_t* tsum = offs1 + o_1;
    uint32_t* totals = tsum + ts;  // [0] build, [1] probe, [2] first-pass scratch
    nparts = P;
    GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * nparts, (void**)&partials));
    GDB_HIP(hipEventRecord(ctx->ev[0], st));
    if (two) {
      GDB_TRY(run_pass(ctx, r1, rG1, (uint32_t)(P1 * rG1), offs1, totals + 2, tsum, st));
      const PassArgs r2 = second_pass(r_a, r_b, b1, bits, offs1, rG1, totals + 2, K);
      GDB_TRY(run_pass(ctx, r2, P1 * K, (uint32_t)o_r, roffs, totals + 0, tsum, st));
      GDB_TRY(run_pass(ctx, s1, sG1, (uint32_t)(P1 * sG1), offs1, totals + 2, tsum, st));
      const PassArgs s2 = second_pass(s_a, s_b, b1, bits, offs1, sG1, totals + 2, K);
      GDB_TRY(run_pass(ctx, s2, P1 * K, (uint32_t)o_s, soffs, totals + 1, tsum, st));
    } else {
      r1.out = r_b;
      s1.out = s_b;
      GDB_TRY(run_pass(ctx, r1, rG1, (uint32_t)o_r, roffs, totals + 0, tsum, st));
      GDB_TRY(run_pass(ctx, s1, sG1, (uint32_t)o_s, soffs, totals + 1, tsum, st));
    }
    GDB_HIP(kernel_mark(ctx, 1, st));
    rj_join<<<P, JT, 0, st>>>(r_b, roffs, rG, totals + 0, s_b, soffs, sG, totals + 1, P,
                              partials);
    GDB_HIP(kernel_mark(ctx, 2, st));
  }
  GDB_HIP(hipGetLastError());
  sum_columns<<<1, 256, 0, st>>>(partials, nparts, dout);
  GDB_HIP(hipMemcpyAsync(host, dout, 8, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  // algorithmic bytes: each key read once; the main kernel of the radix path
  // (the per-partition join) reads both partitioned sides
  const uint64_t main_bytes = (!global && !broadcast && !direct) ? 4ull * nb + 4ull * np : 4ull * np;
  GDB_TRY(timing_finish(ctx, timing, 4ull * nb + 4ull * np, main_bytes));
  *matches = (uint64_t)host[0];
  return GDB_OK;
}
