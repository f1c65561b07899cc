// TPC-H pipelines over partsupp: Q2, Q11, Q16, Q20.
//
// partsupp is keyed by (ps_partkey, ps_suppkey). Q2 and Q11 aggregate per part
// into direct-address maps; Q16's COUNT(DISTINCT ps_suppkey) per group and
// Q20's per-pair quantity sums go through an open-addressing hash table in HBM
// keyed by a packed 64-bit pair (linear probing, CAS insert, sized to twice
// the candidate count). References: src/qN.cpp and kernels/qN.metal, cited
// per pipeline.
#include <cstring>
#include <vector>

#include "tpch_common.h"

using namespace gdb;
using namespace gdbk;
using namespace gdbt;

namespace {

constexpr uint8_t NONE = 255;
constexpr unsigned long long EMPTY = ~0ull;

__device__ __forceinline__ uint64_t mix(uint64_t z) {  // SplitMix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Inserts key; returns its slot and whether this call inserted it.
__device__ __forceinline__ uint64_t ht_insert(unsigned long long* keys, uint64_t mask,
                                              unsigned long long key, bool* fresh) {
  uint64_t slot = mix(key) & mask;
  for (;;) {
    const unsigned long long prev = atomicCAS(&keys[slot], EMPTY, key);
    if (prev == EMPTY || prev == key) {
      *fresh = prev == EMPTY;
      return slot;
    }
    slot = (slot + 1) & mask;
  }
}

// Slot of key, or -1 if absent.
__device__ __forceinline__ int64_t ht_find(const unsigned long long* keys, uint64_t mask,
                                           unsigned long long key) {
  uint64_t slot = mix(key) & mask;
  for (;;) {
    const unsigned long long k = keys[slot];
    if (k == key) return (int64_t)slot;
    if (k == EMPTY) return -1;
    slot = (slot + 1) & mask;
  }
}

__device__ __forceinline__ unsigned long long pair_key(uint32_t a, uint32_t b) {
  return ((unsigned long long)a << 32) | b;
}

// Order-preserving int64 -> uint64.
__device__ __forceinline__ unsigned long long ord64(int64_t v) {
  return (unsigned long long)v ^ 0x8000000000000000ull;
}

// ---- Q2 (reference kernels/q2.metal:33-124, src/q2.cpp) ------------------------

// psel[partkey] = 0x100 | p_mfgr for parts of size `size` whose type is in mask.
__global__ __launch_bounds__(256) void q2_part(const int32_t* __restrict__ pk,
                                               const int32_t* __restrict__ size,
                                               const uint8_t* __restrict__ type,
                                               const uint8_t* __restrict__ mfgr, uint32_t n,
                                               int32_t want, Mask4 mask,
                                               uint16_t* __restrict__ psel) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (size[i] == want && mask.has(type[i])) psel[pk[i]] = (uint16_t)(0x100u | mfgr[i]);
}

// snat[suppkey] = nation when the nation lies in `region` (else NONE); sbal = s_acctbal.
__global__ __launch_bounds__(256) void q2_supp(const int32_t* __restrict__ sk,
                                               const int32_t* __restrict__ sn,
                                               const int64_t* __restrict__ sb, uint32_t n,
                                               const uint8_t* __restrict__ nreg, int32_t region,
                                               uint8_t* __restrict__ snat,
                                               int64_t* __restrict__ sbal) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t nat = sn[i];
    const bool ok = nat >= 0 && nat < 255 && nreg[nat] == region;
    snat[sk[i]] = ok ? (uint8_t)nat : NONE;
    sbal[sk[i]] = sb[i];
  }
}

__global__ __launch_bounds__(256) void q2_min(const int32_t* __restrict__ pk,
                                              const int32_t* __restrict__ sk,
                                              const int64_t* __restrict__ cost, uint32_t n,
                                              const uint16_t* __restrict__ psel,
                                              const uint8_t* __restrict__ snat,
                                              unsigned long long* __restrict__ mn) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t p = pk[i];
    if (psel[p] && snat[sk[i]] != NONE) atomicMin(&mn[p], ord64(cost[i]));
  }
}

struct Q2Src {  // (s_acctbal desc, nation name rank asc, suppkey asc, partkey asc)
  const int32_t *pk, *sk;
  const int64_t* cost;
  const uint16_t* psel;
  const uint8_t* snat;
  const int64_t* sbal;
  const unsigned long long* mn;
  const uint8_t* rank;  // nation key -> position of its name in name order
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    const int32_t p = pk[i], s = sk[i];
    const uint16_t sel = psel[p];
    const uint8_t nat = snat[s];  // rank has 256 entries, NONE included
    e = TopEntry{sbal[s], (int32_t)(((uint32_t)rank[nat] << 26) | (uint32_t)s), p,
                 (int32_t)nat | ((int32_t)(sel & 0xFF) << 8), 0};
    return sel && nat != NONE && ord64(cost[i]) == mn[p];
  }
};

// ---- Q11 (reference kernels/q11.metal:27-59, src/q11.cpp) ----------------------

__global__ __launch_bounds__(256) void supp_nation_bits(const int32_t* __restrict__ sk,
                                                        const int32_t* __restrict__ sn,
                                                        uint32_t n, int32_t nation,
                                                        uint32_t* __restrict__ bits) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (sn[i] == nation) bit_set(bits, (uint32_t)sk[i]);
}

// val[partkey] += supplycost * availqty over the nation's suppliers; runs of one
// part within a lane's group of 4 rows are summed before the atomic (partsupp
// is stored part by part). total accumulates the same values.
__global__ __launch_bounds__(256) void q11_scan(const int32_t* __restrict__ pk,
                                                const int32_t* __restrict__ sk,
                                                const int32_t* __restrict__ avail,
                                                const int64_t* __restrict__ cost, uint32_t n,
                                                const uint32_t* __restrict__ sbits,
                                                int64_t* __restrict__ val,
                                                uint8_t* __restrict__ seen,
                                                int64_t* __restrict__ total) {
  total = acc_copy(total);  // this workgroup's copy of the result
  __shared__ int64_t red[4];
  int64_t tot = 0;
  walk2(quads(n), [&](uint32_t g) { return ld4nt(sk, g); }, [&](uint32_t g, const int4& s) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      m |= (uint32_t)((4 * g + j < n) && bit_test(sbits, (uint32_t)lane_of(s, j))) << j;
    if (!m) return;
    const int4 p = ld4nt(pk, g), a = ld4nt(avail, g);
    const longlong2 c0 = ld2nt(cost, 2 * g), c1 = ld2nt(cost, 2 * g + 1);
    int key = -1;
    int64_t acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const int kj = lane_of(p, j);
      const int64_t v = lane_of(c0, c1, j) * (int64_t)lane_of(a, j);
      tot += v;
      if (kj != key) {
        if (key >= 0) add_i64(&val[key], acc);
        key = kj;
        acc = 0;
        seen[kj] = 1;
      }
      acc += v;
    }
    if (key >= 0) add_i64(&val[key], acc);
  });
  block_add(tot, red, total);
}

struct KV {
  int32_t key;
  int32_t pad;
  int64_t val;
};

__global__ __launch_bounds__(256) void q11_collect(const int64_t* __restrict__ val,
                                                   const uint8_t* __restrict__ seen, uint32_t n,
                                                   const int64_t* __restrict__ total, int64_t num,
                                                   int64_t den, KV* __restrict__ out,
                                                   unsigned long long* __restrict__ cnt) {
  const __int128 t = (__int128)(*total) * num;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (seen[i] && (__int128)val[i] * den > t) {
      const unsigned long long pos = atomicAdd(cnt, 1ull);
      out[pos] = KV{(int32_t)i, 0, val[i]};
    }
}

// ---- Q16 (reference kernels/q16.metal:12-66, src/q16.cpp) ----------------------

struct Word {
  char c[32];
  int len;
};

// First position >= from where w occurs in s[0, n), or -1.
__device__ __forceinline__ int find_word(const uint8_t* s, int n, const Word& w, int from) {
  for (int i = from; i + w.len <= n; ++i) {
    int j = 0;
    while (j < w.len && s[i + j] == (uint8_t)w.c[j]) ++j;
    if (j == w.len) return i;
  }
  return -1;
}

// bad[suppkey] = s_comment LIKE '%w1%w2%' (the field ends at its first NUL).
__global__ __launch_bounds__(256) void q16_supp(const int32_t* __restrict__ sk,
                                                const uint8_t* __restrict__ com, int width,
                                                uint32_t n, Word w1, Word w2,
                                                uint8_t* __restrict__ bad) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint8_t* s = com + (size_t)i * width;
    int len = 0;
    while (len < width && s[len]) ++len;
    const int a = find_word(s, len, w1, 0);
    bad[sk[i]] = (a >= 0 && find_word(s, len, w2, a + w1.len) >= 0) ? 1 : 0;
  }
}

struct Q16Part {
  int32_t brand, nsizes;
  int32_t sizes[8];
  Mask4 types;
};

// gid[partkey] = (brand * 256 + type) * 8 + size position for qualifying parts.
__global__ __launch_bounds__(256) void q16_part(const int32_t* __restrict__ pk,
                                                const uint8_t* __restrict__ brand,
                                                const uint8_t* __restrict__ type,
                                                const int32_t* __restrict__ size, uint32_t n,
                                                Q16Part q, int32_t* __restrict__ gid) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int pos = -1;
    for (int k = 0; k < q.nsizes; ++k)
      if (size[i] == q.sizes[k]) pos = k;
    if (pos >= 0 && brand[i] != q.brand && !q.types.has(type[i]))
      gid[pk[i]] = ((int32_t)brand[i] * 256 + type[i]) * 8 + pos;
  }
}

// Distinct (group, suppkey) pairs: a pair counts for its group the first time
// it is inserted into the hash set.
__global__ __launch_bounds__(256) void q16_scan(const int32_t* __restrict__ pk,
                                                const int32_t* __restrict__ sk, uint32_t n,
                                                const int32_t* __restrict__ gid,
                                                const uint8_t* __restrict__ bad,
                                                unsigned long long* __restrict__ keys,
                                                uint64_t mask, uint32_t* __restrict__ cnt) {
  struct Ld {
    int4 p, s;
  };
  walk2(quads(n), [&](uint32_t g) { return Ld{ld4nt(pk, g), ld4nt(sk, g)}; },
        [&](uint32_t g, const Ld& ld) {
    const int4 p = ld.p, s = ld.s;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) continue;
      const int32_t grp = gid[lane_of(p, j)];
      const uint32_t sv = (uint32_t)lane_of(s, j);
      if (grp < 0 || bad[sv]) continue;
      bool fresh;
      ht_insert(keys, mask, pair_key((uint32_t)grp, sv), &fresh);
      if (fresh) atomicAdd(&cnt[grp], 1u);
    }
  });
}

// The result's group order (brand, type name, size) as an ordinal: brand << 11 |
// type rank << 3 | size rank. `ord` holds, in device memory, the type code of
// each type rank, the size position of each size rank and the sizes.
struct Q16Ord {
  uint8_t tcode[256];
  uint8_t spos[8];
  int32_t sizes[8];
  int32_t nsizes;
};
constexpr uint32_t Q16_ORDINALS = 1u << 19;

__device__ __forceinline__ uint32_t q16_count_at(const uint32_t* __restrict__ cnt, uint32_t i,
                                                 const Q16Ord* __restrict__ o, int4* row) {
  const uint32_t sr = i & 7, brand = i >> 11;
  if ((int)sr >= o->nsizes) return 0;
  const uint32_t type = o->tcode[(i >> 3) & 255], pos = o->spos[sr];
  const uint32_t c = cnt[(brand * 256 + type) * 8 + pos];
  *row = make_int4((int)brand, (int)type, o->sizes[pos], (int)c);
  return c;
}

// Groups with suppliers per block of 256 ordinals.
__global__ __launch_bounds__(256) void q16_block_counts(const uint32_t* __restrict__ cnt,
                                                        const Q16Ord* __restrict__ o,
                                                        uint32_t* __restrict__ bcount) {
  int4 row;
  const int v = __syncthreads_count(q16_count_at(cnt, blockIdx.x * 256 + threadIdx.x, o, &row) != 0);
  if (threadIdx.x == 0) bcount[blockIdx.x] = (uint32_t)v;
}

// Exclusive scan of the 2048 block counts in place (one workgroup of 1024); the
// total goes to *total.
__global__ __launch_bounds__(1024) void q16_block_offsets(uint32_t* __restrict__ b,
                                                          unsigned long long* __restrict__ total) {
  __shared__ uint32_t sh[1024];
  const int t = threadIdx.x;
  const uint32_t a0 = b[2 * t], a1 = b[2 * t + 1];
  sh[t] = a0 + a1;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const uint32_t v = t >= d ? sh[t - d] : 0u;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  const uint32_t before = sh[t] - (a0 + a1);
  b[2 * t] = before;
  b[2 * t + 1] = before + a0;
  if (t == 1023) *total = sh[t];
}

// Rows of the groups with suppliers, in ordinal order (an order-preserving
// compaction): (brand, type, size, supplier count).
__global__ __launch_bounds__(256) void q16_emit(const uint32_t* __restrict__ cnt,
                                                const Q16Ord* __restrict__ o,
                                                const uint32_t* __restrict__ boff,
                                                int4* __restrict__ rows) {
  __shared__ uint32_t wsum[4];
  int4 row;
  const bool v = q16_count_at(cnt, blockIdx.x * 256 + threadIdx.x, o, &row) != 0;
  const unsigned long long m = __ballot(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t base = boff[blockIdx.x];
  for (int i = 0; i < w; ++i) base += wsum[i];
  if (v) rows[base + __popcll(m & ((1ull << lane) - 1))] = row;
}

// ---- Q20 (reference kernels/q20.metal:15-104, src/q20.cpp) ---------------------

// bits[partkey] = p_name starts with prefix.
__global__ __launch_bounds__(256) void q20_part(const int32_t* __restrict__ pk,
                                                const uint8_t* __restrict__ name, int width,
                                                uint32_t n, Word pre,
                                                uint32_t* __restrict__ bits) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint8_t* s = name + (size_t)i * width;
    bool ok = pre.len <= width;
    for (int j = 0; ok && j < pre.len; ++j) ok = s[j] == (uint8_t)pre.c[j];
    if (ok) bit_set(bits, (uint32_t)pk[i]);
  }
}

// The pair table's size is set on the device from the pair count (at least
// twice the pairs, a power of two), so the host does not wait for the count.
// The host allocated `cap` slots from an estimate; if the pairs need more,
// *slots = 0 turns the build, scan and pair kernels into no-ops and *over = 1
// tells the host to run the query again with a table sized from the count.
__global__ void q20_size(const unsigned long long* __restrict__ count, uint64_t cap,
                         uint64_t* __restrict__ slots, uint64_t* __restrict__ over) {
  const uint64_t want = 2 * (*count > 0 ? *count : 1ull);
  uint64_t s = 1;
  while (s < want) s <<= 1;
  *slots = s <= cap ? s : 0ull;
  *over = s <= cap ? 0ull : 1ull;
}

// Clears the first *slots entries of the pair table.
__global__ __launch_bounds__(256) void q20_clear(const uint64_t* __restrict__ slots,
                                                 unsigned long long* __restrict__ keys,
                                                 int64_t* __restrict__ sum,
                                                 uint32_t* __restrict__ lines) {
  const uint64_t ns = *slots;
  for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns;
       i += (uint64_t)gridDim.x * blockDim.x) {
    keys[i] = EMPTY;
    sum[i] = 0;
    lines[i] = 0u;
  }
}

// Inserts the partsupp pairs of selected parts; row[slot] = partsupp row.
__global__ __launch_bounds__(256) void q20_build(const int32_t* __restrict__ pk,
                                                 const int32_t* __restrict__ sk, uint32_t n,
                                                 const uint32_t* __restrict__ bits,
                                                 unsigned long long* __restrict__ keys,
                                                 const uint64_t* __restrict__ slots,
                                                 uint32_t* __restrict__ row) {
  if (*slots == 0) return;  // the table is too small: the host reruns
  const uint64_t mask = *slots - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t p = pk[i];
    if (!bit_test(bits, (uint32_t)p)) continue;
    bool fresh;
    const uint64_t slot = ht_insert(keys, mask, pair_key((uint32_t)p, (uint32_t)sk[i]), &fresh);
    if (fresh) row[slot] = i;
  }
}

__global__ __launch_bounds__(256) void count_bits(const int32_t* __restrict__ pk, uint32_t n,
                                                  const uint32_t* __restrict__ bits,
                                                  unsigned long long* __restrict__ cnt) {
  cnt = acc_copy(cnt);  // this workgroup's copy of the result
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    c += bit_test(bits, (uint32_t)pk[i]);
  __shared__ int64_t red[4];
  const int64_t b = block_sum((int64_t)c, red);  // one atomic per workgroup
  if (threadIdx.x == 0 && b) atomicAdd(cnt, (unsigned long long)b);
}

__global__ __launch_bounds__(256) void q20_scan(const int32_t* __restrict__ ship,
                                                const int32_t* __restrict__ pk,
                                                const int32_t* __restrict__ sk,
                                                const int32_t* __restrict__ qty, uint32_t n,
                                                int32_t lo, int32_t hi,
                                                const uint32_t* __restrict__ bits,
                                                const unsigned long long* __restrict__ keys,
                                                const uint64_t* __restrict__ slots,
                                                int64_t* __restrict__ sum,
                                                uint32_t* __restrict__ lines) {
  if (*slots == 0) return;  // the table is too small: the host reruns
  const uint64_t mask = *slots - 1;
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ship, g); }, [&](uint32_t g, const int4& d) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = lane_of(d, j);
      m |= (uint32_t)((4 * g + j < n) & (v >= lo) & (v < hi)) << j;
    }
    if (!m) return;
    const int4 p = ld4nt(pk, g);
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      m2 |= (uint32_t)(((m >> j) & 1) && bit_test(bits, (uint32_t)lane_of(p, j))) << j;
    if (!m2) return;
    const int4 s = ld4nt(sk, g), q = ld4nt(qty, g);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m2 >> j) & 1)) continue;
      const int64_t slot =
          ht_find(keys, mask, pair_key((uint32_t)lane_of(p, j), (uint32_t)lane_of(s, j)));
      if (slot < 0) continue;
      add_i64(&sum[slot], lane_of(q, j));
      atomicAdd(&lines[slot], 1u);
    }
  });
}

// good[suppkey] = 1 for pairs with lines and availqty * den > sum * num.
__global__ __launch_bounds__(256) void q20_pairs(const unsigned long long* __restrict__ keys,
                                                 const uint64_t* __restrict__ dslots,
                                                 const uint32_t* __restrict__ row,
                                                 const int64_t* __restrict__ sum,
                                                 const uint32_t* __restrict__ lines,
                                                 const int32_t* __restrict__ avail, int64_t num,
                                                 int64_t den, uint8_t* __restrict__ good) {
  const uint64_t slots = *dslots;
  for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < slots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long k = keys[i];
    if (k == EMPTY || !lines[i]) continue;
    if ((__int128)avail[row[i]] * den > (__int128)sum[i] * num) good[(uint32_t)k] = 1;
  }
}

__global__ __launch_bounds__(256) void q20_supp(const int32_t* __restrict__ sk,
                                                const int32_t* __restrict__ sn, uint32_t n,
                                                int32_t nation, const uint8_t* __restrict__ good,
                                                int32_t* __restrict__ out,
                                                unsigned long long* __restrict__ cnt) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (sn[i] == nation && good[sk[i]]) out[atomicAdd(cnt, 1ull)] = sk[i];
}

uint64_t pow2_at_least(uint64_t v) {
  uint64_t p = 1024;
  while (p < v) p <<= 1;
  return p;
}

int make_word(const char* src, Word* w) {
  w->len = 0;
  while (w->len < 31 && src[w->len]) {
    w->c[w->len] = src[w->len];
    ++w->len;
  }
  if (w->len == 31 && src[31]) return fail(GDB_E_INVALID, "pattern words are at most 31 bytes");
  for (int i = w->len; i < 32; ++i) w->c[i] = 0;
  return GDB_OK;
}

// Positions of the 25 TPC-H nation names in name order (ORDER BY n_name).
void nation_rank(uint8_t* rank) {
  std::vector<int> idx(25);
  for (int i = 0; i < 25; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [](int a, int b) {
    return std::string(gdb_nation_name(a)) < std::string(gdb_nation_name(b));
  });
  for (int i = 0; i < 256; ++i) rank[i] = 255;
  for (int r = 0; r < 25; ++r) rank[idx[r]] = (uint8_t)r;
}

// p_type code -> text, for ORDER BY p_type.
std::string type_name(int code) {
  static const char* s1[6] = {"STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO"};
  static const char* s2[5] = {"ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED"};
  static const char* s3[5] = {"TIN", "NICKEL", "BRASS", "STEEL", "COPPER"};
  if (code < 0 || code >= 150) return std::string(1, (char)code);
  return std::string(s1[code / 25]) + " " + s2[(code / 5) % 5] + " " + s3[code % 5];
}

int build_nreg_ps(gdb_ctx* ctx, const gdb_tpch_db* db, uint8_t* nreg) {
  const Column *nk, *rk;
  GDB_TRY(col(db->nation, "nation", "n_nationkey", GDB_I32, &nk));
  GDB_TRY(col(db->nation, "nation", "n_regionkey", GDB_I32, &rk));
  if (db->nation->nrows > 25 || (db->nation->nrows && (nk->minv < 0 || nk->maxv > 24)))
    return fail(GDB_E_UNSUPPORTED, "nation keys must lie in 0..24");
  nation_regions<<<1, BLOCK, 0, ctx->stream>>>((const int32_t*)nk->dev, (const int32_t*)rk->dev,
                                               (uint32_t)db->nation->nrows, nreg);
  GDB_HIP(hipGetLastError());
  return GDB_OK;
}

}  // namespace

extern "C" {

int gdb_q2(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q2_params* p, gdb_q2_result* out,
           gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  if (p->top_k < 0 || p->top_k > 100) return fail(GDB_E_INVALID, "top_k must lie in 0..100");
  const Column *pp, *pz, *pt, *pm, *sk, *sn, *sb, *qp, *qs, *qc;
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_size", GDB_I32, &pz));
  GDB_TRY(col(db->part, "part", "p_type", GDB_U8, &pt));
  GDB_TRY(col(db->part, "part", "p_mfgr", GDB_U8, &pm));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  GDB_TRY(col(db->supplier, "supplier", "s_acctbal", GDB_I64, &sb));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_partkey", GDB_I32, &qp));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_suppkey", GDB_I32, &qs));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_supplycost", GDB_I64, &qc));
  const uint64_t np = db->part->nrows, ns = db->supplier->nrows, nps = db->partsupp->nrows;
  uint64_t a, pspan, sspan;
  GDB_TRY(key_span(pp, np, &pspan));
  GDB_TRY(key_span(qp, nps, &a));
  pspan = std::max(pspan, a);
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(qs, nps, &a));
  sspan = std::max(sspan, a);
  if (sspan > (1ull << 26)) return fail(GDB_E_UNSUPPORTED, "Q2 supports supplier keys below 2^26");
  const int k = std::max(p->top_k, 1);
  uint16_t* psel;
  unsigned long long* mn;
  uint8_t *snat, *nreg, *rank;
  int64_t* sbal;
  ThkState* thk;  // top-k by account-balance histogram
  TopEntry* tcand;
  uint64_t* dout;
  char* host;
  GDB_TRY(arena(ctx, A_MAP, pspan * 2, (void**)&psel));
  GDB_TRY(arena(ctx, A_AGG, pspan * 8, (void**)&mn));
  GDB_TRY(arena(ctx, A_AUX, sspan * 9 + 512, (void**)&sbal));
  snat = (uint8_t*)(sbal + sspan);
  nreg = snat + sspan;
  rank = nreg + 256;
  GDB_TRY(thk_buffers(ctx, A_AUX2, &thk, &tcand));
  const size_t out_bytes = 16 + (size_t)k * sizeof(TopEntry) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, std::max<size_t>(out_bytes, 256), (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 2);
  uint8_t hrank[256];
  nation_rank(hrank);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(psel, 0, pspan * 2, s));
  GDB_TRY(fill(mn, 0xFF, pspan * 8, s));
  GDB_TRY(fill(snat, NONE, sspan, s));
  GDB_TRY(fill(sbal, 0, sspan * 8, s));
  GDB_TRY(fill(dout, 0, out_bytes, s));
  std::memcpy(host, hrank, 256);
  GDB_HIP(hipMemcpyAsync(rank, host, 256, hipMemcpyHostToDevice, s));
  GDB_TRY(build_nreg_ps(ctx, db, nreg));
  if (np)
    q2_part<<<blocks_for(ctx, q2_part, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const int32_t*)pz->dev, (const uint8_t*)pt->dev,
        (const uint8_t*)pm->dev, (uint32_t)np, p->size, mask4(p->type_mask), psel);
  if (ns)
    q2_supp<<<blocks_for(ctx, q2_supp, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (const int64_t*)sb->dev, (uint32_t)ns,
        nreg, p->region, snat, sbal);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 1, s));
  if (nps)
    q2_min<<<blocks_for(ctx, q2_min, nps), BLOCK, 0, s>>>(
        (const int32_t*)qp->dev, (const int32_t*)qs->dev, (const int64_t*)qc->dev, (uint32_t)nps,
        psel, snat, mn);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 2, s));
  GDB_TRY(thk_run(ctx,
                  Q2Src{(const int32_t*)qp->dev, (const int32_t*)qs->dev,
                        (const int64_t*)qc->dev, psel, snat, sbal, mn, rank},
                  nps, k, thk, tcand, fin, (unsigned long long*)dout));
  GDB_HIP(hipMemcpyAsync(host, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: part key, size (8 B), type, mfgr (2 B); supplier key, nation (8 B),
  // acctbal (8 B); partsupp key pair (8 B), supplycost (8 B), read twice. Mandatory in
  // q2_min: ps_partkey, 4 B per row (suppkey and cost only for selected parts).
  GDB_TRY(timing_finish(ctx, timing, 10 * np + 16 * ns + 32 * nps, 4 * nps));
  const uint64_t* h = (const uint64_t*)host;
  const TopEntry* hf = (const TopEntry*)(h + 2);
  out->matches = (int64_t)h[0];
  out->nrows = 0;
  for (int i = 0; i < p->top_k; ++i) {
    if (hf[i].key == INT32_MAX) break;
    gdb_q2_row& r = out->rows[out->nrows++];
    r.s_acctbal = hf[i].rev;
    r.suppkey = (int32_t)((uint32_t)hf[i].date & ((1u << 26) - 1));
    r.partkey = hf[i].key;
    r.nationkey = hf[i].prio & 0xFF;
    r.mfgr = (hf[i].prio >> 8) & 0xFF;
  }
  return GDB_OK;
}

int gdb_q11(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q11_params* p, gdb_q11_row* rows,
            int64_t cap, int64_t* nrows, int64_t* total, gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, nrows));
  if (cap < 0 || (cap > 0 && !rows) || p->frac_den <= 0 || p->frac_num < 0)
    return fail(GDB_E_INVALID, "Q11 needs cap >= 0 with rows, frac_num >= 0, frac_den > 0");
  const Column *qp, *qs, *qa, *qc, *sk, *sn;
  GDB_TRY(col(db->partsupp, "partsupp", "ps_partkey", GDB_I32, &qp));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_suppkey", GDB_I32, &qs));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_availqty", GDB_I32, &qa));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_supplycost", GDB_I64, &qc));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  const uint64_t nps = db->partsupp->nrows, ns = db->supplier->nrows;
  uint64_t a, pspan, sspan;
  GDB_TRY(key_span(qp, nps, &pspan));
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(qs, nps, &a));
  sspan = std::max(sspan, a);
  const uint64_t words = (sspan + 31) / 32;
  int64_t* val;
  uint8_t* seen;
  uint32_t* sbits;
  KV* kv;
  int64_t* dout;
  int64_t* host;
  GDB_TRY(arena(ctx, A_AGG, pspan * 8, (void**)&val));
  GDB_TRY(arena(ctx, A_MAP, pspan, (void**)&seen));
  GDB_TRY(arena(ctx, A_BITMAP, words * 4, (void**)&sbits));
  GDB_TRY(arena(ctx, A_AUX, pspan * sizeof(KV), (void**)&kv));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(val, 0, pspan * 8, s));
  GDB_TRY(fill(seen, 0, pspan, s));
  GDB_TRY(fill(sbits, 0, words * 4, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (ns)
    supp_nation_bits<<<blocks_for(ctx, supp_nation_bits, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (uint32_t)ns, p->nation, sbits);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 1, s));
  if (nps)
    q11_scan<<<blocks_for(ctx, q11_scan, quads(nps)), BLOCK, 0, s>>>(
        (const int32_t*)qp->dev, (const int32_t*)qs->dev, (const int32_t*)qa->dev,
        (const int64_t*)qc->dev, (uint32_t)nps, sbits, val, seen, dout);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 2, s));
  GDB_TRY(acc_finish(ctx, dout, 0, 1));  // q11_collect reads the total
  q11_collect<<<blocks_for(ctx, q11_collect, pspan), BLOCK, 0, s>>>(
      val, seen, (uint32_t)pspan, dout, p->frac_num, p->frac_den, kv,
      (unsigned long long*)(dout + 1));
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, dout, 16, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipStreamSynchronize(s));
  const uint64_t n = (uint64_t)host[1];
  *total = host[0];
  std::vector<KV> got(n);
  if (n) GDB_HIP(hipMemcpyAsync(got.data(), kv, n * sizeof(KV), hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: partsupp key pair, availqty (12 B), supplycost (8 B); supplier
  // key + nation (8 B). Mandatory in q11_scan: ps_suppkey, 4 B per row.
  GDB_TRY(timing_finish(ctx, timing, 20 * nps + 8 * ns, 4 * nps));
  std::sort(got.begin(), got.end(), [](const KV& x, const KV& y) {
    return x.val != y.val ? x.val > y.val : x.key < y.key;
  });
  *nrows = (int64_t)n;
  for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)n); ++i)
    rows[i] = gdb_q11_row{got[i].key, 0, got[i].val};
  return GDB_OK;
}

int gdb_q16(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q16_params* p, gdb_q16_row* rows,
            int64_t cap, int64_t* nrows, gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, nrows));
  if (cap < 0 || (cap > 0 && !rows) || p->nsizes < 0 || p->nsizes > 8)
    return fail(GDB_E_INVALID, "Q16 needs cap >= 0 with rows and 0..8 sizes");
  Word w1, w2;
  GDB_TRY(make_word(p->word1, &w1));
  GDB_TRY(make_word(p->word2, &w2));
  const Column *pp, *pb, *pt, *pz, *qp, *qs, *sk, *sc;
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_brand", GDB_U8, &pb));
  GDB_TRY(col(db->part, "part", "p_type", GDB_U8, &pt));
  GDB_TRY(col(db->part, "part", "p_size", GDB_I32, &pz));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_partkey", GDB_I32, &qp));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_suppkey", GDB_I32, &qs));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_comment", GDB_BYTES, &sc));
  const uint64_t np = db->part->nrows, nps = db->partsupp->nrows, ns = db->supplier->nrows;
  uint64_t a, pspan, sspan;
  GDB_TRY(key_span(pp, np, &pspan));
  GDB_TRY(key_span(qp, nps, &a));
  pspan = std::max(pspan, a);
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(qs, nps, &a));
  sspan = std::max(sspan, a);
  const uint64_t ngroups = 256ull * 256 * 8;
  const uint64_t slots = pow2_at_least(2 * std::max<uint64_t>(nps, 1));
  const uint32_t nblk = Q16_ORDINALS / 256;
  int32_t* gid;
  uint8_t* bad;
  unsigned long long* keys;
  uint32_t *cnt, *boff;
  int4* grp;
  uint64_t* dout;
  Q16Ord* dord;
  char* host;
  GDB_TRY(arena(ctx, A_MAP, pspan * 4, (void**)&gid));
  GDB_TRY(arena(ctx, A_AUX2, sspan, (void**)&bad));
  GDB_TRY(arena(ctx, A_HT, slots * 8, (void**)&keys));
  GDB_TRY(arena(ctx, A_AGG, ngroups * 4, (void**)&cnt));
  GDB_TRY(arena(ctx, A_AUX, Q16_ORDINALS * sizeof(int4), (void**)&grp));
  GDB_TRY(arena(ctx, A_OUT, 64 + sizeof(Q16Ord) + nblk * 4, (void**)&dout));
  dord = (Q16Ord*)(dout + 8);
  boff = (uint32_t*)(dord + 1);
  const size_t host_bytes = 64 + sizeof(Q16Ord) + Q16_ORDINALS * sizeof(int4);
  GDB_TRY(pinned(ctx, host_bytes, (void**)&host));
  // Group order: type codes by name, size positions by value (ORDER BY p_size).
  static const std::vector<int> trank = [] {  // type code -> position in name order
    std::vector<int> idx(256), rank(256);
    for (int c = 0; c < 256; ++c) idx[c] = c;
    std::sort(idx.begin(), idx.end(), [](int a, int b) { return type_name(a) < type_name(b); });
    for (int r = 0; r < 256; ++r) rank[idx[r]] = r;
    return rank;
  }();
  Q16Ord* hord = (Q16Ord*)(host + 64);
  for (int c = 0; c < 256; ++c) hord->tcode[trank[c]] = (uint8_t)c;
  for (int i = 0; i < p->nsizes; ++i) {
    int r = 0;
    for (int j = 0; j < p->nsizes; ++j)
      r += p->sizes[j] < p->sizes[i] || (p->sizes[j] == p->sizes[i] && j < i);
    hord->spos[r] = (uint8_t)i;
  }
  for (int i = 0; i < 8; ++i) hord->sizes[i] = i < p->nsizes ? p->sizes[i] : 0;
  hord->nsizes = p->nsizes;
  Q16Part qpart{};
  qpart.brand = p->brand;
  qpart.nsizes = p->nsizes;
  for (int i = 0; i < p->nsizes; ++i) qpart.sizes[i] = p->sizes[i];
  qpart.types = mask4(p->type_mask);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(hipMemcpyAsync(dord, hord, sizeof(Q16Ord), hipMemcpyHostToDevice, s));
  GDB_TRY(fill(gid, 0xFF, pspan * 4, s));
  GDB_TRY(fill(bad, 0, sspan, s));  // NOT IN keeps suppliers without a row
  GDB_TRY(fill(keys, 0xFF, slots * 8, s));
  GDB_TRY(fill(cnt, 0, ngroups * 4, s));
  GDB_TRY(fill(dout, 0, 16, s));
  if (ns)
    q16_supp<<<blocks_for(ctx, q16_supp, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const uint8_t*)sc->dev, sc->width, (uint32_t)ns, w1, w2, bad);
  if (np)
    q16_part<<<blocks_for(ctx, q16_part, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const uint8_t*)pb->dev, (const uint8_t*)pt->dev,
        (const int32_t*)pz->dev, (uint32_t)np, qpart, gid);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 1, s));
  if (nps)
    q16_scan<<<blocks_for(ctx, q16_scan, quads(nps)), BLOCK, 0, s>>>(
        (const int32_t*)qp->dev, (const int32_t*)qs->dev, (uint32_t)nps, gid, bad, keys,
        slots - 1, cnt);
  GDB_HIP(hipGetLastError());
  GDB_HIP(kernel_mark(ctx, 2, s));
  q16_block_counts<<<nblk, 256, 0, s>>>(cnt, dord, boff);
  q16_block_offsets<<<1, 1024, 0, s>>>(boff, (unsigned long long*)dout);
  q16_emit<<<nblk, 256, 0, s>>>(cnt, dord, boff, grp);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, dout, 8, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipStreamSynchronize(s));
  const uint64_t n = *(const uint64_t*)host;
  int4* got = (int4*)(host + 64 + sizeof(Q16Ord));
  if (n) GDB_HIP(hipMemcpyAsync(got, grp, n * sizeof(int4), hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: partsupp key pair (8 B); part key, size (8 B), brand, type (2 B);
  // supplier key + comment. Mandatory in q16_scan: the partsupp key pair.
  GDB_TRY(timing_finish(ctx, timing, 8 * nps + 10 * np + (4 + sc->width) * ns, 8 * nps));
  // Rows arrive in (brand, type name, size) order; a stable counting sort by
  // supplier count, descending, completes ORDER BY.
  uint32_t maxc = 0;
  for (uint64_t i = 0; i < n; ++i) maxc = std::max(maxc, (uint32_t)got[i].w);
  *nrows = (int64_t)n;
  if (maxc <= 4 * n + 1024) {
    std::vector<uint64_t> at(maxc + 2, 0);
    for (uint64_t i = 0; i < n; ++i) ++at[maxc - (uint32_t)got[i].w + 1];
    for (uint32_t c = 0; c <= maxc; ++c) at[c + 1] += at[c];
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t pos = at[maxc - (uint32_t)got[i].w]++;
      if ((int64_t)pos < cap) rows[pos] = gdb_q16_row{got[i].x, got[i].y, got[i].z, got[i].w};
    }
    return GDB_OK;
  }
  std::vector<gdb_q16_row> out(n);
  for (uint64_t i = 0; i < n; ++i) out[i] = gdb_q16_row{got[i].x, got[i].y, got[i].z, got[i].w};
  std::stable_sort(out.begin(), out.end(), [](const gdb_q16_row& a, const gdb_q16_row& b) {
    return a.supplier_cnt > b.supplier_cnt;
  });
  for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)n); ++i) rows[i] = out[i];
  return GDB_OK;
}

static int q20_run(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q20_params* p, int32_t* suppkeys,
                   int64_t cap, int64_t* nrows, gdb_timing* timing, bool* again);

int gdb_q20(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q20_params* p, int32_t* suppkeys,
            int64_t cap, int64_t* nrows, gdb_timing* timing) {
  bool again = false;
  GDB_TRY(q20_run(ctx, db, p, suppkeys, cap, nrows, timing, &again));
  if (again) GDB_TRY(q20_run(ctx, db, p, suppkeys, cap, nrows, timing, &again));
  return again ? fail(GDB_E_INVALID, "Q20 pair table sizing did not converge") : GDB_OK;
}

// One run of Q20. The pair table is sized from the last run's pair count
// (ctx->q20_pairs), or from 1/32 of partsupp on the first run; when the
// device finds more pairs than that table holds, the run produces nothing,
// records the count and sets *again, and the caller runs it once more.
static int q20_run(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q20_params* p, int32_t* suppkeys,
                   int64_t cap, int64_t* nrows, gdb_timing* timing, bool* again) {
  *again = false;
  GDB_TRY(args_ok(ctx, db, p, nrows));
  if (cap < 0 || (cap > 0 && !suppkeys) || p->frac_den <= 0 || p->frac_num < 0)
    return fail(GDB_E_INVALID, "Q20 needs cap >= 0 with suppkeys, frac_num >= 0, frac_den > 0");
  Word pre;
  GDB_TRY(make_word(p->prefix, &pre));
  const Column *pp, *pn, *qp, *qs, *qa, *ls, *lp, *lk, *lq, *sk, *sn;
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_name", GDB_BYTES, &pn));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_partkey", GDB_I32, &qp));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_suppkey", GDB_I32, &qs));
  GDB_TRY(col(db->partsupp, "partsupp", "ps_availqty", GDB_I32, &qa));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipdate", GDB_I32, &ls));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_partkey", GDB_I32, &lp));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_suppkey", GDB_I32, &lk));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_quantity", GDB_I32, &lq));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  const uint64_t np = db->part->nrows, nps = db->partsupp->nrows, nl = db->lineitem->nrows,
                 ns = db->supplier->nrows;
  uint64_t a, pspan, sspan;
  GDB_TRY(key_span(pp, np, &pspan));
  GDB_TRY(key_span(qp, nps, &a));
  pspan = std::max(pspan, a);
  GDB_TRY(key_span(lp, nl, &a));
  pspan = std::max(pspan, a);
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(qs, nps, &a));
  sspan = std::max(sspan, a);
  GDB_TRY(key_span(lk, nl, &a));
  sspan = std::max(sspan, a);
  const uint64_t words = (pspan + 31) / 32;
  uint32_t* bits;
  uint8_t* good;
  uint64_t* dout;
  uint64_t* hall;  // the count, then at most one key per supplier
  GDB_TRY(arena(ctx, A_BITMAP, words * 4, (void**)&bits));
  GDB_TRY(arena(ctx, A_AUX2, sspan, (void**)&good));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 32 + std::max<uint64_t>(ns, 1) * 4, (void**)&hall));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(bits, 0, words * 4, s));
  GDB_TRY(fill(good, 0, sspan, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (np)
    q20_part<<<blocks_for(ctx, q20_part, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const uint8_t*)pn->dev, pn->width, (uint32_t)np, pre, bits);
  if (nps)
    count_bits<<<blocks_for(ctx, count_bits, nps), BLOCK, 0, s>>>(
        (const int32_t*)qp->dev, (uint32_t)nps, bits, (unsigned long long*)dout);
  GDB_HIP(hipGetLastError());
  GDB_TRY(acc_finish(ctx, (int64_t*)dout, 0, 1));
  // The pair table holds twice the expected pairs (the last run's count, or
  // 1/32 of partsupp); its used size comes from the count on the device
  // (dout[2]), and dout[3] flags a count beyond the table.
  const uint64_t expect = ctx->q20_pairs ? ctx->q20_pairs : std::max<uint64_t>(nps / 32, 1);
  const uint64_t smax = pow2_at_least(2 * expect);
  uint64_t* dslots = dout + 2;
  q20_size<<<1, 1, 0, s>>>((const unsigned long long*)dout, smax, dslots, dout + 3);
  unsigned long long* keys;
  uint32_t *row, *lines;
  int64_t* sum;
  int32_t* out_keys;
  GDB_TRY(arena(ctx, A_HT, smax * 8, (void**)&keys));
  GDB_TRY(arena(ctx, A_AGG, smax * 8, (void**)&sum));
  GDB_TRY(arena(ctx, A_AUX, smax * 8, (void**)&row));
  lines = row + smax;
  GDB_TRY(arena(ctx, A_MAP, std::max<uint64_t>(ns, 1) * 4, (void**)&out_keys));
  q20_clear<<<blocks_for(ctx, q20_clear, smax), BLOCK, 0, s>>>(dslots, keys, sum, lines);
  if (nps)
    q20_build<<<blocks_for(ctx, q20_build, nps), BLOCK, 0, s>>>(
        (const int32_t*)qp->dev, (const int32_t*)qs->dev, (uint32_t)nps, bits, keys, dslots,
        row);
  GDB_HIP(hipGetLastError());
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {ls, lp, lk, lq}, [&](const void* const* d, uint32_t rows) {
    q20_scan<<<blocks_for(ctx, q20_scan, quads(rows)), BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
        rows, p->date_lo, p->date_hi, bits, keys, dslots, sum, lines);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  q20_pairs<<<blocks_for(ctx, q20_pairs, smax), BLOCK, 0, s>>>(
      keys, dslots, row, sum, lines, (const int32_t*)qa->dev, p->frac_num, p->frac_den, good);
  if (ns)
    q20_supp<<<blocks_for(ctx, q20_supp, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (uint32_t)ns, p->nation, good,
        out_keys, (unsigned long long*)(dout + 1));
  GDB_HIP(hipGetLastError());
  // The counts and every possible supplier key (at most one per supplier) come
  // back in one copy into pinned memory: no wait for the count before the keys.
  GDB_HIP(hipMemcpyAsync(hall, dout, 32, hipMemcpyDeviceToHost, s));
  if (ns)
    GDB_HIP(hipMemcpyAsync(hall + 4, out_keys, ns * 4, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem shipdate, partkey, suppkey, quantity (16 B); partsupp key
  // pair + availqty (12 B); part key + name; supplier key + nation. Mandatory in
  // q20_scan: l_shipdate, 4 B per line.
  GDB_TRY(timing_finish(ctx, timing, 16 * nl + 12 * nps + (4 + pn->width) * np + 8 * ns,
                        4 * nl, &rs));
  ctx->q20_pairs = std::max<uint64_t>(hall[0], 1);  // pairs found: sizes the next table
  if (hall[3]) {
    *again = true;
    return GDB_OK;
  }
  const uint64_t n = hall[1];
  int32_t* got = (int32_t*)(hall + 4);
  std::sort(got, got + n);
  *nrows = (int64_t)n;
  for (int64_t i = 0; i < std::min<int64_t>(cap, (int64_t)n); ++i) suppkeys[i] = got[i];
  return GDB_OK;
}

}  // extern "C"
