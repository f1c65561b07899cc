// TPC-H join pipelines over orders and lineitem: Q5, Q7, Q8, Q10, Q18, Q21.
//
// The dimension side of each join collapses into one byte or int32 per
// orderkey (a nation, a year, a customer), built by a scan of orders over the
// customer / supplier maps, so the lineitem scan needs a single gather per
// surviving row, in orderkey order (lineitem is clustered by order, so the
// gathers walk the map nearly sequentially). References: src/qN.cpp and
// kernels/qN.metal, cited per pipeline.
#include "tpch_common.h"

using namespace gdb;
using namespace gdbk;
using namespace gdbt;

namespace {

constexpr uint8_t NONE = 255;

struct Ld3 {  // three streamed int32 columns of one vector group
  int4 a, b, c;
};

// ---- Q5 (reference kernels/q5.metal:19-125, src/q5.cpp) ------------------------

// cmap[custkey] = nation if the nation's region is `region`, else NONE.
__global__ __launch_bounds__(256) void cust_nation_in_region(const int32_t* __restrict__ ck,
                                                             const int32_t* __restrict__ cn,
                                                             uint32_t n,
                                                             const uint8_t* __restrict__ nreg,
                                                             int32_t region,
                                                             uint8_t* __restrict__ cmap) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t nat = cn[i];
    const bool ok = nat >= 0 && nat < 25 && nreg[nat] == region;  // LDS bins hold 25 nations
    cmap[ck[i]] = ok ? (uint8_t)nat : NONE;
  }
}

// omap[orderkey] = cmap[custkey] for orders dated in [lo, hi) (NONE otherwise).
__global__ __launch_bounds__(256) void orders_cust_byte(const int32_t* __restrict__ ok,
                                                        const int32_t* __restrict__ oc,
                                                        const int32_t* __restrict__ od, uint32_t n,
                                                        int32_t lo, int32_t hi,
                                                        const uint8_t* __restrict__ cmap,
                                                        uint32_t cspan,
                                                        uint8_t* __restrict__ omap) {
  walk2(quads(n), [&](uint32_t g) { return Ld3{ld4nt(ok, g), ld4nt(oc, g), ld4nt(od, g)}; },
        [&](uint32_t g, const Ld3& l) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) continue;
      const int dv = lane_of(l.c, j);
      const uint32_t cv = (uint32_t)lane_of(l.b, j);
      omap[lane_of(l.a, j)] = (dv >= lo && dv < hi && cv < cspan) ? cmap[cv] : NONE;
    }
  });
}

__global__ __launch_bounds__(256) void q5_scan(const int32_t* __restrict__ ok,
                                               const int32_t* __restrict__ sk,
                                               const int32_t* __restrict__ disc,
                                               const int64_t* __restrict__ price, uint32_t n,
                                               const uint8_t* __restrict__ omap,
                                               const uint8_t* __restrict__ smap,
                                               int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t rev[32];
  __shared__ uint32_t hit;
  if (threadIdx.x < 32) rev[threadIdx.x] = 0;
  if (threadIdx.x == 0) hit = 0;
  __syncthreads();
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ok, g); }, [&](uint32_t g, const int4& k) {
    uint32_t on[4];
    gather4(omap, k, on);
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m |= (uint32_t)(4 * g + j < n && on[j] != NONE) << j;
    if (!m) return;
    const int4 s = ld4nt(sk, g);
    uint32_t sn[4];
    gather4(smap, s, sn);
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m2 |= (uint32_t)(((m >> j) & 1) && sn[j] == on[j]) << j;
    if (!m2) return;
    const int4 d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((m2 >> j) & 1) {
        lds_add_i64(&rev[on[j]], lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j)));
        atomicOr(&hit, 1u << on[j]);
      }
  });
  __syncthreads();
  if (threadIdx.x < 25 && ((hit >> threadIdx.x) & 1)) add_i64(&out[threadIdx.x], rev[threadIdx.x]);
  if (threadIdx.x == 0 && hit) atomicOr(reinterpret_cast<unsigned int*>(out + 25), hit);
}

// ---- Q7 (reference kernels/q7.metal:13-109, src/q7.cpp) ------------------------

// map[key] = 0 for nation n1, 1 for n2, NONE otherwise.
__global__ __launch_bounds__(256) void pair_side(const int32_t* __restrict__ key,
                                                 const int32_t* __restrict__ nat, uint32_t n,
                                                 int32_t n1, int32_t n2,
                                                 uint8_t* __restrict__ map) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    map[key[i]] = nat[i] == n1 ? 0 : nat[i] == n2 ? 1 : NONE;
}

// omap[orderkey] = cmap[custkey] for every order.
__global__ __launch_bounds__(256) void orders_side(const int32_t* __restrict__ ok,
                                                   const int32_t* __restrict__ oc, uint32_t n,
                                                   const uint8_t* __restrict__ cmap,
                                                   uint32_t cspan, uint8_t* __restrict__ omap) {
  constexpr int QB = 8;  // groups per lane per trip, each load phase issued for all
  walk_batch<QB>(quads(n), [&](const uint32_t* g, const bool* live) {
    int4 k[QB], c[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (live[b]) {
        k[b] = ld4nt(ok, g[b]);
        c[b] = ld4nt(oc, g[b]);
      }
    uint32_t cb[QB][4];
#pragma unroll
    for (int b = 0; b < QB; ++b)
      if (live[b])
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t cv = (uint32_t)lane_of(c[b], j);
          cb[b][j] = cv < cspan ? cmap[cv] : NONE;
        }
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      if (!live[b]) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g[b] + j < n) omap[lane_of(k[b], j)] = (uint8_t)cb[b][j];
    }
  });
}

__global__ __launch_bounds__(256) void q7_scan(const int32_t* __restrict__ ship,
                                               const int32_t* __restrict__ ok,
                                               const int32_t* __restrict__ sk,
                                               const int32_t* __restrict__ disc,
                                               const int64_t* __restrict__ price, uint32_t n,
                                               int32_t lo, int32_t hi, int same,
                                               const uint8_t* __restrict__ omap,
                                               const uint8_t* __restrict__ smap,
                                               int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ int64_t rev[16];
  __shared__ uint32_t hit;
  if (threadIdx.x < 16) rev[threadIdx.x] = 0;
  if (threadIdx.x == 0) hit = 0;
  __syncthreads();
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ship, g); }, [&](uint32_t g, const int4& s) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int v = lane_of(s, j);
      m |= (uint32_t)((4 * g + j < n) & (v >= lo) & (v <= hi)) << j;
    }
    if (!m) return;
    const int4 k = ld4nt(ok, g);
    uint32_t cs[4];
    gather4(omap, k, cs);
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m2 |= (uint32_t)(((m >> j) & 1) && cs[j] != NONE) << j;
    if (!m2) return;
    const int4 sp = ld4nt(sk, g);
    uint32_t ss[4];
    gather4(smap, sp, ss);
    uint32_t m3 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      m3 |= (uint32_t)(((m2 >> j) & 1) && ss[j] != NONE && (ss[j] != cs[j] || same)) << j;
    if (!m3) return;
    const int4 d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((m3 >> j) & 1) {
        const int y = year_index(lane_of(s, j));
        const int bin = (int)ss[j] * 8 + y;
        lds_add_i64(&rev[bin], lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j)));
        atomicOr(&hit, 1u << bin);
      }
  });
  __syncthreads();
  if (threadIdx.x < 16 && ((hit >> threadIdx.x) & 1)) add_i64(&out[threadIdx.x], rev[threadIdx.x]);
  if (threadIdx.x == 0 && hit) atomicOr(reinterpret_cast<unsigned int*>(out + 16), hit);
}

// ---- Q8 (reference kernels/q8.metal:10-76, src/q8.cpp) -------------------------

// omap[orderkey] = year index of orders dated in [lo, hi] whose customer's
// nation lies in `region`; NONE otherwise.
__global__ __launch_bounds__(256) void q8_orders(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ oc,
                                                 const int32_t* __restrict__ od, uint32_t n,
                                                 int32_t lo, int32_t hi,
                                                 const uint8_t* __restrict__ cnat, uint32_t cspan,
                                                 const uint8_t* __restrict__ nreg, int32_t region,
                                                 uint8_t* __restrict__ omap) {
  walk2(quads(n), [&](uint32_t g) { return Ld3{ld4nt(ok, g), ld4nt(oc, g), ld4nt(od, g)}; },
        [&](uint32_t g, const Ld3& l) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) continue;
      const int dv = lane_of(l.c, j);
      const uint32_t cv = (uint32_t)lane_of(l.b, j);
      uint8_t y = NONE;
      if (dv >= lo && dv <= hi && cv < cspan) {
        const uint8_t nat = cnat[cv];
        if (nat != NONE && nreg[nat] == region) y = (uint8_t)year_index(dv);
      }
      omap[lane_of(l.a, j)] = y;
    }
  });
}

__global__ __launch_bounds__(FOLD_BLOCK) void q8_scan(const int32_t* __restrict__ pk,
                                                      const int32_t* __restrict__ ok,
                                                      const int32_t* __restrict__ sk,
                                                      const int32_t* __restrict__ disc,
                                                      const int64_t* __restrict__ price,
                                                      uint32_t n,
                                                      const uint32_t* __restrict__ pbits,
                                                      const uint32_t* __restrict__ pfold,
                                                      const uint8_t* __restrict__ omap,
                                                      const uint8_t* __restrict__ smap,
                                                      int32_t nation, int64_t* __restrict__ out) {
  out = acc_copy(out);  // this workgroup's copy of the result
  __shared__ __align__(16) uint32_t lfold[FOLD_WORDS];
  __shared__ int64_t vol[16];  // [0..8) nation volume, [8..16) total volume by year
  __shared__ uint32_t hit;
  if (threadIdx.x < 16) vol[threadIdx.x] = 0;
  if (threadIdx.x == 0) hit = 0;
  fold_load(lfold, pfold);
  walk2(quads(n), [&](uint32_t g) { return ld4nt(pk, g); }, [&](uint32_t g, const int4& p) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t key = (uint32_t)lane_of(p, j);
      m |= (uint32_t)((4 * g + j < n) && fold_test(lfold, key)) << j;
    }
    if (!m) return;
    m &= bits4(pbits, p);
    if (!m) return;
    const int4 k = ld4nt(ok, g);
    uint32_t y[4];
    gather4(omap, k, y);
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m2 |= (uint32_t)(((m >> j) & 1) && y[j] < 8) << j;
    if (!m2) return;
    const int4 sp = ld4nt(sk, g), d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
    uint32_t sn[4];
    gather4(smap, sp, sn);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m2 >> j) & 1)) continue;
      const int64_t v = lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j));
      lds_add_i64(&vol[8 + y[j]], v);
      if (sn[j] == (uint32_t)nation) lds_add_i64(&vol[y[j]], v);
      atomicOr(&hit, 1u << y[j]);
    }
  });
  __syncthreads();
  if (threadIdx.x < 16 && vol[threadIdx.x]) add_i64(&out[threadIdx.x], vol[threadIdx.x]);
  if (threadIdx.x == 0 && hit) atomicOr(reinterpret_cast<unsigned int*>(out + 16), hit);
}

// ---- Q10 (reference kernels/q10.metal:9-54, src/q10.cpp) -----------------------

// ocust[orderkey] = custkey for orders dated in [lo, hi) (-1 otherwise).
// Orders in the date window: bit orderkey of `win` set and ocust[orderkey] =
// custkey (ocust is read only where the bit is set, so it needs no clearing).
__global__ __launch_bounds__(256) void q10_orders(const int32_t* __restrict__ ok,
                                                  const int32_t* __restrict__ oc,
                                                  const int32_t* __restrict__ od, uint32_t n,
                                                  int32_t lo, int32_t hi,
                                                  uint32_t* __restrict__ win,
                                                  int32_t* __restrict__ ocust) {
  constexpr int QB = 8;  // groups per lane per trip, each load phase issued for all
  walk_batch<QB>(quads(n), [&](const uint32_t* g, const bool* live) {
    int4 d[QB], k[QB], c[QB];
    uint32_t m[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) d[b] = live[b] ? ld4nt(od, g[b]) : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      m[b] = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int dv = lane_of(d[b], j);
        m[b] |= (uint32_t)(live[b] && 4 * g[b] + j < n && dv >= lo && dv < hi) << j;
      }
      if (m[b]) {
        k[b] = ld4nt(ok, g[b]);
        c[b] = ld4nt(oc, g[b]);
      }
    }
#pragma unroll
    for (int b = 0; b < QB; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if ((m[b] >> j) & 1) {
          bit_set(win, (uint32_t)lane_of(k[b], j));
          ocust[lane_of(k[b], j)] = lane_of(c[b], j);
        }
  });
}

// Q10 over lineitem: lines with the return flag whose order is in the window
// (bit set in `win`) add their revenue to their order's customer. Flag and key
// load together, so the probe's only dependent load is the bitmap word. The rare
// hits are first listed as chunk-relative rows in the wave's own region of `cap`
// rows of `hits` (at least the rows it visits; the count stays in a register, so
// no atomics), then the wave finishes its list: order -> customer, revenue.
__global__ __launch_bounds__(256) void q10_probe(const uint8_t* __restrict__ rflag,
                                                 const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ disc,
                                                 const int64_t* __restrict__ price, uint32_t n,
                                                 uint32_t flag, const uint32_t* __restrict__ win,
                                                 const int32_t* __restrict__ ocust, uint32_t cspan,
                                                 uint32_t cap, uint32_t* __restrict__ hits,
                                                 int64_t* __restrict__ crev,
                                                 uint8_t* __restrict__ seen) {
  const int lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nq = quads(n), stride = gridDim.x * blockDim.x;
  uint32_t* out = hits + (size_t)wave * cap;
  uint32_t cnt = 0;  // wave-uniform: every trip runs on all lanes
  auto probe = [&](uint32_t g, uint32_t fw, const int4& k) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      m |= (uint32_t)((4 * g + j < n) && byte_of(fw, j) == flag &&
                      bit_test(win, (uint32_t)lane_of(k, j))) << j;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool h = (m >> j) & 1;
      const unsigned long long b = __ballot(h);
      if (h) out[cnt + __popcll(b & ((1ull << lane) - 1))] = 4 * g + j;
      cnt += (uint32_t)__popcll(b);
    }
  };
  const int4 z = make_int4(0, 0, 0, 0);
  // Two quads per lane in flight: g and g + stride.
  for (uint32_t gw = wave * 64; gw < nq; gw += 2 * stride) {
    const uint32_t g0 = gw + lane, g1 = g0 + stride;
    const bool l0 = g0 < nq, l1 = g1 < nq;
    const uint32_t f0 = l0 ? ld_u8x4nt(rflag, g0) : 0u, f1 = l1 ? ld_u8x4nt(rflag, g1) : 0u;
    const int4 k0 = l0 ? ld4nt(ok, g0) : z, k1 = l1 ? ld4nt(ok, g1) : z;
    probe(g0, f0, k0);
    probe(g1, f1, k1);
  }
  __threadfence_block();  // the list's stores are seen by the wave's other lanes
  for (uint32_t i = lane; i < cnt; i += 64) {
    const uint32_t r = out[i];
    const int32_t cu = ocust[ok[r]];
    if (cu < 0 || (uint32_t)cu >= cspan) continue;
    add_i64(&crev[cu], price[r] * (int64_t)(100 - disc[r]));
    seen[cu] = 1;
  }
}

struct KeyedRevSrc {  // (revenue desc, key asc) over a per-key revenue array
  const int64_t* rev;
  const uint8_t* seen;
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    e = TopEntry{rev[i], 0, (int32_t)i, 0, 0};
    return seen[i] != 0;
  }
};

// rank[key] = position of key among the k winners (written for winners only).
__global__ void mark_ranks(const TopEntry* __restrict__ win, int k, uint8_t* __restrict__ rank) {
  for (int i = threadIdx.x; i < k; i += blockDim.x)
    if (win[i].key != INT32_MAX) rank[win[i].key] = (uint8_t)i;
}

// Gathers c_acctbal / c_nationkey of the winners by scanning customer rows.
__global__ __launch_bounds__(256) void gather_customers(const int32_t* __restrict__ ck,
                                                        const int32_t* __restrict__ cn,
                                                        const int64_t* __restrict__ cb, uint32_t n,
                                                        const uint8_t* __restrict__ rank,
                                                        uint32_t span, int64_t* __restrict__ bal,
                                                        int32_t* __restrict__ nat) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t key = (uint32_t)ck[i];
    if (key >= span) continue;
    const uint8_t r = rank[key];
    if (r != NONE) {
      bal[r] = cb[i];
      nat[r] = cn[i];
    }
  }
}

// ---- Q18 (reference kernels/q18.metal:9-61, src/q18.cpp) -----------------------

// sq[orderkey] += quantity; consecutive rows of one order within a lane's group
// of 4 are summed before the atomic (lineitem is clustered by order).
__global__ __launch_bounds__(256) void q18_sums(const int32_t* __restrict__ ok,
                                                const int32_t* __restrict__ qty, uint32_t n,
                                                int32_t* __restrict__ sq) {
  const uint32_t nq = quads(n), stride = gridDim.x * blockDim.x;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nq; g += stride) {
    const int4 k = ld4nt(ok, g), q = ld4nt(qty, g);
    int key = lane_of(k, 0), acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) break;
      const int kj = lane_of(k, j);
      if (kj != key) {
        atomicAdd(&sq[key], acc);
        key = kj;
        acc = 0;
      }
      acc += lane_of(q, j);
    }
    if (acc) atomicAdd(&sq[key], acc);
  }
}

struct Q18Src {  // orders over the threshold: (totalprice desc, orderdate asc, orderkey asc)
  const int32_t *ok, *od;
  const int64_t* tp;
  const int32_t* sq;
  int32_t thr;
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    const int32_t key = ok[i];
    e = TopEntry{tp[i], od[i], key, (int32_t)i, 0};
    return sq[key] > thr;
  }
};

__global__ void q18_gather(const TopEntry* __restrict__ win, int k,
                           const int32_t* __restrict__ oc, const int32_t* __restrict__ sq,
                           int32_t* __restrict__ out) {
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    if (win[i].key == INT32_MAX) continue;
    out[2 * i] = oc[win[i].prio];
    out[2 * i + 1] = sq ? sq[win[i].key] : win[i].pad;
  }
}

// ---- Q21 (reference kernels/q21.metal:24-119, src/q21.cpp) ---------------------

// fst[orderkey] = (o_orderstatus == status), plain byte stores (no atomics).
__global__ __launch_bounds__(256) void order_status_bytes(const int32_t* __restrict__ ok,
                                                          const uint8_t* __restrict__ st,
                                                          uint32_t n, uint32_t status,
                                                          uint8_t* __restrict__ fst) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    fst[ok[i]] = st[i] == status;
}

// Per-order state in one 64-bit word:
//   bits 32..62 first supplier + 1, bit 63 "a second supplier seen",
//   bits 0..30 first late supplier + 1, bit 31 "a second late supplier seen".

__device__ __forceinline__ uint64_t q21_merge(uint64_t a, uint64_t b) {
  const uint32_t af = (uint32_t)(a >> 32) & 0x7FFFFFFFu, bf = (uint32_t)(b >> 32) & 0x7FFFFFFFu;
  const uint32_t al = (uint32_t)a & 0x7FFFFFFFu, bl = (uint32_t)b & 0x7FFFFFFFu;
  const uint32_t f = af ? af : bf, l = al ? al : bl;
  const uint32_t mf = (uint32_t)((a >> 63) | (b >> 63) | (af && bf && af != bf));
  const uint32_t ml = ((uint32_t)(a >> 31) & 1) | ((uint32_t)(b >> 31) & 1) | (al && bl && al != bl);
  return ((uint64_t)(f | (mf << 31)) << 32) | (uint64_t)(l | (ml << 31));
}

__device__ __forceinline__ void q21_apply(unsigned long long* st, uint64_t v) {
  unsigned long long old = *st;
  for (;;) {
    const unsigned long long nw = q21_merge(old, v);
    if (nw == old) return;
    const unsigned long long prev = atomicCAS(st, old, nw);
    if (prev == old) return;
    old = prev;
  }
}

// Pass 1 over lines of orders with the wanted status: merge each run of one
// order's rows within a lane's group of 4 into the order's state.
__global__ __launch_bounds__(256) void q21_track(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ sk,
                                                 const int32_t* __restrict__ commit,
                                                 const int32_t* __restrict__ receipt, uint32_t n,
                                                 const uint8_t* __restrict__ fst,
                                                 unsigned long long* __restrict__ state) {
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ok, g); }, [&](uint32_t g, const int4& k) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m |= (uint32_t)((4 * g + j < n) && fst[lane_of(k, j)]) << j;
    if (!m) return;
    const int4 s = ld4nt(sk, g), c = ld4nt(commit, g), r = ld4nt(receipt, g);
    int key = -1;
    uint64_t acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1)) continue;
      const int kj = lane_of(k, j);
      const uint32_t sv = (uint32_t)lane_of(s, j) + 1u;
      const bool late = lane_of(r, j) > lane_of(c, j);
      const uint64_t v = ((uint64_t)sv << 32) | (late ? sv : 0u);
      if (kj != key) {
        if (key >= 0) q21_apply(&state[key], acc);
        key = kj;
        acc = v;
      } else {
        acc = q21_merge(acc, v);
      }
    }
    if (key >= 0) q21_apply(&state[key], acc);
  });
}

// Pass 2 over lineitem once every order's state is final: a late line of the
// nation's supplier on an order of the wanted status counts when the order has
// a second supplier and no second late supplier (the line's own supplier is
// then the only late one).
__global__ __launch_bounds__(256) void q21_count_rows(const int32_t* __restrict__ ok,
                                                      const int32_t* __restrict__ sk,
                                                      const int32_t* __restrict__ commit,
                                                      const int32_t* __restrict__ receipt,
                                                      uint32_t n, const uint8_t* __restrict__ fst,
                                                      const uint8_t* __restrict__ smap,
                                                      int32_t nation,
                                                      const unsigned long long* __restrict__ state,
                                                      uint32_t* __restrict__ cnt) {
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ok, g); }, [&](uint32_t g, const int4& k) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m |= (uint32_t)((4 * g + j < n) && fst[lane_of(k, j)]) << j;
    if (!m) return;
    const int4 s = ld4nt(sk, g), c = ld4nt(commit, g), r = ld4nt(receipt, g);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m >> j) & 1) || !(lane_of(r, j) > lane_of(c, j))) continue;
      const uint32_t sv = (uint32_t)lane_of(s, j);
      if (smap[sv] != nation) continue;
      const uint64_t st = state[lane_of(k, j)];
      if ((st >> 63) && !((st >> 31) & 1)) atomicAdd(&cnt[sv], 1u);
    }
  });
}

// ---- clustered lineitem: one lane resolves each order ---------------------------
// Per-order results over a clustered lineitem (l_orderkey non-decreasing, so
// an order's rows are contiguous), as a segmented suffix scan across the wave.
// A lane holds 4 rows; a boundary is a row that starts an order (or lies past n).
// The order state St is associative: St::merge(a, b) combines the rows of a with
// the later rows of b, and St{} is empty. Within a lane, r[j] folds rows j.. up
// to the next boundary and e[j] says that run reaches the lane's end; `lead`
// folds the lane's rows before its first boundary and `pass` says the lane has
// none. The suffix scan gives lane L the fold of the leads of lanes L+1.. up to
// and including the first lane with a boundary. Past the wave's last lane the
// run continues into quads g0+64 and g0+65, which lanes 0 and 1 load beside their
// own, and for orders longer than that into rows read from memory. The order
// starting at row j of lane L is then r[j], merged with that carry when e[j]:
// the lane holding an order's first row has its whole state and acts alone.

// Lanes 0 and 1 load quads g0+64 and g0+65 (zero past the column's nq quads).
__device__ __forceinline__ int4 ext_quad(const int32_t* __restrict__ c, uint32_t g0, uint32_t nq,
                                         int lane) {
  const uint32_t e = g0 + 64 + (uint32_t)lane;
  return lane < 2 && e < nq ? ld4nt(c, e) : make_int4(0, 0, 0, 0);
}

// Lane 0 of a wave reads the row before its quad (the last row of quad g-1).
__device__ __forceinline__ int prev_row(const int32_t* __restrict__ ok, uint32_t g, bool live,
                                        int lane) {
  return lane == 0 && live && g > 0 ? ok[4 * g - 1] : 0;
}

__device__ __forceinline__ int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <class St>
struct LaneRuns {
  St r[4];
  bool e[4];
  uint32_t head;  // bit j: row j starts an order
  St lead;
  bool pass;
};

// Runs of the 4 rows at row0.. with keys k and row states row[]; `prev` is the
// key of row row0-1 (ignored for row 0 of the table).
template <class St>
__device__ __forceinline__ void lane_runs(const int4& k, const St* row, int prev, uint32_t row0,
                                          uint32_t n, LaneRuns<St>& o) {
  const int key[4] = {k.x, k.y, k.z, k.w};
  bool b[4], valid[4];
  o.head = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    valid[j] = row0 + j < n;
    const bool h = valid[j] && (row0 + j == 0 || key[j] != (j ? key[j - 1] : prev));
    b[j] = !valid[j] || h;
    o.head |= (uint32_t)h << j;
  }
  o.r[3] = valid[3] ? row[3] : St{};
  o.e[3] = valid[3];
#pragma unroll
  for (int j = 2; j >= 0; --j) {
    const St rj = valid[j] ? row[j] : St{};
    o.r[j] = b[j + 1] ? rj : St::merge(rj, o.r[j + 1]);
    o.e[j] = !b[j + 1] && o.e[j + 1];
  }
  o.lead = b[0] ? St{} : o.r[0];
  o.pass = !b[0] && o.e[0];
}

// The carry each lane merges into runs that reach its end. `ext` continues the
// wave's last run (the leads past lane 63).
template <class St>
__device__ __forceinline__ St wave_carry(const LaneRuns<St>& lr, const St& ext, int lane) {
  St v = lr.lead;
  bool closed = !lr.pass;
  auto step = [&](int d) {
    const St v2 = v.shfl_down(d);
    const int c2 = __shfl_down((int)closed, d, 64);
    if (lane + d < 64 && !closed) {
      v = St::merge(v, v2);
      closed = c2 != 0;
    }
  };
  // After distances 1 and 2 a lane's fold covers 4 lanes (16 rows): every run is
  // closed or reaches the wave's end unless an order spans more, which TPC-H's
  // 1..7 lines never do. The longer distances run only when a lane needs them
  // (decided for the whole wave).
  step(1);
  step(2);
  if (__any(!closed && lane + 4 < 64)) {
#pragma unroll
    for (int d = 4; d < 64; d <<= 1) step(d);
  }
  if (!closed) v = St::merge(v, ext);
  const St nxt = v.shfl_down(1);
  return lane < 63 ? nxt : ext;
}

// The fold past lane 63: lanes 0 and 1 run over their extra quads (keys ek, row
// states xrow), then `more(r, key)` folds rows from r on while they carry `key`
// (wave-uniform; only for orders longer than the two extra quads).
template <class St, class More>
__device__ __forceinline__ St ext_carry(const int4& k, const int4& ek, const St* xrow,
                                        uint32_t gw, uint32_t nq, uint32_t n, int lane,
                                        More more) {
  const int prev = lane == 0 ? read_lane(k.w, 63) : read_lane(ek.w, 0);
  const uint32_t e = gw + 64 + (uint32_t)lane;
  LaneRuns<St> x;
  lane_runs(ek, xrow, prev, 4 * e, lane < 2 && e < nq ? n : 0, x);
  St ext = x.lead.read(0);
  if (read_lane((int)x.pass, 0)) {
    ext = St::merge(ext, x.lead.read(1));
    if (read_lane((int)x.pass, 1)) ext = St::merge(ext, more(4 * (gw + 66), read_lane(ek.w, 1)));
  }
  return ext;
}

// Q18: quantity sum per order.
struct QtySt {
  int v = 0;
  __device__ static QtySt merge(const QtySt& a, const QtySt& b) { return QtySt{a.v + b.v}; }
  __device__ QtySt shfl_down(int d) const { return QtySt{__shfl_down(v, d, 64)}; }
  __device__ QtySt read(int l) const { return QtySt{read_lane(v, l)}; }
};

// Q21: per order, the least and greatest supplier (+1) over its lines and over
// its late lines, and the late line count. A second (late) supplier appears iff
// the least and greatest differ. Merging is min / max / add, so the empty state
// (least ~0, greatest 0, count 0) is the identity.
struct SuppSt {
  uint32_t lo = ~0u, hi = 0u, llo = ~0u, lhi = 0u, lc = 0u;
  __device__ static SuppSt row(int s, int c, int r) {
    const uint32_t s1 = (uint32_t)s + 1u;
    const bool late = r > c;
    return SuppSt{s1, s1, late ? s1 : ~0u, late ? s1 : 0u, late ? 1u : 0u};
  }
  __device__ static SuppSt merge(const SuppSt& a, const SuppSt& b) {
    return SuppSt{min(a.lo, b.lo), max(a.hi, b.hi), min(a.llo, b.llo), max(a.lhi, b.lhi),
                  a.lc + b.lc};
  }
  __device__ SuppSt shfl_down(int d) const {
    return SuppSt{(uint32_t)__shfl_down((int)lo, d, 64), (uint32_t)__shfl_down((int)hi, d, 64),
                  (uint32_t)__shfl_down((int)llo, d, 64), (uint32_t)__shfl_down((int)lhi, d, 64),
                  (uint32_t)__shfl_down((int)lc, d, 64)};
  }
  __device__ SuppSt read(int l) const {
    return SuppSt{(uint32_t)read_lane((int)lo, l), (uint32_t)read_lane((int)hi, l),
                  (uint32_t)read_lane((int)llo, l), (uint32_t)read_lane((int)lhi, l),
                  (uint32_t)read_lane((int)lc, l)};
  }
};

// Q21 over a clustered lineitem: an order of the wanted status counts its late
// lines for their supplier when it has a second supplier and every late line has
// the same supplier (of `nation`).
__global__ __launch_bounds__(256) void q21_owner(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ sk,
                                                 const int32_t* __restrict__ commit,
                                                 const int32_t* __restrict__ receipt, uint32_t n,
                                                 const uint8_t* __restrict__ fst,
                                                 const uint8_t* __restrict__ smap, int32_t nation,
                                                 uint32_t* __restrict__ cnt) {
  const uint32_t nq = quads(n), stride = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  const int4 z = make_int4(0, 0, 0, 0);
  for (uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; gw < nq; gw += stride) {
    const uint32_t g = gw + lane;
    const bool live = g < nq;
    const int4 k = live ? ld4nt(ok, g) : z, s = live ? ld4nt(sk, g) : z,
               c = live ? ld4nt(commit, g) : z, rc = live ? ld4nt(receipt, g) : z;
    const int4 ek = ext_quad(ok, gw, nq, lane), es = ext_quad(sk, gw, nq, lane),
               ec = ext_quad(commit, gw, nq, lane), er = ext_quad(receipt, gw, nq, lane);
    const int p0 = prev_row(ok, g, live, lane);
    // Order status of every row's key, loaded at once (keys of dead lanes are 0).
    const uint8_t f[4] = {fst[k.x], fst[k.y], fst[k.z], fst[k.w]};
    const int up = __shfl_up(k.w, 1, 64);
    SuppSt row[4], xrow[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      row[j] = SuppSt::row(lane_of(s, j), lane_of(c, j), lane_of(rc, j));
      xrow[j] = SuppSt::row(lane_of(es, j), lane_of(ec, j), lane_of(er, j));
    }
    LaneRuns<SuppSt> lr;
    lane_runs(k, row, lane > 0 ? up : p0, 4 * g, live ? n : 0, lr);
    const SuppSt ext = ext_carry<SuppSt>(k, ek, xrow, gw, nq, n, lane, [&](uint32_t r, int key) {
      SuppSt a;
      for (; r < n && ok[r] == key; ++r) a = SuppSt::merge(a, SuppSt::row(sk[r], commit[r], receipt[r]));
      return a;
    });
    const SuppSt carry = wave_carry(lr, ext, lane);
    uint32_t late[4], supp[4];
    bool want[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const SuppSt t = lr.e[j] ? SuppSt::merge(lr.r[j], carry) : lr.r[j];
      late[j] = t.lc;
      want[j] = ((lr.head >> j) & 1) && f[j] && t.lo != t.hi && late[j] && t.llo == t.lhi;
      supp[j] = want[j] ? t.llo - 1 : 0u;
    }
    // Supplier nations of the candidates, loaded at once.
    const uint8_t sn[4] = {smap[supp[0]], smap[supp[1]], smap[supp[2]], smap[supp[3]]};
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (want[j] && sn[j] == nation) atomicAdd(&cnt[supp[j]], late[j]);
  }
}

// Q18 over a clustered lineitem: the quantity sum per order. EMIT = false: a
// plain store into sq[orderkey]. EMIT = true: orders over `thr` are appended as
// (key, sum) to `cand` (one atomic per wave and row slot), so nothing keyed by
// orderkey needs clearing and the top-k runs over the few candidates.
template <bool EMIT>
__global__ __launch_bounds__(256) void q18_owner(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ qty, uint32_t n,
                                                 int32_t* __restrict__ sq, int32_t thr,
                                                 int2* __restrict__ cand,
                                                 uint32_t* __restrict__ ncand) {
  const uint32_t nq = quads(n), stride = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  const int4 z = make_int4(0, 0, 0, 0);
  for (uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; gw < nq; gw += stride) {
    const uint32_t g = gw + lane;
    const bool live = g < nq;
    const int4 k = live ? ld4nt(ok, g) : z, q = live ? ld4nt(qty, g) : z;
    const int4 ek = ext_quad(ok, gw, nq, lane), eq = ext_quad(qty, gw, nq, lane);
    const int p0 = prev_row(ok, g, live, lane);
    const int up = __shfl_up(k.w, 1, 64);
    const QtySt row[4] = {{q.x}, {q.y}, {q.z}, {q.w}}, xrow[4] = {{eq.x}, {eq.y}, {eq.z}, {eq.w}};
    LaneRuns<QtySt> lr;
    lane_runs(k, row, lane > 0 ? up : p0, 4 * g, live ? n : 0, lr);
    const QtySt ext = ext_carry<QtySt>(k, ek, xrow, gw, nq, n, lane, [&](uint32_t r, int key) {
      QtySt a;
      for (; r < n && ok[r] == key; ++r) a.v += qty[r];
      return a;
    });
    const QtySt carry = wave_carry(lr, ext, lane);
    const int key[4] = {k.x, k.y, k.z, k.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool mine = (lr.head >> j) & 1;
      const int sum = lr.r[j].v + (lr.e[j] ? carry.v : 0);
      if (EMIT) {  // every lane reaches the ballot
        const bool emit = mine && sum > thr;
        const unsigned long long m = __ballot(emit);
        if (m) {
          const int leader = __ffsll((long long)m) - 1;
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(ncand, (uint32_t)__popcll(m));
          base = __shfl(base, leader, 64);
          if (emit) cand[base + __popcll(m & ((1ull << lane) - 1))] = make_int2(key[j], sum);
        }
      } else if (mine) {
        sq[key[j]] = sum;
      }
    }
  }
}

// Candidates (orderkey, sum) looked up in the orders table, whose keys ascend
// strictly: the first row with o_orderkey >= key by binary search. Keys with no
// orders row are not candidates (the join drops them).
struct Q18CandSrc {
  const int2* cand;
  const int32_t *ok, *od;
  const int64_t* tp;
  uint32_t no;
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    const int2 c = cand[i];
    uint32_t lo = 0, hi = no;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ok[mid] < c.x) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t row = lo < no ? lo : no - 1;
    e = TopEntry{tp[row], od[row], c.x, (int32_t)row, c.y};
    return lo < no && ok[row] == c.x;
  }
};

struct CountSrc {  // (count desc, key asc)
  const uint32_t* cnt;
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    const uint32_t c = cnt[i];
    e = TopEntry{(int64_t)c, 0, (int32_t)i, 0, 0};
    return c != 0;
  }
};

// Shared setup: nreg (256 bytes) from the nation table.
int build_nreg(gdb_ctx* ctx, const gdb_tpch_db* db, uint8_t* nreg) {
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

int top_k_arg(int k) {
  if (k < 0 || k > 100) return fail(GDB_E_INVALID, "top_k must lie in 0..100");
  return GDB_OK;
}

}  // namespace

extern "C" {

int gdb_q5(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q5_params* p, gdb_q5_result* out,
           gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *lo, *lk, *ld, *le, *oo, *oc, *od, *ck, *cn, *sk, *sn;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_suppkey", GDB_I32, &lk));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_custkey", GDB_I32, &oc));
  GDB_TRY(col(db->orders, "orders", "o_orderdate", GDB_I32, &od));
  GDB_TRY(col(db->customer, "customer", "c_custkey", GDB_I32, &ck));
  GDB_TRY(col(db->customer, "customer", "c_nationkey", GDB_I32, &cn));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows, nc = db->customer->nrows,
                 ns = db->supplier->nrows;
  uint64_t a, b, ospan, cspan, sspan;
  GDB_TRY(key_span(oo, no, &a));
  GDB_TRY(key_span(lo, nl, &b));
  ospan = std::max(a, b);
  GDB_TRY(key_span(ck, nc, &cspan));
  GDB_TRY(key_span(oc, no, &a));
  cspan = std::max(cspan, a);
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(lk, nl, &a));
  sspan = std::max(sspan, a);
  uint8_t *omap, *cmap, *smap, *nreg;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, ospan, (void**)&omap));
  GDB_TRY(arena(ctx, A_AUX, cspan + 256, (void**)&cmap));
  nreg = cmap + cspan;
  GDB_TRY(arena(ctx, A_AUX2, sspan, (void**)&smap));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(omap, NONE, ospan, s));
  GDB_TRY(fill(cmap, NONE, cspan, s));
  GDB_TRY(fill(smap, NONE, sspan, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  GDB_TRY(build_nreg(ctx, db, nreg));
  if (nc)
    cust_nation_in_region<<<blocks_for(ctx, cust_nation_in_region, nc), BLOCK, 0, s>>>(
        (const int32_t*)ck->dev, (const int32_t*)cn->dev, (uint32_t)nc, nreg, p->region, cmap);
  if (ns)
    nation_bytes<<<blocks_for(ctx, nation_bytes, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (uint32_t)ns, smap);
  if (no)
    orders_cust_byte<<<blocks_for(ctx, orders_cust_byte, quads(no)), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const int32_t*)oc->dev, (const int32_t*)od->dev, (uint32_t)no,
        p->date_lo, p->date_hi, cmap, (uint32_t)cspan, omap);
  GDB_HIP(hipGetLastError());
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lo, lk, ld, le}, [&](const void* const* d, uint32_t rows) {
    q5_scan<<<blocks_for(ctx, q5_scan, quads(rows)), BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int64_t*)d[3],
        rows, omap, smap, dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(acc_finish(ctx, dout, 0, 26, 1u << 25));
  GDB_HIP(hipMemcpyAsync(host, dout, 26 * 8, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem orderkey, suppkey, discount, extendedprice (20 B); orders
  // orderkey, custkey, orderdate (12 B); customer and supplier key + nation (8 B).
  // Mandatory in the scan: l_orderkey, 4 B per line.
  GDB_TRY(timing_finish(ctx, timing, 20 * nl + 12 * no + 8 * (nc + ns), 4 * nl, &rs));
  for (int i = 0; i < 25; ++i) out->revenue[i] = host[i];
  out->has = (uint32_t)host[25];
  return GDB_OK;
}

int gdb_q7(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q7_params* p, gdb_q7_result* out,
           gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *ls, *lo, *lk, *ld, *le, *oo, *oc, *ck, *cn, *sk, *sn;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_shipdate", GDB_I32, &ls));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_suppkey", GDB_I32, &lk));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_custkey", GDB_I32, &oc));
  GDB_TRY(col(db->customer, "customer", "c_custkey", GDB_I32, &ck));
  GDB_TRY(col(db->customer, "customer", "c_nationkey", GDB_I32, &cn));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows, nc = db->customer->nrows,
                 ns = db->supplier->nrows;
  uint64_t a, b, ospan, cspan, sspan;
  GDB_TRY(key_span(oo, no, &a));
  GDB_TRY(key_span(lo, nl, &b));
  ospan = std::max(a, b);
  GDB_TRY(key_span(ck, nc, &cspan));
  GDB_TRY(key_span(oc, no, &a));
  cspan = std::max(cspan, a);
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(lk, nl, &a));
  sspan = std::max(sspan, a);
  uint8_t *omap, *cmap, *smap;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, ospan, (void**)&omap));
  GDB_TRY(arena(ctx, A_AUX, cspan, (void**)&cmap));
  GDB_TRY(arena(ctx, A_AUX2, sspan, (void**)&smap));
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(omap, NONE, ospan, s));
  GDB_TRY(fill(cmap, NONE, cspan, s));
  GDB_TRY(fill(smap, NONE, sspan, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  if (nc)
    pair_side<<<blocks_for(ctx, pair_side, nc), BLOCK, 0, s>>>(
        (const int32_t*)ck->dev, (const int32_t*)cn->dev, (uint32_t)nc, p->nation1, p->nation2,
        cmap);
  if (ns)
    pair_side<<<blocks_for(ctx, pair_side, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (uint32_t)ns, p->nation1, p->nation2,
        smap);
  if (no)
    orders_side<<<blocks_for(ctx, orders_side, quads(no)), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const int32_t*)oc->dev, (uint32_t)no, cmap, (uint32_t)cspan,
        omap);
  GDB_HIP(hipGetLastError());
  // Years index 0..7 (1992..1999): clamp the window so year_index stays in range.
  const int32_t lo_d = std::max(p->date_lo, 0), hi_d = std::min(p->date_hi, 2921);
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {ls, lo, lk, ld, le},
                     [&](const void* const* d, uint32_t rows) {
    if (lo_d > hi_d) return GDB_OK;
    q7_scan<<<blocks_for(ctx, q7_scan, quads(rows)), BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
        (const int64_t*)d[4], rows, lo_d, hi_d, p->nation1 == p->nation2 ? 1 : 0, omap, smap,
        dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(acc_finish(ctx, dout, 0, 17, 1u << 16));
  GDB_HIP(hipMemcpyAsync(host, dout, 17 * 8, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem shipdate, orderkey, suppkey, discount, extendedprice (24 B);
  // orders key + custkey (8 B); customer and supplier key + nation (8 B). Mandatory
  // in the scan: l_shipdate, 4 B per line.
  GDB_TRY(timing_finish(ctx, timing, 24 * nl + 8 * (no + nc + ns), 4 * nl, &rs));
  for (int d = 0; d < 2; ++d) {
    for (int y = 0; y < 8; ++y) out->revenue[d][y] = host[d * 8 + y];
    out->has[d] = ((uint32_t)host[16] >> (8 * d)) & 0xFFu;
  }
  return GDB_OK;
}

int gdb_q8(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q8_params* p, gdb_q8_result* out,
           gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  const Column *lp, *lo, *lk, *ld, *le, *oo, *oc, *od, *ck, *cn, *sk, *sn, *pp, *pt;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_partkey", GDB_I32, &lp));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_suppkey", GDB_I32, &lk));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_custkey", GDB_I32, &oc));
  GDB_TRY(col(db->orders, "orders", "o_orderdate", GDB_I32, &od));
  GDB_TRY(col(db->customer, "customer", "c_custkey", GDB_I32, &ck));
  GDB_TRY(col(db->customer, "customer", "c_nationkey", GDB_I32, &cn));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  GDB_TRY(col(db->part, "part", "p_partkey", GDB_I32, &pp));
  GDB_TRY(col(db->part, "part", "p_type", GDB_U8, &pt));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows, nc = db->customer->nrows,
                 ns = db->supplier->nrows, np = db->part->nrows;
  uint64_t a, b, ospan, cspan, sspan, pspan;
  GDB_TRY(key_span(oo, no, &a));
  GDB_TRY(key_span(lo, nl, &b));
  ospan = std::max(a, b);
  GDB_TRY(key_span(ck, nc, &cspan));
  GDB_TRY(key_span(oc, no, &a));
  cspan = std::max(cspan, a);
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(lk, nl, &a));
  sspan = std::max(sspan, a);
  GDB_TRY(key_span(pp, np, &pspan));
  GDB_TRY(key_span(lp, nl, &a));
  pspan = std::max(pspan, a);
  const uint64_t words = (pspan + 31) / 32;
  const uint64_t words4 = (words + 3) & ~3ull;  // the folded filter starts 16-byte aligned
  uint8_t *omap, *cmap, *smap, *nreg;
  uint32_t *pbits, *pfold;
  int64_t *dout, *host;
  GDB_TRY(arena(ctx, A_MAP, ospan, (void**)&omap));
  GDB_TRY(arena(ctx, A_AUX, cspan + 256, (void**)&cmap));
  nreg = cmap + cspan;
  GDB_TRY(arena(ctx, A_AUX2, sspan, (void**)&smap));
  GDB_TRY(arena(ctx, A_BITMAP, (words4 + FOLD_WORDS) * 4, (void**)&pbits));
  pfold = pbits + words4;
  GDB_TRY(arena(ctx, A_OUT, ACC_BYTES, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(omap, NONE, ospan, s));
  GDB_TRY(fill(cmap, NONE, cspan, s));
  GDB_TRY(fill(smap, NONE, sspan, s));
  GDB_TRY(fill(pbits, 0, (words4 + FOLD_WORDS) * 4, s));
  GDB_TRY(fill(dout, 0, ACC_BYTES, s));
  GDB_TRY(build_nreg(ctx, db, nreg));
  if (np) {
    Mask4 m{{0, 0, 0, 0}};
    if (p->type_code >= 0 && p->type_code < 256) m.w[p->type_code >> 6] = 1ull << (p->type_code & 63);
    code_bits<<<blocks_for(ctx, code_bits, np), BLOCK, 0, s>>>(
        (const int32_t*)pp->dev, (const uint8_t*)pt->dev, (uint32_t)np, m, pbits, pfold);
  }
  if (nc)
    nation_bytes<<<blocks_for(ctx, nation_bytes, nc), BLOCK, 0, s>>>(
        (const int32_t*)ck->dev, (const int32_t*)cn->dev, (uint32_t)nc, cmap);
  if (ns)
    nation_bytes<<<blocks_for(ctx, nation_bytes, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (uint32_t)ns, smap);
  const int32_t lo_d = std::max(p->date_lo, 0), hi_d = std::min(p->date_hi, 2921);
  if (no)
    q8_orders<<<blocks_for(ctx, q8_orders, quads(no)), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const int32_t*)oc->dev, (const int32_t*)od->dev, (uint32_t)no,
        lo_d, hi_d, cmap, (uint32_t)cspan, nreg, p->region, omap);
  GDB_HIP(hipGetLastError());
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lp, lo, lk, ld, le},
                     [&](const void* const* d, uint32_t rows) {
    q8_scan<<<fold_blocks(ctx, (const void*)q8_scan, quads(rows)), FOLD_BLOCK, 0, s>>>(
        (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
        (const int64_t*)d[4], rows, pbits, pfold, omap, smap, p->nation, dout);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(acc_finish(ctx, dout, 0, 17, 1u << 16));
  GDB_HIP(hipMemcpyAsync(host, dout, 17 * 8, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem partkey, orderkey, suppkey, discount, extendedprice (24 B);
  // orders key, custkey, date (12 B); customer, supplier key + nation (8 B); part key
  // + type (5 B). Mandatory in the scan: l_partkey, 4 B per line.
  GDB_TRY(timing_finish(ctx, timing, 24 * nl + 12 * no + 8 * (nc + ns) + 5 * np, 4 * nl, &rs));
  for (int y = 0; y < 8; ++y) {
    out->nation_volume[y] = host[y];
    out->total_volume[y] = host[8 + y];
  }
  out->has = (uint32_t)host[16];
  return GDB_OK;
}

int gdb_q10(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q10_params* p, gdb_q10_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  GDB_TRY(top_k_arg(p->top_k));
  const Column *lf, *lo, *ld, *le, *oo, *oc, *od, *ck, *cn, *cb;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_returnflag", GDB_U8, &lf));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_discount", GDB_I32, &ld));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_extendedprice", GDB_I64, &le));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_custkey", GDB_I32, &oc));
  GDB_TRY(col(db->orders, "orders", "o_orderdate", GDB_I32, &od));
  GDB_TRY(col(db->customer, "customer", "c_custkey", GDB_I32, &ck));
  GDB_TRY(col(db->customer, "customer", "c_nationkey", GDB_I32, &cn));
  GDB_TRY(col(db->customer, "customer", "c_acctbal", GDB_I64, &cb));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows, nc = db->customer->nrows;
  uint64_t a, b, ospan, cspan;
  GDB_TRY(key_span(oo, no, &a));
  GDB_TRY(key_span(lo, nl, &b));
  ospan = std::max(a, b);
  GDB_TRY(key_span(ck, nc, &cspan));
  GDB_TRY(key_span(oc, no, &a));
  cspan = std::max(cspan, a);
  const int k = p->top_k;
  int32_t* ocust;
  uint32_t* owin;
  const uint64_t wwords = ospan / 32 + 1;
  int64_t* crev;
  uint8_t *seen, *rank;
  uint64_t* dout;
  char* host;
  GDB_TRY(arena(ctx, A_MAP, ospan * 4, (void**)&ocust));
  GDB_TRY(arena(ctx, A_BITMAP, wwords * 4, (void**)&owin));
  uint32_t* hits = nullptr;  // at most one per line of a chunk (sized per chunk)
  GDB_TRY(arena(ctx, A_AGG, cspan * 8, (void**)&crev));
  GDB_TRY(arena(ctx, A_AUX, cspan * 2, (void**)&seen));
  rank = seen + cspan;
  ThkState* thk;  // top-k by revenue histogram
  TopEntry* cand;
  GDB_TRY(thk_buffers(ctx, A_AUX2, &thk, &cand));
  // dout: [0] groups, [1] dummy, then k TopEntry, k balances, k nations
  const size_t out_bytes = 16 + (size_t)k * (sizeof(TopEntry) + 12) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, out_bytes, (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 2);
  int64_t* bal = (int64_t*)(fin + k);
  int32_t* nat = (int32_t*)(bal + k);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_TRY(fill(owin, 0, wwords * 4, s));
  GDB_TRY(fill(crev, 0, cspan * 8, s));
  GDB_TRY(fill(seen, 0, cspan, s));
  GDB_TRY(fill(rank, NONE, cspan, s));
  GDB_TRY(fill(dout, 0, out_bytes, s));
  if (no)
    q10_orders<<<blocks_for(ctx, q10_orders, quads(no)), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const int32_t*)oc->dev, (const int32_t*)od->dev, (uint32_t)no,
        p->date_lo, p->date_hi, owin, ocust);
  GDB_HIP(hipGetLastError());
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lf, lo, ld, le}, [&](const void* const* d,
                                                              uint32_t rows) -> int {
    const int grid = blocks_for(ctx, q10_probe, quads(rows));
    const uint64_t stride = (uint64_t)grid * BLOCK, nwaves = stride / 64;
    const uint32_t cap = (uint32_t)(256 * ((quads(rows) + stride - 1) / stride));
    GDB_TRY(arena(ctx, A_HT, std::max<uint64_t>(nwaves * cap, 1) * 4, (void**)&hits));
    q10_probe<<<grid, BLOCK, 0, s>>>((const uint8_t*)d[0], (const int32_t*)d[1],
                                     (const int32_t*)d[2], (const int64_t*)d[3], rows,
                                     (uint32_t)(uint8_t)p->returnflag, owin, ocust,
                                     (uint32_t)cspan, cap, hits, crev, seen);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  if (k > 0) {
    GDB_TRY(thk_run(ctx, KeyedRevSrc{crev, seen}, cspan, k, thk, cand, fin,
                    (unsigned long long*)dout));
    mark_ranks<<<1, 128, 0, s>>>(fin, k, rank);
    if (nc)
      gather_customers<<<blocks_for(ctx, gather_customers, nc), BLOCK, 0, s>>>(
          (const int32_t*)ck->dev, (const int32_t*)cn->dev, (const int64_t*)cb->dev,
          (uint32_t)nc, rank, (uint32_t)cspan, bal, nat);
    GDB_HIP(hipGetLastError());
  } else {
    GDB_TRY(thk_run(ctx, KeyedRevSrc{crev, seen}, cspan, 0, thk, cand, nullptr,
                    (unsigned long long*)dout));
  }
  GDB_HIP(hipMemcpyAsync(host, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem returnflag (1 B), orderkey, discount (4 B), price (8 B);
  // orders key, custkey, date (12 B); customer key, nation, acctbal (16 B).
  // Mandatory in q10_probe: l_returnflag and l_orderkey, 5 B per line (loaded together).
  GDB_TRY(timing_finish(ctx, timing, 17 * nl + 12 * no + 16 * nc, 5 * nl, &rs));
  const uint64_t* h = (const uint64_t*)host;
  const TopEntry* hf = (const TopEntry*)(h + 2);
  const int64_t* hb = (const int64_t*)(hf + k);
  const int32_t* hn = (const int32_t*)(hb + k);
  out->groups = (int64_t)h[0];
  out->nrows = 0;
  for (int i = 0; i < k; ++i) {
    if (hf[i].key == INT32_MAX) break;
    gdb_q10_row& r = out->rows[out->nrows++];
    r.custkey = hf[i].key;
    r.revenue = hf[i].rev;
    r.acctbal = hb[i];
    r.nationkey = hn[i];
  }
  return GDB_OK;
}

int gdb_q18(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q18_params* p, gdb_q18_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  GDB_TRY(top_k_arg(p->top_k));
  const Column *lo, *lq, *oo, *oc, *od, *ot;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_quantity", GDB_I32, &lq));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_custkey", GDB_I32, &oc));
  GDB_TRY(col(db->orders, "orders", "o_orderdate", GDB_I32, &od));
  GDB_TRY(col(db->orders, "orders", "o_totalprice", GDB_I64, &ot));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows;
  uint64_t a, b;
  GDB_TRY(key_span(oo, no, &a));
  GDB_TRY(key_span(lo, nl, &b));
  const uint64_t ospan = std::max(a, b);
  const int k = std::max(p->top_k, 1);
  // Clustered, resident lineitem: owners emit the orders over the threshold
  // (at most one per distinct l_orderkey), looked up in orders by binary search
  // when o_orderkey ascends strictly. Otherwise sums go to sq[orderkey] and the
  // top-k runs over every orders row.
  const bool clustered = lo->sorted && !db->lineitem->host;  // an order's rows are contiguous
  const bool emit = clustered && oo->strict && no > 0;
  const uint64_t cap = std::max<uint64_t>(1, std::min<uint64_t>(nl, ospan));
  int32_t* sq = nullptr;
  int2* cand = nullptr;
  ThkState* thk;  // top-k by totalprice histogram
  TopEntry* tcand;
  uint64_t* dout;
  char* host;
  if (emit) GDB_TRY(arena(ctx, A_MAP, cap * sizeof(int2), (void**)&cand));
  else GDB_TRY(arena(ctx, A_MAP, ospan * 4, (void**)&sq));
  GDB_TRY(thk_buffers(ctx, A_AUX2, &thk, &tcand));
  const size_t out_bytes = 16 + (size_t)k * (sizeof(TopEntry) + 8) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, out_bytes, (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 2);
  int32_t* extra = (int32_t*)(fin + k);
  uint32_t* ncand = (uint32_t*)(extra + 2 * k);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  if (!emit) GDB_TRY(fill(sq, 0, ospan * 4, s));
  GDB_TRY(fill(dout, 0, out_bytes, s));
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lo, lq}, [&](const void* const* d, uint32_t rows) {
    if (emit)
      q18_owner<true><<<blocks_for(ctx, q18_owner<true>, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], rows, nullptr, p->quantity, cand, ncand);
    else if (clustered)  // one plain store per order
      q18_owner<false><<<blocks_for(ctx, q18_owner<false>, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], rows, sq, 0, nullptr, nullptr);
    else
      q18_sums<<<blocks_for(ctx, q18_sums, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], rows, sq);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  if (emit)
    GDB_TRY(thk_run(ctx,
                    Q18CandSrc{cand, (const int32_t*)oo->dev, (const int32_t*)od->dev,
                               (const int64_t*)ot->dev, (uint32_t)no},
                    cap, k, thk, tcand, fin, (unsigned long long*)dout, ncand));
  else
    GDB_TRY(thk_run(ctx,
                    Q18Src{(const int32_t*)oo->dev, (const int32_t*)od->dev,
                           (const int64_t*)ot->dev, sq, p->quantity},
                    no, k, thk, tcand, fin, (unsigned long long*)dout));
  q18_gather<<<1, 128, 0, s>>>(fin, k, (const int32_t*)oc->dev, sq, extra);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipMemcpyAsync(host, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: l_orderkey, l_quantity (8 B per line, all read by q18_sums); orders
  // key, custkey, date (12 B) and totalprice (8 B).
  GDB_TRY(timing_finish(ctx, timing, 8 * nl + 20 * no, 8 * nl, &rs));
  const uint64_t* h = (const uint64_t*)host;
  const TopEntry* hf = (const TopEntry*)(h + 2);
  const int32_t* hx = (const int32_t*)(hf + k);
  out->orders_over = (int64_t)h[0];
  out->nrows = 0;
  for (int i = 0; i < p->top_k; ++i) {
    if (hf[i].key == INT32_MAX) break;
    gdb_q18_row& r = out->rows[out->nrows++];
    r.orderkey = hf[i].key;
    r.orderdate = hf[i].date;
    r.totalprice = hf[i].rev;
    r.custkey = hx[2 * i];
    r.sum_qty = hx[2 * i + 1];
  }
  return GDB_OK;
}

int gdb_q21(gdb_ctx* ctx, const gdb_tpch_db* db, const gdb_q21_params* p, gdb_q21_result* out,
            gdb_timing* timing) {
  GDB_TRY(args_ok(ctx, db, p, out));
  GDB_TRY(top_k_arg(p->top_k));
  const Column *lo, *lk, *lc, *lr, *oo, *os, *sk, *sn;
  GDB_TRY(scol(db->lineitem, "lineitem", "l_orderkey", GDB_I32, &lo));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_suppkey", GDB_I32, &lk));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_commitdate", GDB_I32, &lc));
  GDB_TRY(scol(db->lineitem, "lineitem", "l_receiptdate", GDB_I32, &lr));
  GDB_TRY(col(db->orders, "orders", "o_orderkey", GDB_I32, &oo));
  GDB_TRY(col(db->orders, "orders", "o_orderstatus", GDB_U8, &os));
  GDB_TRY(col(db->supplier, "supplier", "s_suppkey", GDB_I32, &sk));
  GDB_TRY(col(db->supplier, "supplier", "s_nationkey", GDB_I32, &sn));
  const uint64_t nl = db->lineitem->nrows, no = db->orders->nrows, ns = db->supplier->nrows;
  uint64_t a, b;
  GDB_TRY(key_span(oo, no, &a));
  GDB_TRY(key_span(lo, nl, &b));
  const uint64_t ospan = std::max(a, b);
  uint64_t sspan;
  GDB_TRY(key_span(sk, ns, &sspan));
  GDB_TRY(key_span(lk, nl, &a));
  sspan = std::max(sspan, a);
  const int k = std::max(p->top_k, 1);
  const bool clustered = lo->sorted && !db->lineitem->host;  // an order's rows are contiguous
  unsigned long long* state = nullptr;
  uint32_t* cnt;
  uint8_t *smap, *fst;
  ThkState* thk;  // top-k by count histogram
  TopEntry* tcand;
  uint64_t* dout;
  char* host;
  if (!clustered) GDB_TRY(arena(ctx, A_AGG, ospan * 8, (void**)&state));
  GDB_TRY(arena(ctx, A_BITMAP, ospan, (void**)&fst));
  GDB_TRY(arena(ctx, A_MAP, sspan * 5, (void**)&cnt));
  smap = (uint8_t*)(cnt + sspan);
  GDB_TRY(thk_buffers(ctx, A_AUX2, &thk, &tcand));
  const size_t out_bytes = 32 + (size_t)k * sizeof(TopEntry) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, out_bytes, (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 4);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  if (!clustered) GDB_TRY(fill(state, 0, ospan * 8, s));
  GDB_TRY(fill(fst, 0, ospan, s));
  GDB_TRY(fill(cnt, 0, sspan * 4, s));
  GDB_TRY(fill(smap, NONE, sspan, s));
  GDB_TRY(fill(dout, 0, out_bytes, s));
  if (no)
    order_status_bytes<<<blocks_for(ctx, order_status_bytes, no), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const uint8_t*)os->dev, (uint32_t)no,
        (uint32_t)(uint8_t)p->status, fst);
  if (ns)
    nation_bytes<<<blocks_for(ctx, nation_bytes, ns), BLOCK, 0, s>>>(
        (const int32_t*)sk->dev, (const int32_t*)sn->dev, (uint32_t)ns, smap);
  GDB_HIP(hipGetLastError());
  RunStats rs, rs2;
  if (clustered) {
    GDB_TRY(scan_table(ctx, db->lineitem, {lo, lk, lc, lr}, [&](const void* const* d, uint32_t rows) {
      q21_owner<<<blocks_for(ctx, q21_owner, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
          rows, fst, smap, p->nation, cnt);
      GDB_HIP(hipGetLastError());
      return GDB_OK;
    }, &rs));
  } else {
    // every order's state over all of lineitem first, then the counting pass
    // (a host table is streamed twice)
    GDB_TRY(scan_table(ctx, db->lineitem, {lo, lk, lc, lr}, [&](const void* const* d, uint32_t rows) {
      q21_track<<<blocks_for(ctx, q21_track, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
          rows, fst, state);
      GDB_HIP(hipGetLastError());
      return GDB_OK;
    }, &rs));
    auto count = [&](const void* const* d, uint32_t rows) -> int {
      q21_count_rows<<<blocks_for(ctx, q21_count_rows, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int32_t*)d[3],
          rows, fst, smap, p->nation, state, cnt);
      GDB_HIP(hipGetLastError());
      return GDB_OK;
    };
    if (!db->lineitem->host) {  // resident: launched directly, ev[1..2] keep timing pass 1
      const void* d[4] = {lo->dev, lk->dev, lc->dev, lr->dev};
      if (nl) GDB_TRY(count(d, (uint32_t)nl));
    } else {
      GDB_TRY(scan_table(ctx, db->lineitem, {lo, lk, lc, lr}, count, &rs2));
      rs.kernel_ms += rs2.kernel_ms;
      rs.h2d_bytes += rs2.h2d_bytes;
      rs.chunks += rs2.chunks;
    }
  }
  GDB_TRY(thk_run(ctx, CountSrc{cnt}, sspan, k, thk, tcand, fin, (unsigned long long*)dout));
  GDB_HIP(hipMemcpyAsync(host, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem orderkey, suppkey, commitdate, receiptdate (16 B); orders
  // key + status (5 B); supplier key + nation (8 B). Mandatory in the lineitem pass:
  // all four lineitem columns (16 B per line) for q21_owner, whose scan folds every
  // row; l_orderkey (4 B) for q21_track, which loads the rest only for orders of the
  // wanted status.
  GDB_TRY(timing_finish(ctx, timing, 16 * nl + 5 * no + 8 * ns, (clustered ? 16 : 4) * nl, &rs));
  const uint64_t* h = (const uint64_t*)host;
  const TopEntry* hf = (const TopEntry*)(h + 4);
  out->suppliers = (int64_t)h[0];
  out->nrows = 0;
  for (int i = 0; i < p->top_k; ++i) {
    if (hf[i].key == INT32_MAX) break;
    gdb_q21_row& r = out->rows[out->nrows++];
    r.suppkey = hf[i].key;
    r.pad = 0;
    r.numwait = hf[i].rev;
  }
  return GDB_OK;
}

}  // extern "C"
