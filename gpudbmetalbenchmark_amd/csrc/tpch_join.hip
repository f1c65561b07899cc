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
  __shared__ int64_t rev[32];
  __shared__ uint32_t hit;
  if (threadIdx.x < 32) rev[threadIdx.x] = 0;
  if (threadIdx.x == 0) hit = 0;
  __syncthreads();
  walk2(quads(n), [&](uint32_t g) { return ld4nt(ok, g); }, [&](uint32_t g, const int4& k) {
    uint32_t on[4];
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      on[j] = 4 * g + j < n ? omap[lane_of(k, j)] : NONE;
      m |= (uint32_t)(on[j] != NONE) << j;
    }
    if (!m) return;
    const int4 s = ld4nt(sk, g);
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      m2 |= (uint32_t)(((m >> j) & 1) && smap[lane_of(s, j)] == on[j]) << j;
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
  walk2(quads(n), [&](uint32_t g) { return Ld3{ld4nt(ok, g), ld4nt(oc, g), make_int4(0, 0, 0, 0)}; },
        [&](uint32_t g, const Ld3& l) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) continue;
      const uint32_t cv = (uint32_t)lane_of(l.b, j);
      omap[lane_of(l.a, j)] = cv < cspan ? cmap[cv] : NONE;
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
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cs[j] = (m >> j) & 1 ? omap[lane_of(k, j)] : NONE;
      m2 |= (uint32_t)(cs[j] != NONE) << j;
    }
    if (!m2) return;
    const int4 sp = ld4nt(sk, g);
    uint32_t ss[4];
    uint32_t m3 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ss[j] = (m2 >> j) & 1 ? smap[lane_of(sp, j)] : NONE;
      m3 |= (uint32_t)(ss[j] != NONE && (ss[j] != cs[j] || same)) << j;
    }
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
      m |= (uint32_t)((4 * g + j < n) && fold_test(lfold, key) && bit_test(pbits, key)) << j;
    }
    if (!m) return;
    const int4 k = ld4nt(ok, g);
    uint32_t y[4];
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      y[j] = (m >> j) & 1 ? omap[lane_of(k, j)] : NONE;
      m2 |= (uint32_t)(y[j] < 8) << j;
    }
    if (!m2) return;
    const int4 sp = ld4nt(sk, g), d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((m2 >> j) & 1)) continue;
      const int64_t v = lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j));
      lds_add_i64(&vol[8 + y[j]], v);
      if (smap[lane_of(sp, j)] == nation) lds_add_i64(&vol[y[j]], v);
      atomicOr(&hit, 1u << y[j]);
    }
  });
  __syncthreads();
  if (threadIdx.x < 16 && vol[threadIdx.x]) add_i64(&out[threadIdx.x], vol[threadIdx.x]);
  if (threadIdx.x == 0 && hit) atomicOr(reinterpret_cast<unsigned int*>(out + 16), hit);
}

// ---- Q10 (reference kernels/q10.metal:9-54, src/q10.cpp) -----------------------

// ocust[orderkey] = custkey for orders dated in [lo, hi) (-1 otherwise).
__global__ __launch_bounds__(256) void q10_orders(const int32_t* __restrict__ ok,
                                                  const int32_t* __restrict__ oc,
                                                  const int32_t* __restrict__ od, uint32_t n,
                                                  int32_t lo, int32_t hi,
                                                  int32_t* __restrict__ ocust) {
  walk2(quads(n), [&](uint32_t g) { return Ld3{ld4nt(ok, g), ld4nt(oc, g), ld4nt(od, g)}; },
        [&](uint32_t g, const Ld3& l) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (4 * g + j >= n) continue;
      const int dv = lane_of(l.c, j);
      ocust[lane_of(l.a, j)] = (dv >= lo && dv < hi) ? lane_of(l.b, j) : -1;
    }
  });
}

__global__ __launch_bounds__(256) void q10_scan(const uint8_t* __restrict__ rflag,
                                                const int32_t* __restrict__ ok,
                                                const int32_t* __restrict__ disc,
                                                const int64_t* __restrict__ price, uint32_t n,
                                                uint32_t flag, const int32_t* __restrict__ ocust,
                                                uint32_t cspan, int64_t* __restrict__ crev,
                                                uint8_t* __restrict__ seen) {
  walk2(quads(n), [&](uint32_t g) { return ld_u8x4nt(rflag, g); }, [&](uint32_t g, uint32_t fw) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) m |= (uint32_t)((4 * g + j < n) & (byte_of(fw, j) == flag)) << j;
    if (!m) return;
    const int4 k = ld4nt(ok, g);
    int32_t cu[4];
    uint32_t m2 = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cu[j] = (m >> j) & 1 ? ocust[lane_of(k, j)] : -1;
      m2 |= (uint32_t)(cu[j] >= 0 && (uint32_t)cu[j] < cspan) << j;
    }
    if (!m2) return;
    const int4 d = ld4nt(disc, g);
    const longlong2 a = ld2nt(price, 2 * g), b = ld2nt(price, 2 * g + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((m2 >> j) & 1) {
        add_i64(&crev[cu[j]], lane_of(a, b, j) * (int64_t)(100 - lane_of(d, j)));
        seen[cu[j]] = 1;
      }
  });
}

struct KeyedRevSrc {  // (revenue desc, key asc) over a per-key revenue array
  const int64_t* rev;
  const uint8_t* seen;
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    if (!seen[i]) return false;
    e = TopEntry{rev[i], 0, (int32_t)i, 0, 0};
    return true;
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
    if (sq[key] <= thr) return false;
    e = TopEntry{tp[i], od[i], key, (int32_t)i, 0};
    return true;
  }
};

__global__ void q18_gather(const TopEntry* __restrict__ win, int k,
                           const int32_t* __restrict__ oc, const int32_t* __restrict__ sq,
                           int32_t* __restrict__ out) {
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    if (win[i].key == INT32_MAX) continue;
    out[2 * i] = oc[win[i].prio];
    out[2 * i + 1] = sq[win[i].key];
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
// When l_orderkey is non-decreasing, the rows of an order are contiguous. The lane
// whose group of 4 rows holds an order's first row owns the order and computes its
// result alone, so no per-order state is shared. The owner sees a 12-row window in
// registers: its own 4 rows and the next two lanes' 4 rows each (wave shuffles;
// TPC-H orders have at most 7 rows). Only rows past the window (the last two lanes
// of a wave, or longer orders) are read again from memory. Loops run wave-uniform
// so every lane takes part in the shuffles.

constexpr int WIN = 12;

// w[0..3] = this lane's rows, w[4..11] = the next two lanes' rows.
__device__ __forceinline__ void row_window(const int4& v, int* w) {
  w[0] = v.x;
  w[1] = v.y;
  w[2] = v.z;
  w[3] = v.w;
#pragma unroll
  for (int d = 1; d < WIN / 4; ++d) {
    w[4 * d + 0] = __shfl_down(v.x, d, 64);
    w[4 * d + 1] = __shfl_down(v.y, d, 64);
    w[4 * d + 2] = __shfl_down(v.z, d, 64);
    w[4 * d + 3] = __shfl_down(v.w, d, 64);
  }
}

// Bit j set when row 4g+j starts an order (rows >= n never do).
__device__ __forceinline__ uint32_t order_heads(const int32_t* __restrict__ ok, const int4& k,
                                                uint32_t g, uint32_t n, int lane) {
  const int up = __shfl_up(k.w, 1, 64);
  const int prev = lane > 0 ? up : (g > 0 && 4 * g <= n ? ok[4 * g - 1] : 0);
  uint32_t h = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t r = 4 * g + j;
    const int pj = j == 0 ? prev : lane_of(k, j - 1);
    h |= (uint32_t)(r < n && (r == 0 || pj != lane_of(k, j))) << j;
  }
  return h;
}

// Window rows this lane may read from registers (the last lanes see fewer).
__device__ __forceinline__ int window_rows(int lane) {
  return lane <= 63 - (WIN / 4 - 1) ? WIN : 4 * (64 - lane);
}

// Q21: per order of the wanted status, suppliers and late suppliers; the order's
// late lines count for their supplier when the order has a second supplier and
// every late line has the same supplier (of `nation`).
struct Q21Acc {
  uint32_t first = 0, lfirst = 0, lcount = 0;
  bool multi = false, lmulti = false;
  __device__ void add(int sv, int cv, int rv) {
    const uint32_t s1 = (uint32_t)sv + 1u;
    if (!first) first = s1;
    else multi |= first != s1;
    if (rv > cv) {
      if (!lfirst) lfirst = s1;
      else lmulti |= lfirst != s1;
      ++lcount;
    }
  }
};

__global__ __launch_bounds__(256) void q21_owner(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ sk,
                                                 const int32_t* __restrict__ commit,
                                                 const int32_t* __restrict__ receipt, uint32_t n,
                                                 const uint8_t* __restrict__ fst,
                                                 const uint8_t* __restrict__ smap, int32_t nation,
                                                 uint32_t* __restrict__ cnt) {
  const uint32_t nq = quads(n), stride = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  const int avail = window_rows(lane);
  for (uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; gw < nq; gw += stride) {
    const uint32_t g = gw + lane;
    const bool live = g < nq;
    const int4 z = make_int4(0, 0, 0, 0);
    const int4 k = live ? ld4nt(ok, g) : z;
    uint32_t heads = order_heads(ok, k, g, live ? n : 0, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (((heads >> j) & 1) && !fst[lane_of(k, j)]) heads &= ~(1u << j);
    // every lane takes part in the shuffles; only lanes with owned orders load more
    const bool need = __any(heads != 0) != 0;
    if (!need) continue;
    const int4 s = live ? ld4nt(sk, g) : z, c = live ? ld4nt(commit, g) : z,
               rc = live ? ld4nt(receipt, g) : z;
    int wk[WIN], ws[WIN], wc[WIN], wr[WIN];
    row_window(k, wk);
    row_window(s, ws);
    row_window(c, wc);
    row_window(rc, wr);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((heads >> j) & 1)) continue;
      const int key = wk[j];
      Q21Acc acc;
      acc.add(ws[j], wc[j], wr[j]);
      bool open = true;
      uint32_t r = 4 * g + j + 1;
#pragma unroll
      for (int i = j + 1; i < WIN; ++i) {
        if (!open) continue;
        if (i >= avail || 4 * g + i >= n) break;
        if (wk[i] != key) {
          open = false;
          continue;
        }
        acc.add(ws[i], wc[i], wr[i]);
        r = 4 * g + i + 1;
      }
      if (open)
        for (; r < n && ok[r] == key; ++r) acc.add(sk[r], commit[r], receipt[r]);
      if (acc.multi && acc.lcount && !acc.lmulti && smap[acc.lfirst - 1] == nation)
        atomicAdd(&cnt[acc.lfirst - 1], acc.lcount);
    }
  }
}

// Q18: the owner writes the order's quantity sum with a plain store.
__global__ __launch_bounds__(256) void q18_owner(const int32_t* __restrict__ ok,
                                                 const int32_t* __restrict__ qty, uint32_t n,
                                                 int32_t* __restrict__ sq) {
  const uint32_t nq = quads(n), stride = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  const int avail = window_rows(lane);
  for (uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; gw < nq; gw += stride) {
    const uint32_t g = gw + lane;
    const bool live = g < nq;
    const int4 z = make_int4(0, 0, 0, 0);
    const int4 k = live ? ld4nt(ok, g) : z, q = live ? ld4nt(qty, g) : z;
    const uint32_t heads = order_heads(ok, k, g, live ? n : 0, lane);
    int wk[WIN], wq[WIN];
    row_window(k, wk);
    row_window(q, wq);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!((heads >> j) & 1)) continue;
      const int key = wk[j];
      int sum = wq[j];
      bool open = true;
      uint32_t r = 4 * g + j + 1;
#pragma unroll
      for (int i = j + 1; i < WIN; ++i) {
        if (!open) continue;
        if (i >= avail || 4 * g + i >= n) break;
        if (wk[i] != key) {
          open = false;
          continue;
        }
        sum += wq[i];
        r = 4 * g + i + 1;
      }
      if (open)
        for (; r < n && ok[r] == key; ++r) sum += qty[r];
      sq[key] = sum;
    }
  }
}

struct CountSrc {  // (count desc, key asc)
  const uint32_t* cnt;
  __device__ bool operator()(uint64_t i, TopEntry& e) const {
    if (!cnt[i]) return false;
    e = TopEntry{(int64_t)cnt[i], 0, (int32_t)i, 0, 0};
    return true;
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
  GDB_TRY(arena(ctx, A_OUT, 256, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(hipMemsetAsync(omap, NONE, ospan, s));
  GDB_HIP(hipMemsetAsync(cmap, NONE, cspan, s));
  GDB_HIP(hipMemsetAsync(smap, NONE, sspan, s));
  GDB_HIP(hipMemsetAsync(dout, 0, 26 * 8, s));
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
  GDB_TRY(arena(ctx, A_OUT, 256, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(hipMemsetAsync(omap, NONE, ospan, s));
  GDB_HIP(hipMemsetAsync(cmap, NONE, cspan, s));
  GDB_HIP(hipMemsetAsync(smap, NONE, sspan, s));
  GDB_HIP(hipMemsetAsync(dout, 0, 17 * 8, s));
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
  GDB_TRY(arena(ctx, A_OUT, 256, (void**)&dout));
  GDB_TRY(pinned(ctx, 256, (void**)&host));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(hipMemsetAsync(omap, NONE, ospan, s));
  GDB_HIP(hipMemsetAsync(cmap, NONE, cspan, s));
  GDB_HIP(hipMemsetAsync(smap, NONE, sspan, s));
  GDB_HIP(hipMemsetAsync(pbits, 0, (words4 + FOLD_WORDS) * 4, s));
  GDB_HIP(hipMemsetAsync(dout, 0, 17 * 8, s));
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
  const int tblocks = 256;
  int32_t* ocust;
  int64_t* crev;
  uint8_t *seen, *rank;
  TopEntry* win;
  uint64_t* dout;
  char* host;
  GDB_TRY(arena(ctx, A_MAP, ospan * 4, (void**)&ocust));
  GDB_TRY(arena(ctx, A_AGG, cspan * 8, (void**)&crev));
  GDB_TRY(arena(ctx, A_AUX, cspan * 2, (void**)&seen));
  rank = seen + cspan;
  GDB_TRY(arena(ctx, A_AUX2, sizeof(TopEntry) * (size_t)(tblocks + 1) * std::max(k, 1),
                (void**)&win));
  // dout: [0] groups, [1] dummy, then k TopEntry, k balances, k nations
  const size_t out_bytes = 16 + (size_t)k * (sizeof(TopEntry) + 12) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, out_bytes, (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 2);
  int64_t* bal = (int64_t*)(fin + k);
  int32_t* nat = (int32_t*)(bal + k);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(hipMemsetAsync(ocust, 0xFF, ospan * 4, s));
  GDB_HIP(hipMemsetAsync(crev, 0, cspan * 8, s));
  GDB_HIP(hipMemsetAsync(seen, 0, cspan, s));
  GDB_HIP(hipMemsetAsync(rank, NONE, cspan, s));
  GDB_HIP(hipMemsetAsync(dout, 0, out_bytes, s));
  if (no)
    q10_orders<<<blocks_for(ctx, q10_orders, quads(no)), BLOCK, 0, s>>>(
        (const int32_t*)oo->dev, (const int32_t*)oc->dev, (const int32_t*)od->dev, (uint32_t)no,
        p->date_lo, p->date_hi, ocust);
  GDB_HIP(hipGetLastError());
  RunStats rs;
  GDB_TRY(scan_table(ctx, db->lineitem, {lf, lo, ld, le}, [&](const void* const* d, uint32_t rows) {
    q10_scan<<<blocks_for(ctx, q10_scan, quads(rows)), BLOCK, 0, s>>>(
        (const uint8_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2], (const int64_t*)d[3],
        rows, (uint32_t)(uint8_t)p->returnflag, ocust, (uint32_t)cspan, crev, seen);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  if (k > 0) {
    GDB_TRY(topk_run(ctx, KeyedRevSrc{crev, seen}, cspan, k, win, tblocks, fin,
                     (unsigned long long*)dout, (unsigned long long*)(dout + 1)));
    mark_ranks<<<1, 128, 0, s>>>(fin, k, rank);
    if (nc)
      gather_customers<<<blocks_for(ctx, gather_customers, nc), BLOCK, 0, s>>>(
          (const int32_t*)ck->dev, (const int32_t*)cn->dev, (const int64_t*)cb->dev,
          (uint32_t)nc, rank, (uint32_t)cspan, bal, nat);
    GDB_HIP(hipGetLastError());
  } else {
    GDB_TRY(topk_run(ctx, KeyedRevSrc{crev, seen}, cspan, 1, win, tblocks, win + tblocks,
                     (unsigned long long*)dout, (unsigned long long*)(dout + 1)));
  }
  GDB_HIP(hipMemcpyAsync(host, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem returnflag (1 B), orderkey, discount (4 B), price (8 B);
  // orders key, custkey, date (12 B); customer key, nation, acctbal (16 B).
  // Mandatory in the scan: l_returnflag, 1 B per line.
  GDB_TRY(timing_finish(ctx, timing, 17 * nl + 12 * no + 16 * nc, 1 * nl, &rs));
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
  const int tblocks = 256;
  int32_t* sq;
  TopEntry* win;
  uint64_t* dout;
  char* host;
  GDB_TRY(arena(ctx, A_MAP, ospan * 4, (void**)&sq));
  GDB_TRY(arena(ctx, A_AUX2, sizeof(TopEntry) * (size_t)tblocks * k, (void**)&win));
  const size_t out_bytes = 16 + (size_t)k * (sizeof(TopEntry) + 8) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, out_bytes, (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 2);
  int32_t* extra = (int32_t*)(fin + k);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  GDB_HIP(hipMemsetAsync(sq, 0, ospan * 4, s));
  GDB_HIP(hipMemsetAsync(dout, 0, out_bytes, s));
  RunStats rs;
  const bool clustered = lo->sorted && !db->lineitem->host;  // an order's rows are contiguous
  GDB_TRY(scan_table(ctx, db->lineitem, {lo, lq}, [&](const void* const* d, uint32_t rows) {
    if (clustered)  // one plain store per order
      q18_owner<<<blocks_for(ctx, q18_owner, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], rows, sq);
    else
      q18_sums<<<blocks_for(ctx, q18_sums, quads(rows)), BLOCK, 0, s>>>(
          (const int32_t*)d[0], (const int32_t*)d[1], rows, sq);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }, &rs));
  GDB_TRY(topk_run(ctx,
                   Q18Src{(const int32_t*)oo->dev, (const int32_t*)od->dev,
                          (const int64_t*)ot->dev, sq, p->quantity},
                   no, k, win, tblocks, fin, (unsigned long long*)dout,
                   (unsigned long long*)(dout + 1)));
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
  const int tblocks = 256;
  const bool clustered = lo->sorted && !db->lineitem->host;  // an order's rows are contiguous
  unsigned long long* state = nullptr;
  uint32_t* cnt;
  uint8_t *smap, *fst;
  TopEntry* win;
  uint64_t* dout;
  char* host;
  if (!clustered) GDB_TRY(arena(ctx, A_AGG, ospan * 8, (void**)&state));
  GDB_TRY(arena(ctx, A_BITMAP, ospan, (void**)&fst));
  GDB_TRY(arena(ctx, A_MAP, sspan * 5, (void**)&cnt));
  smap = (uint8_t*)(cnt + sspan);
  GDB_TRY(arena(ctx, A_AUX2, sizeof(TopEntry) * (size_t)tblocks * k, (void**)&win));
  const size_t out_bytes = 32 + (size_t)k * sizeof(TopEntry) + 64;
  GDB_TRY(arena(ctx, A_OUT, out_bytes, (void**)&dout));
  GDB_TRY(pinned(ctx, out_bytes, (void**)&host));
  TopEntry* fin = (TopEntry*)(dout + 4);
  hipStream_t s = ctx->stream;
  GDB_HIP(hipEventRecord(ctx->ev[0], s));
  if (!clustered) GDB_HIP(hipMemsetAsync(state, 0, ospan * 8, s));
  GDB_HIP(hipMemsetAsync(fst, 0, ospan, s));
  GDB_HIP(hipMemsetAsync(cnt, 0, sspan * 4, s));
  GDB_HIP(hipMemsetAsync(smap, NONE, sspan, s));
  GDB_HIP(hipMemsetAsync(dout, 0, out_bytes, s));
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
  GDB_TRY(topk_run(ctx, CountSrc{cnt}, sspan, k, win, tblocks, fin, (unsigned long long*)dout,
                   (unsigned long long*)(dout + 1)));
  GDB_HIP(hipMemcpyAsync(host, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GDB_HIP(hipEventRecord(ctx->ev[5], s));
  // SQL columns: lineitem orderkey, suppkey, commitdate, receiptdate (16 B); orders
  // key + status (5 B); supplier key + nation (8 B). Mandatory in the lineitem pass:
  // l_orderkey, 4 B per line.
  GDB_TRY(timing_finish(ctx, timing, 16 * nl + 5 * no + 8 * ns, 4 * nl, &rs));
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
