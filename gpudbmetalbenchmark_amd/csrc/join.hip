// Equi-join pair count, the `join` primitive (reference hash_join_build /
// hash_join_probe, kernels/microbenchmarks.metal:224-293; host
// src/microbenchmarks.cpp:135-251).
//
// Build: open-addressing table (key, multiplicity) with capacity nextpow2(2*nb),
// a locality-preserving fold hash, linear probing, CAS insert. Probe: dwordx4 loads of 4 probe
// keys per lane, each matched key adds its multiplicity; per-block partial
// counts are summed by the finaliser. Keys may repeat on both sides.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "kcommon.h"

using namespace gdb;
using namespace gdbk;

namespace {

constexpr int32_t EMPTY = INT32_MIN;

// Keys below the capacity map to themselves and higher bits are folded in, so
// keys arriving in (near) key order touch neighbouring slots and the CAS
// inserts and probes coalesce (the reference hashes by key modulo size,
// kernels/microbenchmarks.metal:240).
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

}  // namespace

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
  GDB_HIP(hipSetDevice(ctx->device));
  const uint32_t nb = (uint32_t)build->nrows, np = (uint32_t)probe->nrows;
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
  const int B = 256;
  int blocks = resident_blocks(ctx, (const void*)join_probe, B, 0, ctx->tune[T_JOIN].blocks_per_cu);
  blocks = std::max(1, std::min<int>(blocks, (int)((np / 4 + B - 1) / B)));
  int64_t *partials, *dout, *host;
  GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * blocks, (void**)&partials));
  GDB_TRY(arena(ctx, A_OUT, 64, (void**)&dout));
  GDB_TRY(pinned(ctx, 64, (void**)&host));
  hipStream_t st = ctx->stream;
  const int fill = ctx->info.cu_count * 4;
  GDB_HIP(hipEventRecord(ctx->ev[0], st));
  join_clear<<<fill, B, 0, st>>>(keys, mult, cap);
  if (nb) join_build<<<fill, B, 0, st>>>((const int32_t*)cb->dev, nb, keys, mult, cap, shift);
  GDB_HIP(hipEventRecord(ctx->ev[1], st));
  join_probe<<<blocks, B, 0, st>>>((const int32_t*)cp->dev, np, keys, mult, cap, shift, partials);
  GDB_HIP(hipGetLastError());
  GDB_HIP(hipEventRecord(ctx->ev[2], st));
  sum_columns<<<1, 256, 0, st>>>(partials, (uint32_t)blocks, dout);
  GDB_HIP(hipMemcpyAsync(host, dout, 8, hipMemcpyDeviceToHost, st));
  GDB_HIP(hipEventRecord(ctx->ev[5], st));
  GDB_TRY(timing_finish(ctx, timing, 4ull * nb + 4ull * np, 4ull * np));
  *matches = (uint64_t)host[0];
  return GDB_OK;
}
