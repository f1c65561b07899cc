// Device helpers shared by the gfx950 kernels: wave64 reductions and scans,
// block reductions, and the partial-sum finaliser.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gdbk {

constexpr int WAVE = 64;

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor((long long)v, off, WAVE);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    uint32_t o = __shfl_up(v, off, WAVE);
    if (lane >= off) v += o;
  }
  return v;
}

// Sum of `v` over the block; the result is valid in thread 0. `red` holds one
// slot per wave.
__device__ __forceinline__ int64_t block_sum(int64_t v, int64_t* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  if (lane == 0) red[w] = v;
  __syncthreads();
  int64_t s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    for (int i = 0; i < nw; ++i) s += red[i];
  }
  __syncthreads();
  return s;
}

// Global-load helpers: dwordx4 vector loads of 4 int32 rows / 2 int64 rows.
// The *nt variants are nontemporal (streamed once, not kept in the caches):
// on gfx950 a grid-stride read with 4 such loads in flight per lane reaches
// ~88% of HBM peak, against ~74% for plain loads.
typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef long long v2l_t __attribute__((ext_vector_type(2)));
// Orders a wave's own LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int4 ld4nt(const int32_t* p, uint32_t vec) {
  const v4i_t v = __builtin_nontemporal_load(reinterpret_cast<const v4i_t*>(p) + vec);
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ longlong2 ld2nt(const int64_t* p, uint32_t idx2) {
  const v2l_t v = __builtin_nontemporal_load(reinterpret_cast<const v2l_t*>(p) + idx2);
  longlong2 r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
__device__ __forceinline__ uint32_t ld_u8x4nt(const uint8_t* p, uint32_t vec) {
  return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p) + vec);
}
__device__ __forceinline__ int4 ld4(const int32_t* p, uint32_t vec) {
  return reinterpret_cast<const int4*>(p)[vec];
}
__device__ __forceinline__ longlong2 ld2(const int64_t* p, uint32_t idx2) {
  return reinterpret_cast<const longlong2*>(p)[idx2];
}
__device__ __forceinline__ uint32_t ld_u8x4(const uint8_t* p, uint32_t vec) {
  return reinterpret_cast<const uint32_t*>(p)[vec];
}

// Exclusive prefix sum of v over the block (blockDim a multiple of 64, at most
// 1024); *total receives the block sum. `scratch` holds 17 words of LDS. Every
// thread of the block must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* scratch,
                                                    uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) scratch[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < nw; ++i) {
      const uint32_t t = scratch[i];
      scratch[i] = run;
      run += t;
    }
    scratch[16] = run;
  }
  __syncthreads();
  const uint32_t r = scratch[w] + incl - v;
  *total = scratch[16];
  __syncthreads();
  return r;
}

// True in every thread of the workgroup that finishes last among the grid's
// workgroups (counted on `done`, zero before the launch); all earlier
// workgroups' global writes are visible to it afterwards. Every workgroup
// passes through one atomic on `done`, so this is for grids of a few hundred
// workgroups at most (same-address atomics serialise at tens of ns each).
__device__ __forceinline__ bool last_block(uint32_t* done) {
  __shared__ uint32_t last_sh;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last_sh = atomicAdd(done, 1u) == gridDim.x - 1;
  __syncthreads();
  const bool last = last_sh != 0;
  if (last) __threadfence();
  return last;
}

// Monotone bucket of an int64 in [0, 4096): log-linear with 32 sub-buckets per
// power of two (about 3% wide), negatives mirrored below the non-negatives.
__device__ __forceinline__ uint32_t log_bucket_mag(uint64_t x) {
  if (x == 0) return 0;
  const int e = 63 - __clzll((long long)x);
  const uint32_t m = e >= 5 ? (uint32_t)(x >> (e - 5)) & 31u : (uint32_t)(x << (5 - e)) & 31u;
  return 1u + (uint32_t)e * 32u + m;  // 1 .. 2048
}
__device__ __forceinline__ uint32_t value_bucket(int64_t v) {
  if (v >= 0) return 2047u + log_bucket_mag((uint64_t)v);  // 2047 .. 4095
  const uint64_t mag = v == INT64_MIN ? (1ull << 63) : (uint64_t)(-v);
  return 2047u - min(2047u, log_bucket_mag(mag));          // 0 .. 2046
}

// Finaliser: block k of the grid sums column k of a k-major partials array
// ([K][nparts]) into out[k] (ACC = 1: adds to out[k], for streamed chunks),
// and, when host_out is set, also stores it at host_out[k] (mapped pinned
// host memory: the result reaches the host with no readback copy).
// Launch with 256 threads.
template <int ACC = 0>
__global__ __launch_bounds__(256) void sum_columns(const int64_t* __restrict__ partials,
                                                   uint32_t nparts, int64_t* __restrict__ out,
                                                   int64_t* host_out = nullptr) {
  __shared__ int64_t red[4];
  const int64_t* col = partials + (size_t)blockIdx.x * nparts;
  int64_t s = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) s += col[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    const int64_t v = ACC ? out[blockIdx.x] + s : s;
    out[blockIdx.x] = v;
    if (host_out) host_out[blockIdx.x] = v;
  }
}

// Result readback without a copy engine: one workgroup copies `words` words
// (W: 4 or 8 bytes) from device memory to mapped pinned host memory.
// Launch <<<1, 256>>>.
template <class W>
__global__ __launch_bounds__(256) void publish_words(const W* __restrict__ src, uint32_t words,
                                                     W* dst) {
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

}  // namespace gdbk
