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
__device__ __forceinline__ int4 ld4(const int32_t* p, uint32_t vec) {
  return reinterpret_cast<const int4*>(p)[vec];
}
__device__ __forceinline__ longlong2 ld2(const int64_t* p, uint32_t idx2) {
  return reinterpret_cast<const longlong2*>(p)[idx2];
}
__device__ __forceinline__ uint32_t ld_u8x4(const uint8_t* p, uint32_t vec) {
  return reinterpret_cast<const uint32_t*>(p)[vec];
}

// Finaliser: block k of the grid sums column k of a k-major partials array
// ([K][nparts]) into out[k] (ACC = 1: adds to out[k], for streamed chunks).
// Launch with 256 threads.
template <int ACC = 0>
__global__ __launch_bounds__(256) void sum_columns(const int64_t* __restrict__ partials,
                                                   uint32_t nparts, int64_t* __restrict__ out) {
  __shared__ int64_t red[4];
  const int64_t* col = partials + (size_t)blockIdx.x * nparts;
  int64_t s = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) s += col[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = ACC ? out[blockIdx.x] + s : s;
}

}  // namespace gdbk
