// Device-wide exclusive prefix sum of uint32 values in three launches:
// per-tile scans (4096 values per 1024-thread block) with tile sums, a scan of
// the tile sums, and the tile offsets added back. Used to turn per-block,
// per-bucket counts into scatter offsets without global atomics.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "kcommon.h"
#include "scan_util.h"

using namespace gdbk;

namespace {

constexpr uint32_t TILE = 4096;  // 1024 threads x 4 values

__device__ uint32_t block_scan_1024(uint32_t v, uint32_t* wsum, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (threadIdx.x < 64) {
    const uint32_t x = threadIdx.x < (blockDim.x >> 6) ? wsum[threadIdx.x] : 0u;
    const uint32_t xi = wave_incl_scan(x);
    if (threadIdx.x < (blockDim.x >> 6)) wsum[threadIdx.x] = xi - x;
    if (threadIdx.x == 63) wsum[16] = xi;
  }
  __syncthreads();
  const uint32_t r = wsum[w] + incl - v;
  *total = wsum[16];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(1024) void scan_tiles(const uint32_t* __restrict__ in, uint32_t n,
                                                   uint32_t* __restrict__ out,
                                                   uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t wsum[17];
  const uint32_t base = blockIdx.x * TILE + threadIdx.x * 4;
  uint32_t v[4];
  for (int j = 0; j < 4; ++j) v[j] = base + j < n ? in[base + j] : 0u;
  const uint32_t s = v[0] + v[1] + v[2] + v[3];
  uint32_t total;
  uint32_t run = block_scan_1024(s, wsum, &total);
  for (int j = 0; j < 4; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block: exclusive scan of the tile sums in place; *grand = their total.
__global__ __launch_bounds__(1024) void scan_sums(uint32_t* __restrict__ sums, uint32_t m,
                                                  uint32_t* __restrict__ grand) {
  __shared__ uint32_t wsum[17];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < m; b += blockDim.x) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = i < m ? sums[i] : 0u;
    uint32_t total;
    const uint32_t pre = block_scan_1024(v, wsum, &total);
    if (i < m) sums[i] = carry + pre;
    __syncthreads();
    if (threadIdx.x == 0) carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0 && grand) *grand = carry;
}

__global__ __launch_bounds__(1024) void scan_add(uint32_t* __restrict__ out, uint32_t n,
                                                 const uint32_t* __restrict__ tile_offs) {
  const uint32_t base = blockIdx.x * TILE + threadIdx.x * 4;
  const uint32_t add = tile_offs[blockIdx.x];
  for (int j = 0; j < 4; ++j)
    if (base + j < n) out[base + j] += add;
}

}  // namespace

namespace gdb {

int exclusive_scan_u32(gdb_ctx* ctx, const uint32_t* in, uint32_t n, uint32_t* out,
                       uint32_t* total, uint32_t* tile_sums, hipStream_t s) {
  const uint32_t tiles = (n + TILE - 1) / TILE;
  if (tiles == 0) {
    scan_sums<<<1, 1024, 0, s>>>(tile_sums, 0, total);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  }
  scan_tiles<<<tiles, 1024, 0, s>>>(in, n, out, tile_sums);
  scan_sums<<<1, 1024, 0, s>>>(tile_sums, tiles, total);
  scan_add<<<tiles, 1024, 0, s>>>(out, n, tile_sums);
  GDB_HIP(hipGetLastError());
  (void)ctx;
  return GDB_OK;
}

uint32_t scan_tile_sums_needed(uint32_t n) { return (n + TILE - 1) / TILE + 1; }

}  // namespace gdb
