// HBM bandwidth probe: a plain streaming read (sum of 16-byte words) and a
// plain copy over a buffer of known size. Their rocprofv3 FETCH_SIZE /
// WRITE_SIZE readings calibrate the counters used to measure the pipelines'
// HBM traffic, and their GB/s is the achievable ceiling beside the 8 TB/s
// published peak. No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "kcommon.h"

using namespace gdb;
using namespace gdbk;

namespace {

__global__ __launch_bounds__(256) void probe_read(const int4* __restrict__ src, uint64_t n16,
                                                  int64_t* __restrict__ partials) {
  __shared__ int64_t red[4];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t acc = 0;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    int4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = ld4nt(reinterpret_cast<const int32_t*>(src), (uint32_t)(i + u * stride));
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += (uint32_t)(v[u].x ^ v[u].y ^ v[u].z ^ v[u].w);
  }
  for (; i < n16; i += stride) {
    const int4 v = src[i];
    acc += (uint32_t)(v.x ^ v.y ^ v.z ^ v.w);
  }
  const int64_t s = block_sum((int64_t)acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void probe_copy(const int4* __restrict__ src,
                                                  int4* __restrict__ dst, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

}  // namespace

extern "C" int gdb_bandwidth_probe(gdb_ctx* ctx, uint64_t bytes, int reps, double* read_gbps,
                                   double* copy_gbps) {
  if (!ctx || reps < 1 || bytes < 4096 || bytes / 16 >= (1ull << 32))
    return fail(GDB_E_INVALID, "bad probe arguments");
  GDB_HIP(hipSetDevice(ctx->device));
  bytes &= ~uint64_t(4095);
  const uint64_t n16 = bytes / 16;
  char *a, *b;
  GDB_TRY(arena(ctx, A_STREAM0, bytes, (void**)&a));
  GDB_TRY(arena(ctx, A_STREAM1, bytes, (void**)&b));
  hipStream_t s = ctx->stream;
  GDB_HIP(hipMemsetAsync(a, 1, bytes, s));
  const int blocks = resident_blocks(ctx, (const void*)probe_read, 256, 0, 2);
  int64_t* partials;
  GDB_TRY(arena(ctx, A_PARTIALS, sizeof(int64_t) * blocks, (void**)&partials));
  float tr = 0.f, tc = 0.f;
  for (int r = 0; r <= reps; ++r) {  // rep 0 warms up
    GDB_HIP(hipEventRecord(ctx->ev[1], s));
    probe_read<<<blocks, 256, 0, s>>>((const int4*)a, n16, partials);
    GDB_HIP(hipEventRecord(ctx->ev[2], s));
    probe_copy<<<blocks, 256, 0, s>>>((const int4*)a, (int4*)b, n16);
    GDB_HIP(hipEventRecord(ctx->ev[3], s));
    GDB_HIP(hipEventSynchronize(ctx->ev[3]));
    float x = 0.f, y = 0.f;
    GDB_HIP(hipEventElapsedTime(&x, ctx->ev[1], ctx->ev[2]));
    GDB_HIP(hipEventElapsedTime(&y, ctx->ev[2], ctx->ev[3]));
    if (r) {
      tr += x;
      tc += y;
    }
  }
  if (read_gbps) *read_gbps = (double)bytes * reps / (tr / 1000.0) / 1e9;
  if (copy_gbps) *copy_gbps = 2.0 * (double)bytes * reps / (tc / 1000.0) / 1e9;
  return GDB_OK;
}

// Host-to-device copy rate of `bytes` from caller memory (pinned from
// gdb_host_alloc for the full PCIe rate), on the copy stream the streamed
// pipelines use; rep 0 warms up. The ceiling a streamed pipeline can reach.
extern "C" int gdb_h2d_probe(gdb_ctx* ctx, const void* host, uint64_t bytes, int reps,
                             double* gbps) {
  if (!ctx || !host || reps < 1 || bytes < 4096 || !gbps)
    return fail(GDB_E_INVALID, "bad probe arguments");
  GDB_HIP(hipSetDevice(ctx->device));
  char* d;
  GDB_TRY(arena(ctx, A_STREAM0, bytes, (void**)&d));
  hipStream_t s = ctx->copy_stream;
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  float total = 0.f;
  for (int r = 0; r <= reps; ++r) {
    GDB_HIP(hipEventRecord(ctx->ev[3], s));
    GDB_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, s));
    GDB_HIP(hipEventRecord(ctx->ev[4], s));
    GDB_HIP(hipEventSynchronize(ctx->ev[4]));
    float x = 0.f;
    GDB_HIP(hipEventElapsedTime(&x, ctx->ev[3], ctx->ev[4]));
    if (r) total += x;
  }
  *gbps = (double)bytes * reps / (total / 1000.0) / 1e9;
  return GDB_OK;
}
