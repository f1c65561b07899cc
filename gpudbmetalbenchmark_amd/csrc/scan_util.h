// Device-wide exclusive prefix sum of uint32 (three launches on stream `s`).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gdb.h"

namespace gdb {

// out[i] = sum of in[0, i); *total (device) = sum of all. `tile_sums` must hold
// scan_tile_sums_needed(n) words. `in` and `out` may alias.
int exclusive_scan_u32(gdb_ctx* ctx, const uint32_t* in, uint32_t n, uint32_t* out,
                       uint32_t* total, uint32_t* tile_sums, hipStream_t s);
uint32_t scan_tile_sums_needed(uint32_t n);

}  // namespace gdb
