// Chunked streaming of host tables through two device slots (the reference's
// chunkedStreamLoop and ChunkConfig, src/infra.h:37-57, :805-869).
//
// The reference overlaps CPU .tbl parsing of chunk i+1 with the GPU work on
// chunk i and waits on the CPU for every chunk. Here the host data is already
// columnar (ideally pinned), so the overlap is PCIe copy vs. kernels: chunk
// i+1 is copied on the copy stream while the engine stream runs chunk i, with
// events in both directions (copy done -> compute may start; compute done ->
// the slot may be overwritten) and no host wait until the pipeline finishes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

namespace gdb {

int require_resident(const gdb_table* t, const char* what) {
  if (t && t->host)
    return fail(GDB_E_UNSUPPORTED, std::string(what) + " must be resident (gdb_table_create)");
  return GDB_OK;
}

static uint64_t chunk_rows_for(gdb_ctx* ctx, uint64_t row_bytes, uint64_t n) {
  uint64_t rows = ctx->chunk_rows;
  if (!rows) {
    const uint64_t budget = ctx->budget ? ctx->budget : (64ull << 30);
    rows = budget / 8 / 2 / std::max<uint64_t>(1, row_bytes);
    rows = std::min<uint64_t>(std::max<uint64_t>(rows, 1ull << 20), 64ull << 20);
  }
  rows = (rows + 1023) / 1024 * 1024;
  return std::min<uint64_t>(rows, std::max<uint64_t>(n, 1));
}

int run_over(gdb_ctx* ctx, const gdb_table* t, const Column* const* cols, int ncols,
             const Consume& fn, RunStats* rs) {
  RunStats local;
  RunStats& st = rs ? *rs : local;
  st = RunStats{};
  std::vector<const void*> dev(ncols);
  if (!t->host) {
    for (int i = 0; i < ncols; ++i) dev[i] = cols[i]->dev;
    GDB_HIP(kernel_mark(ctx, 1, ctx->stream));
    GDB_TRY(fn(dev.data(), (uint32_t)t->nrows));
    GDB_HIP(kernel_mark(ctx, 2, ctx->stream));
    return GDB_OK;
  }
  const uint64_t n = t->nrows;
  uint64_t row_bytes = 0;
  for (int i = 0; i < ncols; ++i) row_bytes += (uint64_t)cols[i]->width;
  const uint64_t chunk = chunk_rows_for(ctx, row_bytes, n);
  const uint64_t nchunks = n ? (n + chunk - 1) / chunk : 0;
  if (chunk >= (1ull << 32)) return fail(GDB_E_UNSUPPORTED, "chunk too large");
  // Slot layout: every column's chunk buffer, 256-byte aligned, padded by 64 bytes.
  std::vector<size_t> off(ncols);
  size_t slot_bytes = 0;
  for (int i = 0; i < ncols; ++i) {
    off[i] = slot_bytes;
    slot_bytes += ((chunk * cols[i]->width + 64) + 255) & ~size_t(255);
  }
  char* slot[2];
  GDB_TRY(arena(ctx, A_STREAM0, slot_bytes, (void**)&slot[0]));
  GDB_TRY(arena(ctx, A_STREAM1, slot_bytes, (void**)&slot[1]));
  while (ctx->chunk_ev.size() < 2 * nchunks) {
    hipEvent_t e;
    GDB_HIP(hipEventCreate(&e));
    ctx->chunk_ev.push_back(e);
  }
  hipEvent_t* copied = ctx->sev;      // [0], [1]
  hipEvent_t* freed = ctx->sev + 2;   // [2], [3]
  auto issue_copy = [&](uint64_t c) -> int {
    const int s = (int)(c & 1);
    if (c >= 2) GDB_HIP(hipStreamWaitEvent(ctx->copy_stream, freed[s], 0));
    const uint64_t r0 = c * chunk, rows = std::min(chunk, n - r0);
    for (int i = 0; i < ncols; ++i) {
      const size_t w = (size_t)cols[i]->width;
      GDB_HIP(hipMemcpyAsync(slot[s] + off[i], (const char*)cols[i]->host + r0 * w, rows * w,
                             hipMemcpyHostToDevice, ctx->copy_stream));
    }
    GDB_HIP(hipEventRecord(copied[s], ctx->copy_stream));
    st.h2d_bytes += rows * row_bytes;
    return GDB_OK;
  };
  // The copy stream must not start before work already queued on the engine
  // stream (e.g. a previous pipeline still reading the slots).
  GDB_HIP(hipEventRecord(freed[0], ctx->stream));
  GDB_HIP(hipStreamWaitEvent(ctx->copy_stream, freed[0], 0));
  if (nchunks) GDB_TRY(issue_copy(0));
  for (uint64_t c = 0; c < nchunks; ++c) {
    const int s = (int)(c & 1);
    if (c + 1 < nchunks) GDB_TRY(issue_copy(c + 1));
    GDB_HIP(hipStreamWaitEvent(ctx->stream, copied[s], 0));
    for (int i = 0; i < ncols; ++i) dev[i] = slot[s] + off[i];
    const uint64_t rows = std::min(chunk, n - c * chunk);
    GDB_HIP(hipEventRecord(ctx->chunk_ev[2 * c], ctx->stream));
    GDB_TRY(fn(dev.data(), (uint32_t)rows));
    GDB_HIP(hipEventRecord(ctx->chunk_ev[2 * c + 1], ctx->stream));
    GDB_HIP(hipEventRecord(freed[s], ctx->stream));
  }
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  for (uint64_t c = 0; c < nchunks; ++c) {
    float ms = 0.f;
    GDB_HIP(hipEventElapsedTime(&ms, ctx->chunk_ev[2 * c], ctx->chunk_ev[2 * c + 1]));
    st.kernel_ms += ms;
  }
  st.chunks = (int)nchunks;
  return GDB_OK;
}

}  // namespace gdb
