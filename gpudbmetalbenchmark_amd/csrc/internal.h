// Engine-internal declarations shared by the runtime and the pipeline files.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gdb.h"

namespace gdb {

void set_error(const std::string& msg);

#define GDB_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::gdb::set_error(std::string(#call) + ": " + hipGetErrorString(e_));              \
      return GDB_E_HIP;                                                                 \
    }                                                                                   \
  } while (0)

#define GDB_TRY(expr)             \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != GDB_OK) return rc_; \
  } while (0)

int fail(int code, const std::string& msg);

struct Column {
  std::string name;
  int type = 0;
  int width = 0;       // bytes per value
  void* dev = nullptr;
  size_t bytes = 0;
  int64_t minv = 0, maxv = -1;  // value range of integer columns (empty: max < min)
  bool sorted = false;          // int32 column uploaded in non-decreasing order
  bool strict = false;          // ... and strictly increasing (unique keys)
  const void* host = nullptr;   // host tables: caller memory streamed in chunks
};

// Launch shape chosen per kernel; the autotuner may overwrite the defaults.
struct Tuning {
  int block = 256;          // threads per workgroup
  int blocks_per_cu = 0;    // 0: occupancy-derived residency
  int unroll = 2;           // vector groups per thread per loop trip
  int radix_bits = 0;       // join: partition bits (0 = size-derived, -1 = one global table)
};

// Device scratch that grows on demand and is reused across calls (partials,
// maps, hash tables). Owned by the context; released by gdb_close.
struct Arena {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace gdb

struct gdb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev[6] = {};
  // ev[1] / ev[2] bracket a pipeline's dominant kernel only while this is
  // set: each timing event recorded between two kernels leaves the GPU idle
  // for a few microseconds.
  bool kernel_timing = true;
  gdb_device_info info{};
  uint64_t budget = 0;
  uint64_t used = 0;
  int fail_allocs = 0;              // test hook: allocations still to fail
  int fail_skip = 0;                // test hook: allocations to pass before those
  std::mutex mu;
  std::vector<gdb::Arena> arenas;  // indexed by purpose slot
  void* host_pinned = nullptr;     // small pinned staging for results
  size_t host_pinned_bytes = 0;
  void* host_result = nullptr;     // coherent pinned memory the final kernels write
  void* host_result_dev = nullptr; // its address on the device
  size_t host_result_bytes = 0;
  gdb::Tuning tune[16];
  uint64_t chunk_rows = 0;         // streamed chunk size, 0 = adaptive
  uint64_t q20_pairs = 0;          // Q20: (part, supplier) pairs of the last run, sizes the next table
  hipEvent_t sev[4] = {};          // copy-done / slot-free events of the two slots
  std::vector<hipEvent_t> chunk_ev;  // per-chunk compute start/end
  // Occupancy query results by (kernel, threads, dynamic LDS): the runtime's
  // query costs microseconds and pipelines ask on every call.
  std::mutex occ_mu;
  std::map<std::tuple<const void*, int, size_t>, int> occ;
  std::map<const void*, size_t> lds_attr;  // dynamic LDS granted per kernel (occ_mu)
};

struct gdb_table {
  gdb_ctx* ctx = nullptr;
  uint64_t nrows = 0;
  bool host = false;  // columns live in host memory and are streamed
  std::vector<gdb::Column> cols;
};

namespace gdb {

// Device allocation accounted against the context's HBM budget.
int dev_alloc(gdb_ctx* ctx, void** p, size_t bytes);
void dev_free(gdb_ctx* ctx, void* p, size_t bytes);
// Scratch slot `slot` of at least `bytes` (contents undefined).
int arena(gdb_ctx* ctx, int slot, size_t bytes, void** out);
// Sets `bytes` bytes at `p` to `byte` on stream `s` (a fill kernel, no memset node).
int fill(void* p, int byte, size_t bytes, hipStream_t s);
// Pinned host staging of at least `bytes`.
int pinned(gdb_ctx* ctx, size_t bytes, void** out);
// Coherent pinned host memory of at least `bytes`, mapped into the device's
// address space: a pipeline's final kernel stores its result there (at
// `*dev`), and the host reads it (at `*host`) once the step's closing event
// has completed, with no separate readback copy.
int result_buf(gdb_ctx* ctx, size_t bytes, void** host, void** dev);
// Column lookup with type check; sets the error message on failure.
int column(const gdb_table* t, const char* name, int type, const Column** out);

// Largest magnitude of an integer column's values (1 for an empty column).
inline uint64_t magnitude(const Column* c) {
  if (c->maxv < c->minv) return 1;
  const uint64_t a = c->minv < 0 ? (uint64_t)0 - (uint64_t)c->minv : (uint64_t)c->minv;
  const uint64_t b = c->maxv < 0 ? (uint64_t)0 - (uint64_t)c->maxv : (uint64_t)c->maxv;
  return a > b ? a : b;
}
// GDB_OK if n terms each of magnitude at most `term` (a product of column
// magnitudes, given as a 128-bit value) sum without leaving int64; else a
// GDB_E_UNSUPPORTED naming `what`. The pipelines sum money in int64, so a
// table whose bound fails is refused instead of returning a wrapped sum.
int sum_fits(uint64_t n, unsigned __int128 term, const char* what);
// Number of workgroups that keep every CU busy for `kernel` at `block` threads
// and `lds` bytes of dynamic LDS.
int resident_blocks(gdb_ctx* ctx, const void* kernel, int block, size_t lds, int per_cu_cap);
// Lets `kernel` launch with `bytes` of dynamic LDS: hipFuncSetAttribute only
// when the context has not already granted that much to it.
int allow_dynamic_lds(gdb_ctx* ctx, const void* kernel, size_t bytes);

enum ArenaSlot {
  A_PARTIALS = 0, A_BITMAP = 1, A_MAP = 2, A_HT = 3, A_AGG = 4, A_AUX = 5, A_OUT = 6,
  A_AUX2 = 7, A_STREAM0 = 8, A_STREAM1 = 9, A_Q3BYTE = 10, A_Q3MAP = 11, A_JOINPART = 12, A_JOINHIST = 13,
  A_T0 = 14, A_T1 = 15, A_ACC = 16, A_NSLOTS = 17
};
enum TuneSlot { T_Q6 = 0, T_Q1 = 1, T_SELECT = 2, T_SUM = 3, T_Q3 = 4, T_Q9 = 5, T_Q13 = 6, T_JOIN = 7 };

// Launches of a pipeline's consuming stage over one batch of rows: `dev[i]`
// is the device address of column i for these rows.
using Consume = std::function<int(const void* const* dev, uint32_t rows)>;

struct RunStats {
  double kernel_ms = 0;   // streamed: summed per-chunk compute time
  uint64_t h2d_bytes = 0;
  int chunks = 0;         // 0: the table was resident
};

// Runs `fn` once over a resident table (bracketed by ev[1], ev[2]) or, for a
// host table, chunk by chunk: two device slots per column, copies on the copy
// stream overlapped with compute on the engine stream, event hand-off both ways.
// Partition bits the join derives for a build side of nb keys.
int join_derived_bits(uint64_t nb);

int run_over(gdb_ctx* ctx, const gdb_table* t, const Column* const* cols, int ncols,
             const Consume& fn, RunStats* rs);
// Looks up, once per device (that device current), the device address of the
// scratch array gdb_q6_noalloc uses instead of allocations.
int noalloc_scratch_init(int device);
// Rejects host tables where a pipeline needs the table resident.
int require_resident(const gdb_table* t, const char* what);

// Records ev[i] (1 or 2, around a pipeline's dominant kernel) on stream s
// when kernel timing is on.
inline hipError_t kernel_mark(gdb_ctx* ctx, int i, hipStream_t s) {
  return ctx->kernel_timing ? hipEventRecord(ctx->ev[i], s) : hipSuccess;
}

// Event helpers around a pipeline (ev[0], ev[5]) and its dominant kernel (ev[1], ev[2]).
int timing_finish(gdb_ctx* ctx, gdb_timing* t, uint64_t bytes, uint64_t main_bytes,
                  const RunStats* rs = nullptr);

}  // namespace gdb
