// Context, resident tables, scratch arenas, errors and timing: the HIP layer
// that replaces the reference's Metal device/buffer/pipeline helpers
// (reference src/infra.h:177-189, :499-531; src/main.cpp:76-156).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace gdb {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int dev_alloc(gdb_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  if (ctx->fail_allocs > 0) {
    if (ctx->fail_skip > 0) {
      --ctx->fail_skip;
    } else {
      --ctx->fail_allocs;
      return fail(GDB_E_NOMEM, "injected device allocation failure");
    }
  }
  if (ctx->budget && ctx->used + bytes > ctx->budget)
    return fail(GDB_E_NOMEM, "HBM budget exceeded: " + std::to_string(ctx->used + bytes) +
                                 " > " + std::to_string(ctx->budget) + " bytes");
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(GDB_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  ctx->used += bytes;
  return GDB_OK;
}

void dev_free(gdb_ctx* ctx, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  if (bytes == 0) bytes = 256;
  ctx->used -= std::min<uint64_t>(ctx->used, bytes);
}

int arena(gdb_ctx* ctx, int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= A_NSLOTS) return fail(GDB_E_INVALID, "bad arena slot");
  Arena& a = ctx->arenas[slot];
  if (a.bytes < bytes) {
    if (a.ptr) {
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      dev_free(ctx, a.ptr, a.bytes);
      a.ptr = nullptr;
      a.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 4096);
    want = (want + 4095) & ~size_t(4095);
    GDB_TRY(dev_alloc(ctx, &a.ptr, want));
    a.bytes = want;
  }
  *out = a.ptr;
  return GDB_OK;
}

// Byte fill: 16-byte stores over the aligned body, byte stores for the head and
// tail. A kernel of its own, so it can also sit in captured graphs.
__global__ __launch_bounds__(256) void fill_bytes(uint8_t* __restrict__ p, uint32_t word,
                                                  uint64_t head, uint64_t n16, uint64_t bytes) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint4* q = (uint4*)(p + head);
  const uint4 v = make_uint4(word, word, word, word);
  for (uint64_t i = tid; i < n16; i += stride) q[i] = v;
  const uint64_t tail0 = head + n16 * 16;
  if (tid < head) p[tid] = (uint8_t)word;
  if (tid < bytes - tail0) p[tail0 + tid] = (uint8_t)word;
}

int sum_fits(uint64_t n, unsigned __int128 term, const char* what) {
  const unsigned __int128 lim = (unsigned __int128)1 << 63;
  if (n && term && (term >= lim || (unsigned __int128)n * term >= lim))
    return fail(GDB_E_UNSUPPORTED, std::string(what) + " could exceed int64 for this table "
                                   "(rows x the largest term must stay below 2^63)");
  return GDB_OK;
}

int fill(void* ptr, int byte, size_t bytes, hipStream_t s) {
  if (!bytes) return GDB_OK;
  uint8_t* p = (uint8_t*)ptr;
  const uint64_t mis = (uintptr_t)p & 15;
  const uint64_t head = std::min<uint64_t>(bytes, mis ? 16 - mis : 0);
  const uint64_t n16 = (bytes - head) / 16;
  const uint32_t word = 0x01010101u * (uint8_t)byte;
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((n16 + 1023) / 1024, 8192));
  fill_bytes<<<(unsigned)blocks, 256, 0, s>>>(p, word, head, n16, bytes);
  GDB_HIP(hipGetLastError());
  return GDB_OK;
}

int pinned(gdb_ctx* ctx, size_t bytes, void** out) {
  if (ctx->host_pinned_bytes < bytes) {
    if (ctx->host_pinned) {
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      (void)hipHostFree(ctx->host_pinned);
    }
    ctx->host_pinned = nullptr;
    ctx->host_pinned_bytes = 0;
    size_t want = std::max<size_t>(bytes, 65536);
    if (hipHostMalloc(&ctx->host_pinned, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      ctx->host_pinned = nullptr;
      return fail(GDB_E_NOMEM, "hipHostMalloc failed");
    }
    ctx->host_pinned_bytes = want;
  }
  *out = ctx->host_pinned;
  return GDB_OK;
}

int result_buf(gdb_ctx* ctx, size_t bytes, void** host, void** dev) {
  if (ctx->host_result_bytes < bytes) {
    if (ctx->host_result) {
      GDB_HIP(hipStreamSynchronize(ctx->stream));
      (void)hipHostFree(ctx->host_result);
    }
    ctx->host_result = nullptr;
    ctx->host_result_dev = nullptr;
    ctx->host_result_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 16384);
    void* h = nullptr;
    if (hipHostMalloc(&h, want, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
      (void)hipGetLastError();
      return fail(GDB_E_NOMEM, "hipHostMalloc (coherent result buffer) failed");
    }
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess || !d) {
      (void)hipGetLastError();
      (void)hipHostFree(h);
      return fail(GDB_E_HIP, "hipHostGetDevicePointer failed for the result buffer");
    }
    ctx->host_result = h;
    ctx->host_result_dev = d;
    ctx->host_result_bytes = want;
  }
  *host = ctx->host_result;
  *dev = ctx->host_result_dev;
  return GDB_OK;
}

int column(const gdb_table* t, const char* name, int type, const Column** out) {
  if (!t) return fail(GDB_E_INVALID, "null table");
  if (!name) return fail(GDB_E_INVALID, "null column name");
  for (const Column& c : t->cols) {
    if (c.name == name) {
      if (type && c.type != type)
        return fail(GDB_E_INVALID, std::string("column ") + name + " has the wrong type");
      *out = &c;
      return GDB_OK;
    }
  }
  return fail(GDB_E_INVALID, std::string("table has no column ") + name);
}

int allow_dynamic_lds(gdb_ctx* ctx, const void* kernel, size_t bytes) {
  std::lock_guard<std::mutex> g(ctx->occ_mu);
  size_t& granted = ctx->lds_attr[kernel];
  if (granted >= bytes) return GDB_OK;
  GDB_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  granted = bytes;
  return GDB_OK;
}

int resident_blocks(gdb_ctx* ctx, const void* kernel, int block, size_t lds, int per_cu_cap) {
  int per_cu = 0;
  const auto key = std::make_tuple(kernel, block, lds);
  {
    std::lock_guard<std::mutex> g(ctx->occ_mu);
    auto it = ctx->occ.find(key);
    if (it != ctx->occ.end()) per_cu = it->second;
  }
  if (per_cu <= 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess ||
        per_cu <= 0) {
      (void)hipGetLastError();
      per_cu = 2;
    }
    std::lock_guard<std::mutex> g(ctx->occ_mu);
    ctx->occ[key] = per_cu;
  }
  if (per_cu_cap > 0) per_cu = std::min(per_cu, per_cu_cap);
  return std::max(1, per_cu * ctx->info.cu_count);
}

int timing_finish(gdb_ctx* ctx, gdb_timing* t, uint64_t bytes, uint64_t main_bytes,
                  const RunStats* rs) {
  GDB_HIP(hipEventSynchronize(ctx->ev[5]));
  if (t) {
    float a = 0.f, b = 0.f;
    GDB_HIP(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[5]));
    if (rs && rs->chunks > 0) {
      b = (float)rs->kernel_ms;
    } else if (ctx->kernel_timing) {
      GDB_HIP(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
    } else {
      b = -1.f;  // not measured
    }
    t->gpu_ms = a;
    t->main_kernel_ms = b;
    t->bytes = bytes;
    t->main_kernel_bytes = main_bytes;
    t->h2d_bytes = rs ? rs->h2d_bytes : 0;
    t->chunks = rs ? rs->chunks : 0;
    t->post_ms = 0;
  }
  return GDB_OK;
}

// Min / max of an integer host column, split over threads for large columns.
template <typename T>
static void value_range(const T* v, uint64_t n, int64_t* lo, int64_t* hi, bool* sorted = nullptr,
                        bool* strict = nullptr) {
  if (sorted) *sorted = true;
  if (strict) *strict = true;
  if (n == 0) {
    *lo = 0;
    *hi = -1;
    return;
  }
  const int T_ = n > (1u << 22) ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1;
  std::vector<int64_t> mn(T_, INT64_MAX), mx(T_, INT64_MIN);
  std::vector<char> up(T_, 1), inc(T_, 1);
  const uint64_t per = (n + T_ - 1) / T_;
  auto body = [&](int t) {
    const uint64_t a = (uint64_t)t * per, b = std::min(n, a + per);
    int64_t l = INT64_MAX, h = INT64_MIN;
    bool asc = true, str = true;
    for (uint64_t i = a; i < b; ++i) {
      l = std::min<int64_t>(l, (int64_t)v[i]);
      h = std::max<int64_t>(h, (int64_t)v[i]);
      asc &= i == 0 || v[i - 1] <= v[i];
      str &= i == 0 || v[i - 1] < v[i];
    }
    mn[t] = l;
    mx[t] = h;
    up[t] = asc;
    inc[t] = str;
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T_; ++t) th.emplace_back(body, t);
  body(0);
  for (auto& x : th) x.join();
  *lo = *std::min_element(mn.begin(), mn.end());
  *hi = *std::max_element(mx.begin(), mx.end());
  if (sorted) *sorted = std::all_of(up.begin(), up.end(), [](char c) { return c != 0; });
  if (strict) *strict = std::all_of(inc.begin(), inc.end(), [](char c) { return c != 0; });
}

}  // namespace gdb

using namespace gdb;

extern "C" {

const char* gdb_strerror(int code) {
  switch (code) {
    case GDB_OK: return "ok";
    case GDB_E_INVALID: return "invalid argument";
    case GDB_E_NOMEM: return "out of memory";
    case GDB_E_HIP: return "HIP runtime error";
    case GDB_E_ADMISSION: return "request exceeds the HBM admission budget";
    case GDB_E_UNSUPPORTED: return "unsupported input";
    case GDB_E_NODEVICE: return "no GPU device";
    case GDB_E_TIMEOUT: return "timed out";
    case GDB_E_CANCELLED: return "cancelled";
    case GDB_E_NOTFOUND: return "not found";
    default: return "unknown error";
  }
}

const char* gdb_last_error(void) { return g_last_error.c_str(); }

int gdb_open(int device, uint64_t hbm_budget_bytes, gdb_ctx** out) {
  if (!out) return fail(GDB_E_INVALID, "null output pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(GDB_E_NODEVICE, "no HIP device visible");
  }
  if (device < 0 || device >= n) return fail(GDB_E_INVALID, "device index out of range");
  GDB_HIP(hipSetDevice(device));
  gdb_ctx* ctx = new gdb_ctx();
  ctx->device = device;
  ctx->arenas.resize(A_NSLOTS);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return fail(GDB_E_HIP, "hipGetDeviceProperties failed");
  }
  std::strncpy(ctx->info.name, prop.name, sizeof(ctx->info.name) - 1);
  std::strncpy(ctx->info.arch, prop.gcnArchName, sizeof(ctx->info.arch) - 1);
  ctx->info.cu_count = prop.multiProcessorCount;
  int xcc = 1;
  if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, device) != hipSuccess) {
    (void)hipGetLastError();
    xcc = 1;
  }
  ctx->info.xcd_count = xcc;
  ctx->info.hbm_bytes = prop.totalGlobalMem;
  ctx->info.l2_bytes = (uint64_t)prop.l2CacheSize;
  ctx->info.lds_bytes_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
  ctx->info.wave_size = prop.warpSize;
  ctx->info.hbm_peak_gbps = 8000.0;  // MI355X HBM3E published peak
  // Launch-shape defaults (measured on MI355X: nontemporal dwordx4 streams run
  // best with 4 groups in flight per lane at 2 workgroups of 256 per CU).
  ctx->tune[T_Q6] = Tuning{256, 2, 4};
  // Q1's lane-private LDS tables take 12 KiB per wave; the autotuner's pick
  // on 60M rows is four 128-thread workgroups per CU (233 us, against 243 us
  // for three of 256).
  ctx->tune[T_Q1] = Tuning{128, 4, 2};
  // Q9's probe: its launch shape is fixed (one 512-thread workgroup per CU,
  // two 4-row groups per lane and trip).
  ctx->tune[T_Q9] = Tuning{256, 0, 2};
  // The Q13 matcher runs best at its occupancy limit (5 waves per SIMD):
  // 253 us for 15M comments, against 278 us at 4 workgroups per CU.
  ctx->tune[T_Q13] = Tuning{256, 0, 1};
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    free_b = prop.totalGlobalMem;
  }
  const uint64_t reserve = 4ull << 30;
  ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (free_b > reserve ? free_b - reserve : free_b);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    gdb_close(ctx);
    return fail(GDB_E_HIP, "stream creation failed");
  }
  // The allocation-free entry point's scratch address on this device; a
  // failure here is reported by that entry point, which retries the lookup.
  if (noalloc_scratch_init(device) != GDB_OK) (void)hipGetLastError();
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      gdb_close(ctx);
      return fail(GDB_E_HIP, "event creation failed");
    }
  }
  for (auto& e : ctx->sev) {
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      gdb_close(ctx);
      return fail(GDB_E_HIP, "event creation failed");
    }
  }
  *out = ctx;
  return GDB_OK;
}

void gdb_close(gdb_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  for (Arena& a : ctx->arenas) dev_free(ctx, a.ptr, a.bytes);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->host_result) (void)hipHostFree(ctx->host_result);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->sev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->chunk_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  delete ctx;
}

int gdb_device_info_get(gdb_ctx* ctx, gdb_device_info* out) {
  if (!ctx || !out) return fail(GDB_E_INVALID, "null argument");
  *out = ctx->info;
  return GDB_OK;
}

uint64_t gdb_hbm_used(gdb_ctx* ctx) { return ctx ? ctx->used : 0; }

void gdb_debug_fail_allocs(gdb_ctx* ctx, int count) { gdb_debug_fail_allocs_after(ctx, 0, count); }

void gdb_debug_fail_allocs_after(gdb_ctx* ctx, int skip, int count) {
  if (!ctx) return;
  ctx->fail_skip = std::max(skip, 0);
  ctx->fail_allocs = count;
}

int gdb_set_kernel_timing(gdb_ctx* ctx, int on) {
  if (!ctx) return fail(GDB_E_INVALID, "null argument");
  ctx->kernel_timing = on != 0;
  return GDB_OK;
}

int gdb_release_scratch(gdb_ctx* ctx) {
  if (!ctx) return fail(GDB_E_INVALID, "null argument");
  GDB_HIP(hipSetDevice(ctx->device));
  GDB_HIP(hipStreamSynchronize(ctx->stream));
  GDB_HIP(hipStreamSynchronize(ctx->copy_stream));
  for (Arena& a : ctx->arenas) {
    dev_free(ctx, a.ptr, a.bytes);
    a.ptr = nullptr;
    a.bytes = 0;
  }
  return GDB_OK;
}

int gdb_table_create(gdb_ctx* ctx, const gdb_col_desc* cols, int ncols, uint64_t nrows,
                     gdb_table** out) {
  if (!ctx || !out || (ncols > 0 && !cols) || ncols < 0)
    return fail(GDB_E_INVALID, "null argument");
  *out = nullptr;
  if (nrows >= (1ull << 32)) return fail(GDB_E_UNSUPPORTED, "tables are limited to 2^32-1 rows");
  GDB_HIP(hipSetDevice(ctx->device));
  gdb_table* t = new gdb_table();
  t->ctx = ctx;
  t->nrows = nrows;
  for (int i = 0; i < ncols; ++i) {
    const gdb_col_desc& d = cols[i];
    Column c;
    if (!d.name) {
      gdb_table_free(t);
      return fail(GDB_E_INVALID, "column without a name");
    }
    c.name = d.name;
    c.type = d.type;
    switch (d.type) {
      case GDB_I32: c.width = 4; break;
      case GDB_I64: c.width = 8; break;
      case GDB_U8: c.width = 1; break;
      case GDB_BYTES: c.width = d.width; break;
      default: gdb_table_free(t); return fail(GDB_E_INVALID, "bad column type for " + c.name);
    }
    if (c.width <= 0) {
      gdb_table_free(t);
      return fail(GDB_E_INVALID, "bad column width for " + c.name);
    }
    // Pad every column to a 64-byte multiple so dwordx4 loads of the last
    // vector group stay inside the allocation; the pad is zeroed.
    c.bytes = (size_t)nrows * (size_t)c.width;
    size_t alloc = ((c.bytes + 63) & ~size_t(63)) + 64;
    int rc = dev_alloc(ctx, &c.dev, alloc);
    if (rc != GDB_OK) {
      gdb_table_free(t);
      return rc;
    }
    if (d.host_data && d.type == GDB_I32)
      value_range((const int32_t*)d.host_data, nrows, &c.minv, &c.maxv, &c.sorted, &c.strict);
    else if (d.host_data && d.type == GDB_I64)
      value_range((const int64_t*)d.host_data, nrows, &c.minv, &c.maxv);
    t->cols.push_back(c);
    Column& cc = t->cols.back();
    cc.bytes = alloc;  // accounted size
    if (hipMemsetAsync(cc.dev, 0, alloc, ctx->stream) != hipSuccess ||
        (c.bytes && d.host_data &&
         hipMemcpyAsync(cc.dev, d.host_data, (size_t)nrows * c.width, hipMemcpyHostToDevice,
                        ctx->stream) != hipSuccess)) {
      gdb_table_free(t);
      return fail(GDB_E_HIP, "upload of column " + c.name + " failed");
    }
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    gdb_table_free(t);
    return fail(GDB_E_HIP, "upload synchronisation failed");
  }
  *out = t;
  return GDB_OK;
}

void gdb_table_free(gdb_table* t) {
  if (!t) return;
  if (t->ctx && !t->host) {
    (void)hipStreamSynchronize(t->ctx->stream);
    for (Column& c : t->cols) dev_free(t->ctx, c.dev, c.bytes);
  }
  delete t;
}

int gdb_table_create_host(gdb_ctx* ctx, const gdb_col_desc* cols, int ncols, uint64_t nrows,
                          gdb_table** out) {
  if (!ctx || !out || (ncols > 0 && !cols) || ncols < 0)
    return fail(GDB_E_INVALID, "null argument");
  *out = nullptr;
  if (nrows >= (1ull << 40)) return fail(GDB_E_UNSUPPORTED, "host tables are limited to 2^40 rows");
  gdb_table* t = new gdb_table();
  t->ctx = ctx;
  t->nrows = nrows;
  t->host = true;
  for (int i = 0; i < ncols; ++i) {
    const gdb_col_desc& d = cols[i];
    Column c;
    if (!d.name || (!d.host_data && nrows)) {
      delete t;
      return fail(GDB_E_INVALID, "host column without a name or data");
    }
    c.name = d.name;
    c.type = d.type;
    c.width = d.type == GDB_I32 ? 4 : d.type == GDB_I64 ? 8 : d.type == GDB_U8 ? 1
              : d.type == GDB_BYTES ? d.width : 0;
    if (c.width <= 0) {
      delete t;
      return fail(GDB_E_INVALID, "bad column type or width for " + c.name);
    }
    c.host = d.host_data;
    if (d.type == GDB_I32) value_range((const int32_t*)d.host_data, nrows, &c.minv, &c.maxv);
    else if (d.type == GDB_I64) value_range((const int64_t*)d.host_data, nrows, &c.minv, &c.maxv);
    t->cols.push_back(c);
  }
  *out = t;
  return GDB_OK;
}

int gdb_table_is_host(const gdb_table* t) { return t && t->host ? 1 : 0; }

int gdb_host_alloc(gdb_ctx* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return fail(GDB_E_INVALID, "null argument");
  *out = nullptr;
  GDB_HIP(hipSetDevice(ctx->device));
  if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return fail(GDB_E_NOMEM, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return GDB_OK;
}

void gdb_host_free(gdb_ctx* ctx, void* p) {
  if (!p) return;
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->copy_stream);
  }
  (void)hipHostFree(p);
}

int gdb_set_chunk_rows(gdb_ctx* ctx, uint64_t rows) {
  if (!ctx) return fail(GDB_E_INVALID, "null context");
  if (rows && rows < 1024) return fail(GDB_E_INVALID, "chunks must hold at least 1024 rows");
  ctx->chunk_rows = rows;
  return GDB_OK;
}

uint64_t gdb_table_rows(const gdb_table* t) { return t ? t->nrows : 0; }

uint64_t gdb_table_bytes(const gdb_table* t) {
  uint64_t b = 0;
  if (t)
    for (const Column& c : t->cols) b += (uint64_t)t->nrows * c.width;
  return b;
}

}  // extern "C"
