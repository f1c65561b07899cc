// Fused Q6 + Q1 multi-request scan: one pass over the fact table answers every
// Q6 and Q1 request of a batch exactly (see batch.h). The reference runs one
// query per command buffer (src/q6.cpp:52-94, src/q1.cpp:59-105) and has no
// batching; this is new.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>

#include "batch.h"
#include "kcommon.h"

using namespace gdbk;

namespace gdb {

int interval_of(const std::vector<int32_t>& b, int32_t v) {
  return (int)(std::upper_bound(b.begin(), b.end(), v) - b.begin());
}

static void uniq(std::vector<int32_t>* v) {
  std::sort(v->begin(), v->end());
  v->erase(std::unique(v->begin(), v->end()), v->end());
}

bool plan_flat_batch(const std::vector<gdb_q6_params>& q6, const std::vector<gdb_q1_params>& q1,
                     FlatPlan* plan, size_t lds_limit, bool fill_tables) {
  FlatPlan& p = *plan;
  p.date_b.clear();
  p.disc_b.clear();
  p.qty_b.clear();
  for (const auto& r : q6) {
    p.date_b.push_back(r.date_lo);
    p.date_b.push_back(r.date_hi);
    p.disc_b.push_back(r.disc_lo);
    p.disc_b.push_back(r.disc_hi + 1);
    p.qty_b.push_back(r.qty_lt);
  }
  for (const auto& r : q1) {
    if (r.cutoff == INT32_MAX) return false;
    p.date_b.push_back(r.cutoff + 1);
  }
  uniq(&p.date_b);
  uniq(&p.disc_b);
  uniq(&p.qty_b);
  for (int32_t v : p.disc_b)
    if (v < 0 || v > FB_LUT_SMALL) return false;
  for (int32_t v : p.qty_b)
    if (v < 0 || v > FB_LUT_SMALL) return false;
  if (!p.date_b.empty() && (int64_t)p.date_b.back() - p.date_b.front() >= FB_LUT_DAYS) return false;
  FlatBatchDesc& d = p.desc;
  memset(&d, 0, offsetof(FlatBatchDesc, date_lut));
  d.has_q6 = !q6.empty();
  d.has_q1 = !q1.empty();
  d.n_date = (int)p.date_b.size() + 1;
  d.n_disc = d.has_q6 ? (int)p.disc_b.size() + 1 : 1;
  d.n_qty = d.has_q6 ? (int)p.qty_b.size() + 1 : 1;
  if (d.n_date > FB_MAX_DATE || d.n_disc > FB_MAX_DISC || d.n_qty > FB_MAX_QTY) return false;
  d.q6_cells = d.has_q6 ? d.n_date * d.n_disc * d.n_qty : 0;
  d.q1_cells = d.has_q1 ? d.n_date * 6 : 0;
  d.lut_lo = p.date_b.empty() ? 0 : p.date_b.front();
  d.n_date_b = (int)p.date_b.size();
  p.ncells = d.q6_cells * 2 + d.q1_cells * 6;
  p.lds_bytes = (size_t)p.ncells * 8 + FB_LUT_DAYS + 2 * FB_LUT_SMALL;
  if (p.lds_bytes > std::min<size_t>(lds_limit, FB_LDS_LIMIT)) return false;
  if (fill_tables) {
    // Linear sweeps: entry i holds the number of boundaries <= its value.
    auto sweep = [](const std::vector<int32_t>& b, int32_t lo, int n, uint8_t* lut) {
      size_t k = 0;
      for (int i = 0; i < n; ++i) {
        while (k < b.size() && b[k] <= lo + i) ++k;
        lut[i] = (uint8_t)k;
      }
    };
    sweep(p.date_b, d.lut_lo, FB_LUT_DAYS, d.date_lut);
    sweep(p.disc_b, 0, FB_LUT_SMALL, d.disc_lut);
    sweep(p.qty_b, 0, FB_LUT_SMALL, d.qty_lut);
  }
  return true;
}

namespace {

struct FlatCols {
  const int32_t *ship, *qty, *disc, *tax;
  const int64_t* price;
  const uint8_t *rf, *ls;
};

__device__ __forceinline__ uint32_t fb_rf(uint32_t c) {
  return c == 'A' ? 0u : c == 'N' ? 1u : c == 'R' ? 2u : 3u;
}
__device__ __forceinline__ uint32_t fb_ls(uint32_t c) { return c == 'F' ? 0u : c == 'O' ? 1u : 2u; }

__device__ __forceinline__ void lds_add(unsigned long long* p, uint64_t v) {
  __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct FbState {
  unsigned long long* tab;
  const uint8_t *dlut, *clut, *qlut;
  int32_t lut_lo, n_date, n_disc, n_qty, q6_off_q1;
  bool has_q6, has_q1;
};

__device__ __forceinline__ void fb_row(const FbState& s, int32_t ship, int32_t qty, int32_t disc,
                                       int32_t tax, int64_t price, uint32_t rf, uint32_t ls) {
  const int32_t rel = ship - s.lut_lo;
  const int di = rel < 0 ? 0 : rel >= FB_LUT_DAYS ? s.n_date - 1 : s.dlut[rel];
  if (s.has_q6) {
    const int ci = disc < 0 ? 0 : disc >= FB_LUT_SMALL ? s.n_disc - 1 : s.clut[disc];
    const int qi = qty < 0 ? 0 : qty >= FB_LUT_SMALL ? s.n_qty - 1 : s.qlut[qty];
    const int cell = (di * s.n_disc + ci) * s.n_qty + qi;
    lds_add(&s.tab[2 * cell], (uint64_t)(price * (int64_t)disc));
    lds_add(&s.tab[2 * cell + 1], 1ull);
  }
  if (s.has_q1) {
    const uint32_t rc = fb_rf(rf), lc = fb_ls(ls);
    if (rc < 3u && lc < 2u) {
      unsigned long long* t = s.tab + s.q6_off_q1 + (size_t)(di * 6 + rc * 2 + lc) * 6;
      const int64_t dp = price * (int64_t)(100 - disc);
      lds_add(&t[0], 1ull);
      lds_add(&t[1], (uint64_t)(int64_t)qty);
      lds_add(&t[2], (uint64_t)(int64_t)disc);
      lds_add(&t[3], (uint64_t)price);
      lds_add(&t[4], (uint64_t)dp);
      lds_add(&t[5], (uint64_t)(dp * (int64_t)tax));
    }
  }
}

// The cell table is replicated `copies` times in LDS (up to one copy per 16
// lanes): with a batch of a few requests the table has a few dozen cells, and
// 1024 threads adding into one copy serialised on the same addresses.
__global__ __launch_bounds__(1024) void flat_batch_scan(FlatCols c, uint32_t n,
                                                        const FlatBatchDesc* __restrict__ desc,
                                                        int64_t* __restrict__ partials) {
  extern __shared__ unsigned long long sm[];
  const int ncells = desc->q6_cells * 2 + desc->q1_cells * 6;
  const int copies = desc->copies;
  uint8_t* luts = reinterpret_cast<uint8_t*>(sm + (size_t)copies * ncells);
  for (int i = threadIdx.x; i < copies * ncells; i += blockDim.x) sm[i] = 0ull;
  for (int i = threadIdx.x; i < FB_LUT_DAYS / 4; i += blockDim.x)
    reinterpret_cast<uint32_t*>(luts)[i] = reinterpret_cast<const uint32_t*>(desc->date_lut)[i];
  for (int i = threadIdx.x; i < FB_LUT_SMALL / 4; i += blockDim.x) {
    reinterpret_cast<uint32_t*>(luts + FB_LUT_DAYS)[i] =
        reinterpret_cast<const uint32_t*>(desc->disc_lut)[i];
    reinterpret_cast<uint32_t*>(luts + FB_LUT_DAYS + FB_LUT_SMALL)[i] =
        reinterpret_cast<const uint32_t*>(desc->qty_lut)[i];
  }
  __syncthreads();
  FbState s;
  s.tab = sm + (size_t)((threadIdx.x >> 4) % copies) * ncells;
  s.dlut = luts;
  s.clut = luts + FB_LUT_DAYS;
  s.qlut = luts + FB_LUT_DAYS + FB_LUT_SMALL;
  s.lut_lo = desc->lut_lo;
  s.n_date = desc->n_date;
  s.n_disc = desc->n_disc;
  s.n_qty = desc->n_qty;
  s.q6_off_q1 = desc->q6_cells * 2;
  s.has_q6 = desc->has_q6 != 0;
  s.has_q1 = desc->has_q1 != 0;
  const uint32_t nvec = n >> 2;
  const uint32_t nthr = gridDim.x * blockDim.x;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < nvec; g += nthr) {
    const int4 sh = ld4(c.ship, g), q = ld4(c.qty, g), d = ld4(c.disc, g);
    const longlong2 a = ld2(c.price, 2 * g), b = ld2(c.price, 2 * g + 1);
    int4 t = make_int4(0, 0, 0, 0);
    uint32_t rf = 0, ls = 0;
    if (s.has_q1) {
      t = ld4(c.tax, g);
      rf = ld_u8x4(c.rf, g);
      ls = ld_u8x4(c.ls, g);
    }
    fb_row(s, sh.x, q.x, d.x, t.x, a.x, rf & 0xff, ls & 0xff);
    fb_row(s, sh.y, q.y, d.y, t.y, a.y, (rf >> 8) & 0xff, (ls >> 8) & 0xff);
    fb_row(s, sh.z, q.z, d.z, t.z, b.x, (rf >> 16) & 0xff, (ls >> 16) & 0xff);
    fb_row(s, sh.w, q.w, d.w, t.w, b.y, rf >> 24, ls >> 24);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const uint32_t r = nvec * 4 + threadIdx.x;
    fb_row(s, c.ship[r], c.qty[r], c.disc[r], s.has_q1 ? c.tax[r] : 0, c.price[r],
           s.has_q1 ? c.rf[r] : 0u, s.has_q1 ? c.ls[r] : 0u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ncells; i += blockDim.x) {
    uint64_t v = 0;
    for (int k = 0; k < copies; ++k) v += sm[(size_t)k * ncells + i];
    partials[(size_t)i * gridDim.x + blockIdx.x] = (int64_t)v;
  }
}

// Copies of an ncells table that fit `lds` bytes next to the lookup tables.
int table_copies(int ncells, size_t lds) {
  const size_t luts = FB_LUT_DAYS + 2 * FB_LUT_SMALL;
  if (ncells <= 0 || lds <= luts) return 1;
  return (int)std::max<size_t>(1, std::min<size_t>(64, (lds - luts) / ((size_t)ncells * 8)));
}

}  // namespace

namespace {
// Finaliser of a graph-replayed batch: cell c = sum of its per-block partials,
// with the cell count taken from the descriptor (fixed grid, cells strided).
__global__ __launch_bounds__(256) void flat_cells_sum(const FlatBatchDesc* __restrict__ d,
                                                      const int64_t* __restrict__ partials,
                                                      uint32_t nparts, int64_t* __restrict__ out) {
  __shared__ int64_t red[4];
  const int ncells = d->q6_cells * 2 + d->q1_cells * 6;
  for (int c = blockIdx.x; c < ncells; c += gridDim.x) {
    const int64_t* col = partials + (size_t)c * nparts;
    int64_t s = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) s += col[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[c] = s;
  }
}
}  // namespace

int flat_slot_init(gdb_ctx* ctx, FlatSlot* s) {
  // Use up to FB_LDS_LIMIT of dynamic LDS where the device allows it (160 KB per
  // CU on gfx950), else the default 64 KB per workgroup.
  s->lds_limit = 64 * 1024;
  if (hipFuncSetAttribute((const void*)flat_batch_scan, hipFuncAttributeMaxDynamicSharedMemorySize,
                          FB_LDS_LIMIT) == hipSuccess &&
      ctx->info.lds_bytes_per_cu >= FB_LDS_LIMIT)
    s->lds_limit = FB_LDS_LIMIT;
  (void)hipGetLastError();
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)flat_batch_scan, 1024,
                                                   s->lds_limit / 2) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  s->blocks = std::max(1, std::min(per_cu, 2)) * ctx->info.cu_count;
  s->max_cells = FB_LDS_LIMIT / 8;
  GDB_TRY(dev_alloc(ctx, (void**)&s->d_desc, sizeof(FlatBatchDesc)));
  GDB_TRY(dev_alloc(ctx, (void**)&s->partials, s->max_cells * s->blocks * sizeof(int64_t)));
  GDB_TRY(dev_alloc(ctx, (void**)&s->totals, s->max_cells * sizeof(int64_t)));
  if (hipHostMalloc((void**)&s->h_totals, s->max_cells * sizeof(int64_t), 0) != hipSuccess ||
      hipHostMalloc((void**)&s->h_desc, sizeof(FlatBatchDesc), 0) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GDB_E_NOMEM, "pinned allocation for a batch slot failed");
  }
  return GDB_OK;
}

void flat_slot_free(gdb_ctx* ctx, FlatSlot* s) {
  graph_cache_clear(&s->gc);
  dev_free(ctx, s->d_desc, sizeof(FlatBatchDesc));
  dev_free(ctx, s->partials, s->max_cells * s->blocks * sizeof(int64_t));
  dev_free(ctx, s->totals, s->max_cells * sizeof(int64_t));
  if (s->h_totals) (void)hipHostFree(s->h_totals);
  if (s->h_desc) (void)hipHostFree(s->h_desc);
  *s = FlatSlot{};
}

int flat_batch_launch(gdb_ctx* ctx, FlatSlot* s, const FlatPlan& plan, const gdb_table* li,
                      hipStream_t stream) {
  const Column* c[7];
  GDB_TRY(column(li, "l_shipdate", GDB_I32, &c[0]));
  GDB_TRY(column(li, "l_quantity", GDB_I32, &c[1]));
  GDB_TRY(column(li, "l_discount", GDB_I32, &c[2]));
  GDB_TRY(column(li, "l_tax", GDB_I32, &c[3]));
  GDB_TRY(column(li, "l_extendedprice", GDB_I64, &c[4]));
  GDB_TRY(column(li, "l_returnflag", GDB_U8, &c[5]));
  GDB_TRY(column(li, "l_linestatus", GDB_U8, &c[6]));
  if ((size_t)plan.ncells > s->max_cells) return fail(GDB_E_UNSUPPORTED, "batch plan too large");
  // LDS of the launch: the plan's LDS size class (two workgroups per CU up to
  // half the limit); the spare room holds copies of the cell table.
  const size_t lds_class = plan.lds_bytes <= 32768   ? 32768
                           : plan.lds_bytes <= 65536 ? 65536
                           : plan.lds_bytes <= 98304 ? 98304
                                                     : s->lds_limit;
  *s->h_desc = plan.desc;
  s->h_desc->copies = table_copies(plan.ncells, std::min(lds_class, s->lds_limit));
  GDB_HIP(hipMemcpyAsync(s->d_desc, s->h_desc, sizeof(FlatBatchDesc), hipMemcpyHostToDevice, stream));
  int launches = 0;
  auto consume = [&](const void* const* d, uint32_t rows) -> int {
    const FlatCols fc{(const int32_t*)d[0], (const int32_t*)d[1], (const int32_t*)d[2],
                      (const int32_t*)d[3], (const int64_t*)d[4], (const uint8_t*)d[5],
                      (const uint8_t*)d[6]};
    const int blocks = std::max(1, std::min<int>(s->blocks, (int)((rows / 4 + 1023) / 1024)));
    flat_batch_scan<<<blocks, 1024, std::min(lds_class, s->lds_limit), stream>>>(fc, rows, s->d_desc,
                                                                              s->partials);
    GDB_HIP(hipGetLastError());
    if (launches++ == 0)
      sum_columns<0><<<plan.ncells, 256, 0, stream>>>(s->partials, (uint32_t)blocks, s->totals);
    else
      sum_columns<1><<<plan.ncells, 256, 0, stream>>>(s->partials, (uint32_t)blocks, s->totals);
    GDB_HIP(hipGetLastError());
    return GDB_OK;
  };
  if (li->host) {
    if (stream != ctx->stream) return fail(GDB_E_INVALID, "host tables stream on the engine stream");
    RunStats rs;
    GDB_TRY(run_over(ctx, li, c, 7, consume, &rs));
  } else if (li->nrows == 0) {
    GDB_TRY(fill(s->totals, 0, (size_t)plan.ncells * 8, stream));
  } else if (s->graphs) {
    // One graph per table and LDS size class: the scan is launched with the
    // class's LDS (the plan's layout lives in the descriptor) and the
    // finaliser reads the cell count from the descriptor.
    const uint32_t rows = (uint32_t)li->nrows;
    const int blocks = std::max(1, std::min<int>(s->blocks, (int)((rows / 4 + 1023) / 1024)));
    const FlatCols fc{(const int32_t*)c[0]->dev, (const int32_t*)c[1]->dev, (const int32_t*)c[2]->dev,
                      (const int32_t*)c[3]->dev, (const int64_t*)c[4]->dev, (const uint8_t*)c[5]->dev,
                      (const uint8_t*)c[6]->dev};
    const uint64_t key = (reinterpret_cast<uintptr_t>(li) << 8) ^ (lds_class >> 12);
    GDB_TRY(graph_launch(&s->gc, key, stream, [&](hipStream_t st) -> int {
      flat_batch_scan<<<blocks, 1024, lds_class, st>>>(fc, rows, s->d_desc, s->partials);
      flat_cells_sum<<<512, 256, 0, st>>>(s->d_desc, s->partials, (uint32_t)blocks, s->totals);
      GDB_HIP(hipGetLastError());
      return GDB_OK;
    }));
  } else {
    const void* dev[7];
    for (int i = 0; i < 7; ++i) dev[i] = c[i]->dev;
    GDB_TRY(consume(dev, (uint32_t)li->nrows));
  }
  GDB_HIP(hipMemcpyAsync(s->h_totals, s->totals, (size_t)plan.ncells * 8, hipMemcpyDeviceToHost,
                         stream));
  return GDB_OK;
}

void flat_batch_results(const FlatPlan& plan, const int64_t* tot,
                        const std::vector<gdb_q6_params>& q6, const std::vector<gdb_q1_params>& q1,
                        std::vector<gdb_q6_result>* r6, std::vector<gdb_q1_result>* r1) {
  const FlatBatchDesc& d = plan.desc;
  r6->assign(q6.size(), gdb_q6_result{});
  for (size_t i = 0; i < q6.size(); ++i) {
    const gdb_q6_params& p = q6[i];
    if (p.date_hi <= p.date_lo || p.disc_hi < p.disc_lo) continue;
    const int a0 = interval_of(plan.date_b, p.date_lo), a1 = interval_of(plan.date_b, p.date_hi);
    const int c0 = interval_of(plan.disc_b, p.disc_lo), c1 = interval_of(plan.disc_b, p.disc_hi + 1);
    const int q1i = interval_of(plan.qty_b, p.qty_lt);
    int64_t rev = 0, rows = 0;
    for (int a = a0; a < a1; ++a)
      for (int cc = c0; cc < c1; ++cc)
        for (int q = 0; q < q1i; ++q) {
          const int cell = (a * d.n_disc + cc) * d.n_qty + q;
          rev += tot[2 * cell];
          rows += tot[2 * cell + 1];
        }
    (*r6)[i].revenue = rev;
    (*r6)[i].rows = (uint64_t)rows;
  }
  r1->assign(q1.size(), gdb_q1_result{});
  static const char RF[3] = {'A', 'N', 'R'}, LS[2] = {'F', 'O'};
  const int64_t* t1 = tot + d.q6_cells * 2;
  for (size_t i = 0; i < q1.size(); ++i) {
    const int lim = interval_of(plan.date_b, q1[i].cutoff + 1);
    gdb_q1_result& out = (*r1)[i];
    for (int bin = 0; bin < 6; ++bin) {
      int64_t v[6] = {0, 0, 0, 0, 0, 0};
      for (int a = 0; a < lim; ++a)
        for (int k = 0; k < 6; ++k) v[k] += t1[(a * 6 + bin) * 6 + k];
      if (v[0] == 0) continue;
      gdb_q1_row& r = out.rows[out.nrows++];
      r.returnflag = RF[bin / 2];
      r.linestatus = LS[bin % 2];
      r.count = v[0];
      r.sum_qty = v[1];
      r.sum_disc = v[2];
      r.sum_base_price = v[3];
      r.sum_disc_price = v[4];
      const __int128 charge = (__int128)v[4] * 100 + (__int128)v[5];
      r.sum_charge_lo = (uint64_t)charge;
      r.sum_charge_hi = (int64_t)(charge >> 64);
      r.avg_qty = (double)r.sum_qty / (double)r.count;
      r.avg_price = (double)r.sum_base_price / 100.0 / (double)r.count;
      r.avg_disc = (double)r.sum_disc / 100.0 / (double)r.count;
    }
  }
}

}  // namespace gdb
