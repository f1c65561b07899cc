// Fused multi-request execution for the serving loop (no reference
// counterpart: the reference runs one query per command buffer).
#pragma once
#include <stdint.h>

#include <vector>

#include "gdb.h"
#include "graph.h"
#include "internal.h"

namespace gdb {

// ---------------------------------------------------------------------------
// Flat-table batch: any mix of Q6 and Q1 requests over one fact table in one
// pass. The batch's predicate constants cut each attribute's domain into
// intervals (a value's interval = number of boundaries <= value); every row
// lands in exactly one cell of the product grid, so the kernel's work per row
// does not depend on how many requests share the pass. Each request's answer
// is then the exact sum of the cells inside its box.
//   Q6 cube:  date x discount x quantity intervals -> SUM(price*disc), count
//   Q1 table: date interval x group bin -> count, qty, disc, price, dp, dp*tax
constexpr int FB_MAX_DATE = 160;   // date intervals
constexpr int FB_MAX_DISC = 16;
constexpr int FB_MAX_QTY = 8;
constexpr int FB_LUT_DAYS = 4096;  // span of the date lookup table
constexpr int FB_LUT_SMALL = 128;  // discount / quantity lookup range [0, 128)
constexpr int FB_LDS_LIMIT = 144 * 1024;

struct FlatBatchDesc {
  int32_t lut_lo;         // first day of the date lookup table
  int32_t n_date, n_disc, n_qty;
  int32_t has_q6, has_q1;
  int32_t q6_cells, q1_cells;
  int32_t copies;         // copies of the cell table in LDS (set at launch)
  uint8_t date_lut[FB_LUT_DAYS];
  uint8_t disc_lut[FB_LUT_SMALL];
  uint8_t qty_lut[FB_LUT_SMALL];
  int32_t n_date_b;       // date boundaries (n_date - 1), for out-of-table days
};

struct FlatPlan {
  std::vector<int32_t> date_b, disc_b, qty_b;  // sorted unique boundaries
  FlatBatchDesc desc;
  size_t lds_bytes;
  int ncells;  // q6 cells*2 + q1 cells*6 partial columns
};

// Interval index of v among sorted boundaries b.
int interval_of(const std::vector<int32_t>& b, int32_t v);

// Builds the plan for a set of requests; returns false if it does not fit.
// With fill_tables = false only the boundaries and sizes are computed (the
// cheap fit check used while a batch is being formed).
bool plan_flat_batch(const std::vector<gdb_q6_params>& q6, const std::vector<gdb_q1_params>& q1,
                     FlatPlan* plan, size_t lds_limit = FB_LDS_LIMIT, bool fill_tables = true);

// Device state of one flat batch slot: sized for the largest plan, reused.
struct FlatSlot {
  FlatBatchDesc* d_desc = nullptr;
  int64_t* partials = nullptr;  // [ncells][blocks]
  int64_t* totals = nullptr;    // [ncells]
  int64_t* h_totals = nullptr;  // pinned
  FlatBatchDesc* h_desc = nullptr;  // pinned
  int blocks = 0;
  size_t max_cells = 0;
  size_t lds_limit = 0;  // dynamic LDS the scan kernel may use on this device
  bool graphs = false;   // replay the scan + finaliser of resident tables as graphs
  GraphCache gc;         // one graph per (table, LDS size class)
};

int flat_slot_init(gdb_ctx* ctx, FlatSlot* s);
void flat_slot_free(gdb_ctx* ctx, FlatSlot* s);
// Enqueues the batch on `stream` (desc upload, scan per chunk, finaliser,
// result download). Host tables stream through `run_over` (synchronous).
int flat_batch_launch(gdb_ctx* ctx, FlatSlot* s, const FlatPlan& plan, const gdb_table* li,
                      hipStream_t stream);
// Reads the cells into per-request results (after the stream completed).
void flat_batch_results(const FlatPlan& plan, const int64_t* totals,
                        const std::vector<gdb_q6_params>& q6, const std::vector<gdb_q1_params>& q1,
                        std::vector<gdb_q6_result>* r6, std::vector<gdb_q1_result>* r1);

// ---------------------------------------------------------------------------
// Q3 batch: up to 64 (segment, date) requests over one customer/orders/fact
// set, answered from date-ordered join indexes built once per table set: per
// request its window of orders, its groups, a revenue histogram, candidates
// above the k-th bucket and a block top-k.
constexpr int Q3B_MAX = 64;

struct Q3BatchDesc {
  int32_t n;                  // requests
  int32_t k[Q3B_MAX];         // top-k per request
  int32_t segment[Q3B_MAX];
  int32_t date[Q3B_MAX];
};

struct Q3Slot;  // defined in batch_q3.hip
int q3_slot_create(gdb_ctx* ctx, const gdb_table* customer, const gdb_table* orders,
                   const gdb_table* lineitem, Q3Slot** out);
// Capture the batch kernels as one graph per (batch size, grid) and replay it.
void q3_slot_use_graphs(Q3Slot* s, bool on);
void q3_slot_graph_counts(const Q3Slot* s, uint64_t* captures, uint64_t* replays);
void q3_slot_free(gdb_ctx* ctx, Q3Slot* s);
int q3_batch_launch(gdb_ctx* ctx, Q3Slot* s, const std::vector<gdb_q3_params>& reqs,
                    hipStream_t stream);
int q3_batch_results(Q3Slot* s, const std::vector<gdb_q3_params>& reqs,
                     std::vector<gdb_q3_result>* out);
uint64_t q3_slot_bytes(const Q3Slot* s);

}  // namespace gdb
