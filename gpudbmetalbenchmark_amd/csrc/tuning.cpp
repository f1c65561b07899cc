// Autotuner: times candidate launch shapes of a pipeline's dominant kernel on
// real tables and keeps the fastest per context (no counterpart in the
// reference, which hard-codes its dispatch shapes, e.g. src/q6.cpp:72-74).
#include <algorithm>
#include <vector>

#include "internal.h"

using namespace gdb;

namespace {

struct Shape {
  int block, bpc, unroll, bits;
};

std::vector<Shape> candidates(int pipeline, const gdb_table* const* t, int n) {
  std::vector<Shape> c;
  switch (pipeline) {
    case GDB_PIPE_Q6:
      for (int bpc : {1, 2, 3, 4, 8, 0})
        for (int u : {1, 2, 4}) c.push_back({256, bpc, u, 0});
      break;
    case GDB_PIPE_Q1:
      for (int b : {128, 256})
        for (int bpc : {1, 2, 3, 4, 0})
          for (int u : {1, 2, 4}) c.push_back({b, bpc, u, 0});
      break;
    case GDB_PIPE_Q3:
    case GDB_PIPE_Q13:
      for (int bpc : {1, 2, 3, 4, 6, 8, 0}) c.push_back({256, bpc, 2, 0});
      break;
    case GDB_PIPE_Q9:
      // the probe's LDS filter fixes one 512-thread workgroup per CU and its
      // pipelined loop two 4-row groups per lane: a single shape
      c.push_back({256, 0, 2, 0});
      break;
    case GDB_PIPE_JOIN: {
      // one global table, the derived shape, and radix bits around the derived
      c.push_back({256, 0, 1, -1});
      c.push_back({256, 0, 1, -2});
      c.push_back({256, 0, 1, 0});
      if (n >= 1) {
        const int d = join_derived_bits(t[0]->nrows);
        for (int b = std::max(1, d - 2); b <= std::min(14, d + 1); ++b)
          if (b != d) c.push_back({256, 0, 1, b});
      }
      break;
    }
    default:
      break;
  }
  return c;
}

int run_once(gdb_ctx* ctx, int pipeline, const gdb_table* const* t, int n, double* ms) {
  gdb_timing tm{};
  int rc = GDB_E_INVALID;
  switch (pipeline) {
    case GDB_PIPE_Q6: {
      if (n < 1) break;
      gdb_q6_params p{731, 1096, 5, 7, 24};
      gdb_q6_result r;
      rc = gdb_q6(ctx, t[0], &p, &r, &tm);
      break;
    }
    case GDB_PIPE_Q1: {
      if (n < 1) break;
      gdb_q1_params p{2435};
      gdb_q1_result r;
      rc = gdb_q1(ctx, t[0], &p, &r, &tm);
      break;
    }
    case GDB_PIPE_Q3: {
      if (n < 3) break;
      gdb_q3_params p{1, 1169, 10};
      gdb_q3_result r;
      rc = gdb_q3(ctx, t[0], t[1], t[2], &p, &r, &tm);
      break;
    }
    case GDB_PIPE_Q9: {
      if (n < 5) break;
      gdb_q9_params p{};
      const char g[] = "green";
      std::copy(g, g + sizeof(g), p.pattern);
      static gdb_q9_result r;
      rc = gdb_q9(ctx, t[0], t[1], t[2], t[3], t[4], &p, &r, &tm);
      break;
    }
    case GDB_PIPE_Q13: {
      if (n < 2) break;
      gdb_q13_params p{};
      const char a[] = "special", b[] = "requests";
      std::copy(a, a + sizeof(a), p.word1);
      std::copy(b, b + sizeof(b), p.word2);
      static gdb_q13_result r;
      rc = gdb_q13(ctx, t[0], t[1], &p, &r, &tm);
      break;
    }
    case GDB_PIPE_JOIN: {
      if (n < 2) break;
      if (t[0]->cols.empty() || t[1]->cols.empty())
        return fail(GDB_E_INVALID, "join tables need a key column");
      uint64_t m = 0;
      rc = gdb_join_count(ctx, t[0], t[0]->cols[0].name.c_str(), t[1], t[1]->cols[0].name.c_str(),
                          &m, &tm);
      break;
    }
    default:
      return fail(GDB_E_INVALID, "unknown pipeline id");
  }
  if (rc == GDB_E_INVALID && n < 1) return fail(GDB_E_INVALID, "too few tables for the pipeline");
  *ms = tm.gpu_ms;
  return rc;
}

}  // namespace

extern "C" {

int gdb_get_tuning(gdb_ctx* ctx, int pipeline, gdb_tuning* out) {
  if (!ctx || !out || pipeline < 0 || pipeline >= 16) return fail(GDB_E_INVALID, "bad argument");
  const Tuning& t = ctx->tune[pipeline];
  *out = gdb_tuning{pipeline, t.block, t.blocks_per_cu, t.unroll, 0.0, 0, t.radix_bits};
  return GDB_OK;
}

int gdb_set_tuning(gdb_ctx* ctx, const gdb_tuning* t) {
  if (!ctx || !t || t->pipeline < 0 || t->pipeline >= 16) return fail(GDB_E_INVALID, "bad argument");
  if ((t->block != 64 && t->block != 128 && t->block != 256) || t->blocks_per_cu < 0 ||
      t->blocks_per_cu > 64 || (t->unroll != 1 && t->unroll != 2 && t->unroll != 4) ||
      t->radix_bits < -2 || t->radix_bits > 14)
    return fail(GDB_E_INVALID,
                "tuning out of range (block 64/128/256, unroll 1/2/4, radix bits -2..14)");
  ctx->tune[t->pipeline] = Tuning{t->block, t->blocks_per_cu, t->unroll, t->radix_bits};
  return GDB_OK;
}

int gdb_autotune(gdb_ctx* ctx, int pipeline, const gdb_table* const* tables, int ntables, int reps,
                 gdb_tuning* best) {
  if (!ctx || !tables || !best || reps < 1) return fail(GDB_E_INVALID, "bad argument");
  const std::vector<Shape> cands = candidates(pipeline, tables, ntables);
  if (cands.empty()) return fail(GDB_E_INVALID, "pipeline has no tunable kernel");
  const Tuning saved = ctx->tune[pipeline];
  double best_ms = 1e30;
  Shape best_s = {saved.block, saved.blocks_per_cu, saved.unroll, saved.radix_bits};
  for (const Shape& s : cands) {
    ctx->tune[pipeline] = Tuning{s.block, s.bpc, s.unroll, s.bits};
    double ms = 0;
    int rc = run_once(ctx, pipeline, tables, ntables, &ms);  // warm-up
    if (rc != GDB_OK) {
      ctx->tune[pipeline] = saved;
      return rc;
    }
    std::vector<double> v;
    for (int r = 0; r < reps; ++r) {
      GDB_TRY(run_once(ctx, pipeline, tables, ntables, &ms));
      v.push_back(ms);
    }
    std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
    const double med = v[v.size() / 2];
    if (med < best_ms) {
      best_ms = med;
      best_s = s;
    }
  }
  ctx->tune[pipeline] = Tuning{best_s.block, best_s.bpc, best_s.unroll, best_s.bits};
  *best = gdb_tuning{pipeline, best_s.block, best_s.bpc, best_s.unroll, best_ms,
                     (int32_t)cands.size(), best_s.bits};
  return GDB_OK;
}

}  // extern "C"
