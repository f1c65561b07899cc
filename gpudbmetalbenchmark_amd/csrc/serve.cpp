// Serving loop with continuous batching (no counterpart in the reference).
//
// Clients submit parameterised Q6 / Q1 / Q3 requests from any thread. One
// scheduler thread owns the GPU: whenever a batch slot is free it takes the
// oldest pending request's family and every pending request of that family
// (up to max_batch, FIFO), admits the batch if its HBM footprint fits the
// budget next to the batches already in flight, and launches it on the slot's
// stream without waiting; completions are detected by events, so requests that
// arrive while batches run are batched as soon as a slot frees up, and work of
// different slots overlaps on the device. Q6 and Q1 requests share one fused
// pass over the fact table (batch_flat.hip); Q3 requests run together over the
// resident join indexes (batch_q3.hip). Serial mode runs each request alone
// through gdb_q6 / gdb_q1 / gdb_q3, one at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"
#include "gen_core.h"
#include "internal.h"

using namespace gdb;

namespace {

using Clock = std::chrono::steady_clock;
double now_ms() {
  return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count();
}

enum Family { F_FLAT = 0, F_Q3 = 1 };
int family_of(int kind) { return kind == GDB_REQ_Q3 ? F_Q3 : F_FLAT; }

struct Pending {
  uint64_t id;
  gdb_request req;
  double t_submit;
};

struct Batch {
  int family = F_FLAT;
  std::vector<Pending> reqs;
  FlatPlan plan;
  std::vector<gdb_q6_params> q6;
  std::vector<gdb_q1_params> q1;
  std::vector<int> slot_of;  // request -> index in q6/q1/q3
  std::vector<gdb_q3_params> q3;
  uint64_t footprint = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  Batch batch;
  FlatSlot flat;
};

}  // namespace

struct gdb_server {
  gdb_ctx* ctx = nullptr;
  gdb_serve_cfg cfg{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<Pending> pending;
  std::deque<gdb_response> done;
  bool stopping = false;
  std::vector<Slot> slots;
  Q3Slot* q3 = nullptr;
  bool q3_busy = false;
  uint64_t next_id = 1;
  uint64_t reserved = 0, peak = 0;
  double index_build_ms = 0;  // Q3 join indexes: build time at start
  int fatal = GDB_OK;
  // statistics
  std::vector<double> lat;
  uint64_t submitted = 0, completed = 0, rejected = 0, batches = 0, batch_reqs = 0;
  double t_first = 0, t_last = 0;
};

namespace {

uint64_t flat_footprint(const FlatPlan& p, const FlatSlot& s) {
  return (uint64_t)p.ncells * 8ull * (uint64_t)(s.blocks + 1) + sizeof(FlatBatchDesc);
}
uint64_t q3_footprint(const gdb_q3_params& p, const Q3Slot* s);

void respond(gdb_server* S, const Pending& p, int status, int bsize, const gdb_q6_result* r6,
             const gdb_q1_result* r1, const gdb_q3_result* r3) {
  gdb_response r;
  memset(&r, 0, sizeof(r));
  r.id = p.id;
  r.kind = p.req.kind;
  r.status = status;
  const double t = now_ms();
  r.latency_ms = t - p.t_submit;
  r.batch_size = bsize;
  if (r6) r.q6 = *r6;
  if (r1) r.q1 = *r1;
  if (r3) r.q3 = *r3;
  S->done.push_back(r);
  if (status == GDB_OK) {
    S->lat.push_back(r.latency_ms);
    ++S->completed;
  } else {
    ++S->rejected;
  }
  S->t_last = std::max(S->t_last, t);
}

// Builds a batch of the oldest request's family from the pending queue
// (caller holds the lock). Returns false if nothing can be admitted now.
bool form_batch(gdb_server* S, Slot& slot, Batch* b) {
  if (S->pending.empty()) return false;
  const int fam = family_of(S->pending.front().req.kind);
  if (fam == F_Q3 && (S->q3_busy || !S->q3)) {
    // Q3 batches run one at a time on the shared index; try the flat family.
    bool any_flat = false;
    for (const auto& p : S->pending) any_flat |= family_of(p.req.kind) == F_FLAT;
    if (!any_flat) return false;
  }
  const int want = (fam == F_Q3 && (S->q3_busy || !S->q3)) ? F_FLAT : fam;
  *b = Batch();
  b->family = want;
  const uint64_t budget = S->cfg.hbm_budget ? S->cfg.hbm_budget : UINT64_MAX;
  std::vector<size_t> take;
  if (want == F_FLAT) {
    std::vector<gdb_q6_params> q6;
    std::vector<gdb_q1_params> q1;
    FlatPlan trial;
    for (size_t i = 0; i < S->pending.size() && (int)take.size() < S->cfg.max_batch; ++i) {
      const gdb_request& r = S->pending[i].req;
      if (family_of(r.kind) != F_FLAT) continue;
      if (r.kind == GDB_REQ_Q6) q6.push_back(r.q6);
      else q1.push_back(r.q1);
      if (!plan_flat_batch(q6, q1, &trial, slot.flat.lds_limit, false) ||
          S->reserved + flat_footprint(trial, slot.flat) > budget) {
        if (r.kind == GDB_REQ_Q6) q6.pop_back();
        else q1.pop_back();
        break;
      }
      take.push_back(i);
    }
    if (take.empty()) return false;
    for (size_t i : take) {
      const gdb_request& r = S->pending[i].req;
      if (r.kind == GDB_REQ_Q6) {
        b->slot_of.push_back((int)b->q6.size());
        b->q6.push_back(r.q6);
      } else {
        b->slot_of.push_back((int)b->q1.size());
        b->q1.push_back(r.q1);
      }
    }
    plan_flat_batch(b->q6, b->q1, &b->plan, slot.flat.lds_limit);
    b->footprint = flat_footprint(b->plan, slot.flat);
  } else {
    uint64_t fp = 0;
    for (size_t i = 0; i < S->pending.size() && (int)take.size() < S->cfg.max_batch; ++i) {
      const gdb_request& r = S->pending[i].req;
      if (r.kind != GDB_REQ_Q3) continue;
      const uint64_t f = q3_footprint(r.q3, S->q3);
      if (S->reserved + fp + f > budget) break;
      fp += f;
      take.push_back(i);
      b->slot_of.push_back((int)b->q3.size());
      b->q3.push_back(r.q3);
    }
    if (take.empty()) return false;
    b->footprint = fp;
  }
  // Move the taken requests out of the queue (indices ascending).
  std::deque<Pending> rest;
  size_t t = 0;
  for (size_t i = 0; i < S->pending.size(); ++i) {
    if (t < take.size() && take[t] == i) {
      b->reqs.push_back(S->pending[i]);
      ++t;
    } else {
      rest.push_back(S->pending[i]);
    }
  }
  S->pending.swap(rest);
  return true;
}

void finish_batch(gdb_server* S, Slot& slot) {
  Batch& b = slot.batch;
  const int n = (int)b.reqs.size();
  if (b.family == F_FLAT) {
    std::vector<gdb_q6_result> r6;
    std::vector<gdb_q1_result> r1;
    flat_batch_results(b.plan, slot.flat.h_totals, b.q6, b.q1, &r6, &r1);
    for (int i = 0; i < n; ++i) {
      const Pending& p = b.reqs[i];
      if (p.req.kind == GDB_REQ_Q6) respond(S, p, GDB_OK, n, &r6[b.slot_of[i]], nullptr, nullptr);
      else respond(S, p, GDB_OK, n, nullptr, &r1[b.slot_of[i]], nullptr);
    }
  } else {
    std::vector<gdb_q3_result> r3;
    q3_batch_results(S->q3, b.q3, &r3);
    for (int i = 0; i < n; ++i) respond(S, b.reqs[i], GDB_OK, n, nullptr, nullptr, &r3[b.slot_of[i]]);
    S->q3_busy = false;
  }
  S->reserved -= std::min(S->reserved, b.footprint);
  ++S->batches;
  S->batch_reqs += n;
  slot.busy = false;
  slot.batch = Batch();
}

int launch_batch(gdb_server* S, Slot& slot) {
  Batch& b = slot.batch;
  if (b.family == F_FLAT) {
    GDB_TRY(flat_batch_launch(S->ctx, &slot.flat, b.plan, S->cfg.lineitem_flat, slot.stream));
  } else {
    GDB_TRY(q3_batch_launch(S->ctx, S->q3, b.q3, slot.stream));
  }
  GDB_HIP(hipEventRecord(slot.done, slot.stream));
  return GDB_OK;
}

void run_serial(gdb_server* S, const Pending& p) {
  gdb_q6_result r6{};
  gdb_q1_result r1{};
  gdb_q3_result r3{};
  int rc = GDB_E_INVALID;
  switch (p.req.kind) {
    case GDB_REQ_Q6: rc = gdb_q6(S->ctx, S->cfg.lineitem_flat, &p.req.q6, &r6, nullptr); break;
    case GDB_REQ_Q1: rc = gdb_q1(S->ctx, S->cfg.lineitem_flat, &p.req.q1, &r1, nullptr); break;
    case GDB_REQ_Q3:
      rc = gdb_q3(S->ctx, S->cfg.customer, S->cfg.orders, S->cfg.lineitem, &p.req.q3, &r3, nullptr);
      break;
  }
  std::lock_guard<std::mutex> g(S->mu);
  respond(S, p, rc, 1, &r6, &r1, &r3);
  ++S->batches;
  ++S->batch_reqs;
  S->cv_done.notify_all();
}

void scheduler(gdb_server* S) {
  (void)hipSetDevice(S->ctx->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    if (S->cfg.mode == GDB_SERVE_SERIAL) {
      S->cv_sched.wait(lk, [&] { return S->stopping || !S->pending.empty(); });
      if (S->pending.empty()) break;
      Pending p = S->pending.front();
      S->pending.pop_front();
      lk.unlock();
      run_serial(S, p);
      lk.lock();
      continue;
    }
    bool progressed = false;
    for (Slot& s : S->slots) {
      if (s.busy && hipEventQuery(s.done) == hipSuccess) {
        finish_batch(S, s);
        progressed = true;
        S->cv_done.notify_all();
      }
    }
    bool busy = false;
    for (Slot& s : S->slots) busy |= s.busy;
    if (S->stopping && S->pending.empty() && !busy) break;
    for (Slot& s : S->slots) {
      if (s.busy || S->fatal != GDB_OK) continue;
      Batch b;
      if (!form_batch(S, s, &b)) break;
      s.batch = std::move(b);
      s.busy = true;
      if (s.batch.family == F_Q3) S->q3_busy = true;
      S->reserved += s.batch.footprint;
      S->peak = std::max(S->peak, S->reserved);
      const int rc = launch_batch(S, s);
      if (rc != GDB_OK) {
        // Fail the batch's requests; the server keeps running.
        for (const Pending& p : s.batch.reqs) respond(S, p, rc, 0, nullptr, nullptr, nullptr);
        if (s.batch.family == F_Q3) S->q3_busy = false;
        S->reserved -= std::min(S->reserved, s.batch.footprint);
        s.busy = false;
        s.batch = Batch();
        S->cv_done.notify_all();
      }
      progressed = true;
    }
    if (!progressed) {
      bool inflight = false;
      for (Slot& s : S->slots) inflight |= s.busy;
      if (inflight)
        S->cv_sched.wait_for(lk, std::chrono::microseconds(20));
      else
        S->cv_sched.wait(lk, [&] { return S->stopping || !S->pending.empty(); });
    }
  }
}

uint64_t q3_footprint(const gdb_q3_params& p, const Q3Slot* s) {
  (void)p;
  (void)s;
  return 64ull * 1024;  // per-request share of the fixed Q3 slot scratch
}

}  // namespace

extern "C" {

int gdb_serve_start(gdb_ctx* ctx, const gdb_serve_cfg* cfg, gdb_server** out) {
  if (!ctx || !cfg || !out) return fail(GDB_E_INVALID, "null argument");
  *out = nullptr;
  if (!cfg->lineitem_flat && !cfg->lineitem) return fail(GDB_E_INVALID, "no tables to serve");
  if (cfg->max_batch < 1 || cfg->max_batch > Q3B_MAX || cfg->slots < 1 || cfg->slots > 4 ||
      (cfg->mode != GDB_SERVE_BATCHED && cfg->mode != GDB_SERVE_SERIAL))
    return fail(GDB_E_INVALID, "serve config out of range (max_batch 1..64, slots 1..4)");
  if (cfg->lineitem_flat) {
    // The batched passes sum the same int64 terms as gdb_q6 and gdb_q1.
    const gdb_table* t = cfg->lineitem_flat;
    const Column *price, *disc, *tax;
    GDB_TRY(column(t, "l_extendedprice", GDB_I64, &price));
    GDB_TRY(column(t, "l_discount", GDB_I32, &disc));
    GDB_TRY(sum_fits(t->nrows, (unsigned __int128)magnitude(price) * magnitude(disc),
                     "Q6 SUM(extprice*discount)"));
    if (column(t, "l_tax", GDB_I32, &tax) == GDB_OK)
      GDB_TRY(sum_fits(t->nrows,
                       (unsigned __int128)magnitude(price) * (100u + magnitude(disc)) *
                           std::max<uint64_t>(1, magnitude(tax)),
                       "Q1 SUM(extprice*(100-disc)*tax)"));
  }
  GDB_HIP(hipSetDevice(ctx->device));
  gdb_server* S = new gdb_server();
  S->ctx = ctx;
  S->cfg = *cfg;
  auto bail = [&](int rc) {
    for (Slot& s : S->slots) {
      flat_slot_free(ctx, &s.flat);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream && s.stream != ctx->stream) (void)hipStreamDestroy(s.stream);
    }
    q3_slot_free(ctx, S->q3);
    delete S;
    return rc;
  };
  if (cfg->mode == GDB_SERVE_BATCHED) {
    const bool host_flat = cfg->lineitem_flat && cfg->lineitem_flat->host;
    const int nslots = host_flat ? 1 : cfg->slots;  // host tables stream on the engine stream
    S->slots.resize(nslots);
    for (Slot& s : S->slots) {
      if (host_flat) {
        s.stream = ctx->stream;
      } else if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) {
        return bail(fail(GDB_E_HIP, "stream creation failed"));
      }
      if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
        return bail(fail(GDB_E_HIP, "event creation failed"));
      int rc = flat_slot_init(ctx, &s.flat);
      if (rc != GDB_OK) return bail(rc);
      s.flat.graphs = cfg->graphs != 0;
    }
    if (cfg->customer && cfg->orders && cfg->lineitem) {
      const double t0 = now_ms();
      int rc = q3_slot_create(ctx, cfg->customer, cfg->orders, cfg->lineitem, &S->q3);
      if (rc != GDB_OK) return bail(rc);
      S->index_build_ms = now_ms() - t0;
      q3_slot_use_graphs(S->q3, cfg->graphs != 0);
    }
  }
  S->th = std::thread(scheduler, S);
  *out = S;
  return GDB_OK;
}

int gdb_submit(gdb_server* S, const gdb_request* req, uint64_t* id) {
  if (!S || !req) return fail(GDB_E_INVALID, "null argument");
  if (req->kind != GDB_REQ_Q6 && req->kind != GDB_REQ_Q1 && req->kind != GDB_REQ_Q3)
    return fail(GDB_E_INVALID, "unknown request kind");
  if ((req->kind == GDB_REQ_Q3 && !(S->cfg.customer && S->cfg.orders && S->cfg.lineitem)) ||
      (req->kind != GDB_REQ_Q3 && !S->cfg.lineitem_flat))
    return fail(GDB_E_INVALID, "the server has no tables for this request kind");
  if (req->kind == GDB_REQ_Q3 && (req->q3.top_k < 1 || req->q3.top_k > 100))
    return fail(GDB_E_INVALID, "top_k must be in 1..100");
  std::lock_guard<std::mutex> g(S->mu);
  if (S->stopping) return fail(GDB_E_CANCELLED, "server is stopping");
  if (S->cfg.mode == GDB_SERVE_BATCHED && S->cfg.hbm_budget) {
    // A request that cannot fit even alone is rejected, not queued forever.
    uint64_t fp = 0;
    if (req->kind == GDB_REQ_Q3) {
      fp = q3_footprint(req->q3, S->q3);
    } else if (!S->slots.empty()) {
      FlatPlan p;
      std::vector<gdb_q6_params> a;
      std::vector<gdb_q1_params> b;
      if (req->kind == GDB_REQ_Q6) a.push_back(req->q6);
      else b.push_back(req->q1);
      if (!plan_flat_batch(a, b, &p, S->slots[0].flat.lds_limit, false))
        return fail(GDB_E_UNSUPPORTED, "request parameters outside the batched domain");
      fp = flat_footprint(p, S->slots[0].flat);
    }
    if (fp > S->cfg.hbm_budget) {
      ++S->rejected;
      return fail(GDB_E_ADMISSION, "request footprint " + std::to_string(fp) +
                                       " B exceeds the serving HBM budget");
    }
  }
  Pending p;
  p.id = S->next_id++;
  p.req = *req;
  p.t_submit = now_ms();
  if (S->submitted == 0) S->t_first = p.t_submit;
  ++S->submitted;
  S->pending.push_back(p);
  if (id) *id = p.id;
  S->cv_sched.notify_one();
  return GDB_OK;
}

int gdb_poll(gdb_server* S, gdb_response* out, int max, int timeout_ms) {
  if (!S || (!out && max > 0)) return fail(GDB_E_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done.empty() && timeout_ms > 0)
    S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !S->done.empty(); });
  int n = 0;
  while (n < max && !S->done.empty()) {
    out[n++] = S->done.front();
    S->done.pop_front();
  }
  return n;
}

static void fill_stats(const std::vector<double>& lat_in, uint64_t submitted, uint64_t completed,
                       uint64_t rejected, uint64_t batches, uint64_t breqs, double t0, double t1,
                       uint64_t peak, gdb_serve_stats* o) {
  memset(o, 0, sizeof(*o));
  o->submitted = submitted;
  o->completed = completed;
  o->rejected = rejected;
  o->batches = batches;
  o->mean_batch = batches ? (double)breqs / (double)batches : 0.0;
  o->inflight_peak_bytes = peak;
  if (!lat_in.empty()) {
    std::vector<double> lat = lat_in;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) { return lat[std::min(lat.size() - 1, (size_t)std::floor(q * (lat.size() - 1) + 0.5))]; };
    o->p50_ms = pct(0.50);
    o->p99_ms = pct(0.99);
    o->max_ms = lat.back();
    double s = 0;
    for (double x : lat) s += x;
    o->mean_ms = s / lat.size();
  }
  o->qps = (t1 > t0) ? completed / ((t1 - t0) / 1000.0) : 0.0;
}

static void graph_counts(const gdb_server* S, uint64_t* captures, uint64_t* replays) {
  *captures = *replays = 0;
  for (const Slot& s : S->slots) {
    *captures += s.flat.gc.captures;
    *replays += s.flat.gc.replays;
  }
  if (S->q3) {
    uint64_t c = 0, r = 0;
    q3_slot_graph_counts(S->q3, &c, &r);
    *captures += c;
    *replays += r;
  }
}

int gdb_serve_stats_get(gdb_server* S, gdb_serve_stats* out) {
  if (!S || !out) return fail(GDB_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(S->mu);
  fill_stats(S->lat, S->submitted, S->completed, S->rejected, S->batches, S->batch_reqs, S->t_first,
             S->t_last, S->peak, out);
  graph_counts(S, &out->graph_captures, &out->graph_replays);
  out->index_bytes = q3_slot_bytes(S->q3);
  out->index_build_ms = S->index_build_ms;
  return S->fatal;
}

int gdb_serve_stop(gdb_server* S) {
  if (!S) return fail(GDB_E_INVALID, "null argument");
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stopping = true;
    S->cv_sched.notify_all();
  }
  if (S->th.joinable()) S->th.join();
  gdb_ctx* ctx = S->ctx;
  for (Slot& s : S->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    flat_slot_free(ctx, &s.flat);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream && s.stream != ctx->stream) (void)hipStreamDestroy(s.stream);
  }
  q3_slot_free(ctx, S->q3);
  const int rc = S->fatal;
  delete S;
  return rc;
}

void gdb_loadgen_request(uint64_t seed, uint64_t i, double q3_share, gdb_request* out) {
  using namespace gdbgen;
  memset(out, 0, sizeof(*out));
  const uint64_t base = stream_base(seed, 90);
  const double u = (double)(draw(base, i * 8) >> 11) / 9007199254740992.0;
  const double pq6 = (1.0 - q3_share) * 0.625;
  const int32_t start = (int32_t)uni(draw(base, i * 8 + 1), 0, 2160);
  if (u < pq6) {
    const int32_t c = (int32_t)uni(draw(base, i * 8 + 2), 2, 9);
    out->kind = GDB_REQ_Q6;
    out->q6 = gdb_q6_params{start, start + 365, c - 1, c + 1,
                            (int32_t)uni(draw(base, i * 8 + 3), 24, 25)};
  } else if (u < 1.0 - q3_share) {
    out->kind = GDB_REQ_Q1;
    out->q1 = gdb_q1_params{start + 365};
  } else {
    out->kind = GDB_REQ_Q3;
    out->q3 = gdb_q3_params{(int32_t)uni(draw(base, i * 8 + 4), 0, 4), start + 182, 10};
  }
}

int gdb_serve_loadgen(gdb_server* S, double rate, double seconds, uint64_t seed, double q3_share,
                      gdb_serve_stats* stats) {
  if (!S || !stats || rate <= 0 || seconds <= 0) return fail(GDB_E_INVALID, "bad argument");
  const uint64_t n = (uint64_t)std::max(1.0, rate * seconds);
  const uint64_t eb = gdbgen::stream_base(seed, 91);
  uint64_t batches0, breqs0;
  {
    std::lock_guard<std::mutex> g(S->mu);
    batches0 = S->batches;
    breqs0 = S->batch_reqs;
  }
  std::vector<uint64_t> ids;
  ids.reserve(n);
  const double t0 = now_ms();
  double t = 0;
  uint64_t rejected = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const double u = ((double)(gdbgen::draw(eb, i) >> 11) + 0.5) / 9007199254740992.0;
    t += -std::log(u) / rate * 1000.0;  // exponential inter-arrival, ms
    const double target = t0 + t;
    while (now_ms() < target) {
      const double wait = target - now_ms();
      if (wait > 0.2) std::this_thread::sleep_for(std::chrono::microseconds((int)(wait * 500)));
    }
    gdb_request r;
    gdb_loadgen_request(seed, i, q3_share, &r);
    uint64_t id = 0;
    const int rc = gdb_submit(S, &r, &id);
    if (rc == GDB_OK) ids.push_back(id);
    else ++rejected;
  }
  std::vector<double> lat;
  lat.reserve(ids.size());
  uint64_t got = 0, failed = 0;
  double t_last = t0;
  std::vector<gdb_response> buf(256);
  const double deadline = now_ms() + 120000.0;
  while (got + failed < ids.size() && now_ms() < deadline) {
    const int k = gdb_poll(S, buf.data(), (int)buf.size(), 100);
    for (int j = 0; j < k; ++j) {
      if (buf[j].status == GDB_OK) {
        lat.push_back(buf[j].latency_ms);
        ++got;
      } else {
        ++failed;
      }
    }
    t_last = now_ms();
  }
  uint64_t batches, breqs;
  {
    std::lock_guard<std::mutex> g(S->mu);
    batches = S->batches - batches0;
    breqs = S->batch_reqs - breqs0;
    t_last = S->t_last;
  }
  fill_stats(lat, n, got, rejected + failed, batches, breqs, t0, t_last, S->peak, stats);
  {
    std::lock_guard<std::mutex> g(S->mu);
    graph_counts(S, &stats->graph_captures, &stats->graph_replays);
    stats->index_bytes = q3_slot_bytes(S->q3);
    stats->index_build_ms = S->index_build_ms;
  }
  if (got + failed < ids.size()) return fail(GDB_E_TIMEOUT, "responses missing after 120 s");
  return GDB_OK;
}

}  // extern "C"
