// The serving scheduler (serve.cpp) on the CPU, over a fake executor.
//
// serve.cpp is compiled unchanged with g++ and linked against this file, which
// stands in for everything below the scheduler: the HIP stream/event calls (a
// simulated device where a launch completes a fixed time after the previous
// one on its stream), the batch executors (plan_flat_batch, flat_batch_launch,
// q3_batch_launch and their result readers) and the plain pipelines of serial
// mode. Each fake result encodes the parameters of the request it answers, so
// a response delivered to the wrong request is caught. The fake device also
// checks the scheduler's invariants as they happen: batches within max_batch
// and within the plan's fit, at most `slots` batches and one Q3 batch in
// flight, in-flight scratch within the HBM budget, one serial pipeline call at
// a time, and each family launched in submission order.
//
// Built under AddressSanitizer + UndefinedBehaviorSanitizer and, separately,
// ThreadSanitizer, with client threads submitting and polling concurrently.
//
// Usage: sanitize_serve
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"
#include "gdb.h"
#include "internal.h"

namespace {

using Clock = std::chrono::steady_clock;
double now_ms() {
  return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count();
}

int g_failures = 0;
std::mutex g_fail_mu;

#define CHECK(cond)                                                                          \
  do {                                                                                       \
    if (!(cond)) {                                                                           \
      std::lock_guard<std::mutex> g_(g_fail_mu);                                             \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond);         \
      ++g_failures;                                                                          \
    }                                                                                        \
  } while (0)

// ---------------------------------------------------------------- fake device
constexpr int MAX_CELLS = 40;        // plan fit: 2 cells per Q6, 6 per Q1 request
constexpr double LAUNCH_MS = 0.15;   // simulated duration of one batch
constexpr int32_t BAD_QTY = 77;      // a Q6 request with this qty_lt fails its launch
constexpr int32_t BAD_Q3_DATE = -77; // a Q3 request with this date fails its launch

struct FakeStream {
  double done_ms = 0;         // completion time of the last launch
  uint64_t pending_bytes = 0; // footprint of the launch not yet covered by an event
  int pending_family = -1;
};
struct FakeEvent {
  double ready_ms = 0;
  uint64_t bytes = 0;
  int family = -1;
  bool counted = false;  // the batch it closes is still in flight
};

struct Device {
  std::mutex mu;
  uint64_t budget = 0;
  int max_batch = 64, slots = 1;
  uint64_t inflight_bytes = 0;
  int inflight_flat = 0, inflight_q3 = 0;
  std::vector<std::vector<int32_t>> flat_order;  // request tags of each flat batch, in launch order
  std::vector<int64_t> q3_order;                 // Q3 request codes in launch order
  int serial_active = 0;
  std::vector<int64_t> serial_order;
  bool fail_slot_init = false;
  int flat_batches = 0, q3_batches = 0;
} dev;

void reset_device(uint64_t budget, int max_batch, int slots) {
  std::lock_guard<std::mutex> g(dev.mu);
  dev.budget = budget;
  dev.max_batch = max_batch;
  dev.slots = slots;
  dev.inflight_bytes = 0;
  dev.inflight_flat = dev.inflight_q3 = 0;
  dev.flat_order.clear();
  dev.q3_order.clear();
  dev.serial_order.clear();
  dev.serial_active = 0;
  dev.fail_slot_init = false;
  dev.flat_batches = dev.q3_batches = 0;
}

// The tag of a request is its unique parameter (Q6 date_lo, Q1 cutoff, Q3 date).
int64_t q6_code(const gdb_q6_params& p) {
  return ((int64_t)p.date_lo << 24) ^ ((int64_t)p.date_hi << 12) ^ (p.disc_lo << 8) ^
         (p.disc_hi << 4) ^ p.qty_lt;
}
int64_t q1_code(const gdb_q1_params& p) { return (int64_t)p.cutoff * 7 + 3; }
int64_t q3_code(const gdb_q3_params& p) { return (int64_t)p.segment * 1000000007ll + p.date; }

uint64_t flat_fp(int ncells, int blocks) {
  return (uint64_t)ncells * 8ull * (uint64_t)(blocks + 1) + sizeof(gdb::FlatBatchDesc);
}

}  // namespace

// ---------------------------------------------------------------- HIP stand-ins
extern "C" {
hipError_t hipSetDevice(int) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake device error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new FakeStream());
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete reinterpret_cast<FakeStream*>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  double until;
  {
    std::lock_guard<std::mutex> g(dev.mu);
    until = reinterpret_cast<FakeStream*>(s)->done_ms;
  }
  while (now_ms() < until) std::this_thread::sleep_for(std::chrono::microseconds(20));
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) {
  *e = reinterpret_cast<hipEvent_t>(new FakeEvent());
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<FakeEvent*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::lock_guard<std::mutex> g(dev.mu);
  FakeEvent* ev = reinterpret_cast<FakeEvent*>(e);
  FakeStream* st = reinterpret_cast<FakeStream*>(s);
  CHECK(!ev->counted);  // an event is re-recorded only after its batch finished
  ev->ready_ms = st->done_ms;
  ev->bytes = st->pending_bytes;
  ev->family = st->pending_family;
  ev->counted = st->pending_family >= 0;
  st->pending_bytes = 0;
  st->pending_family = -1;
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
  std::lock_guard<std::mutex> g(dev.mu);
  FakeEvent* ev = reinterpret_cast<FakeEvent*>(e);
  if (now_ms() < ev->ready_ms) return hipErrorNotReady;
  if (ev->counted) {
    ev->counted = false;
    dev.inflight_bytes -= ev->bytes;
    if (ev->family == 0) --dev.inflight_flat;
    else --dev.inflight_q3;
  }
  return hipSuccess;
}
}  // extern "C"

// ---------------------------------------------------------------- library stand-ins
namespace gdb {

struct Q3Slot {
  int launches = 0;
};

static thread_local std::string t_error;

void set_error(const std::string& msg) { t_error = msg; }

int fail(int code, const std::string& msg) {
  t_error = msg;
  return code;
}

int sum_fits(uint64_t, unsigned __int128, const char*) { return GDB_OK; }

int column(const gdb_table* t, const char* name, int type, const Column** out) {
  for (const Column& c : t->cols)
    if (c.name == name && c.type == type) {
      *out = &c;
      return GDB_OK;
    }
  return fail(GDB_E_INVALID, std::string("no column ") + name);
}

// The fake plan carries its requests' tags (Q6 date_lo, Q1 cutoff) in date_b
// and a failure flag per Q6 request in disc_b, for the launch to read.
bool plan_flat_batch(const std::vector<gdb_q6_params>& q6, const std::vector<gdb_q1_params>& q1,
                     FlatPlan* plan, size_t, bool fill_tables) {
  plan->ncells = 2 * (int)q6.size() + 6 * (int)q1.size();
  plan->lds_bytes = (size_t)plan->ncells * 8;
  plan->date_b.clear();
  plan->disc_b.clear();
  if (fill_tables) {
    for (const gdb_q6_params& p : q6) {
      plan->date_b.push_back(p.date_lo);
      plan->disc_b.push_back(p.qty_lt == BAD_QTY);
    }
    for (const gdb_q1_params& p : q1) plan->date_b.push_back(p.cutoff);
  }
  return plan->ncells <= MAX_CELLS;
}

int flat_slot_init(gdb_ctx*, FlatSlot* s) {
  std::lock_guard<std::mutex> g(dev.mu);
  if (dev.fail_slot_init) return fail(GDB_E_NOMEM, "fake slot allocation failure");
  s->blocks = 4;
  s->lds_limit = FB_LDS_LIMIT;
  s->h_totals = new int64_t[8]();
  return GDB_OK;
}

void flat_slot_free(gdb_ctx*, FlatSlot* s) {
  delete[] s->h_totals;
  s->h_totals = nullptr;
}

int flat_batch_launch(gdb_ctx*, FlatSlot* s, const FlatPlan& plan, const gdb_table*,
                      hipStream_t stream) {
  std::lock_guard<std::mutex> g(dev.mu);
  CHECK(plan.ncells > 0 && plan.ncells <= MAX_CELLS);
  CHECK(!plan.date_b.empty() && (int)plan.date_b.size() <= dev.max_batch);
  for (int32_t bad : plan.disc_b)
    if (bad) return fail(GDB_E_HIP, "fake flat launch failure");
  dev.flat_order.push_back(plan.date_b);
  FakeStream* st = reinterpret_cast<FakeStream*>(stream);
  const uint64_t fp = flat_fp(plan.ncells, s->blocks);
  if (dev.budget) CHECK(dev.inflight_bytes + fp <= dev.budget);
  dev.inflight_bytes += fp;
  ++dev.inflight_flat;
  ++dev.flat_batches;
  CHECK(dev.inflight_flat + dev.inflight_q3 <= dev.slots);
  st->done_ms = std::max(st->done_ms, now_ms()) + LAUNCH_MS;
  st->pending_bytes = fp;
  st->pending_family = 0;
  return GDB_OK;
}

void flat_batch_results(const FlatPlan& plan, const int64_t*, const std::vector<gdb_q6_params>& q6,
                        const std::vector<gdb_q1_params>& q1, std::vector<gdb_q6_result>* r6,
                        std::vector<gdb_q1_result>* r1) {
  CHECK(plan.ncells == 2 * (int)q6.size() + 6 * (int)q1.size());
  r6->assign(q6.size(), gdb_q6_result{});
  r1->assign(q1.size(), gdb_q1_result{});
  for (size_t i = 0; i < q6.size(); ++i) (*r6)[i] = gdb_q6_result{q6_code(q6[i]), q6.size() + q1.size()};
  for (size_t i = 0; i < q1.size(); ++i) {
    (*r1)[i].nrows = 1;
    (*r1)[i].rows[0].sum_qty = q1_code(q1[i]);
    (*r1)[i].rows[0].count = (int64_t)(q6.size() + q1.size());
  }
}

int q3_slot_create(gdb_ctx*, const gdb_table*, const gdb_table*, const gdb_table*, Q3Slot** out) {
  *out = new Q3Slot();
  return GDB_OK;
}
void q3_slot_use_graphs(Q3Slot*, bool) {}
void q3_slot_graph_counts(const Q3Slot*, uint64_t* c, uint64_t* r) { *c = *r = 0; }
void q3_slot_free(gdb_ctx*, Q3Slot* s) { delete s; }
uint64_t q3_slot_bytes(const Q3Slot* s) { return s ? 4096 : 0; }

int q3_batch_launch(gdb_ctx*, Q3Slot* s, const std::vector<gdb_q3_params>& reqs,
                    hipStream_t stream) {
  std::lock_guard<std::mutex> g(dev.mu);
  CHECK(!reqs.empty() && (int)reqs.size() <= dev.max_batch && (int)reqs.size() <= Q3B_MAX);
  for (const gdb_q3_params& p : reqs)
    if (p.date == BAD_Q3_DATE) return fail(GDB_E_HIP, "fake Q3 launch failure");
  for (const gdb_q3_params& p : reqs) dev.q3_order.push_back(q3_code(p));
  const uint64_t fp = 64ull * 1024 * reqs.size();
  if (dev.budget) CHECK(dev.inflight_bytes + fp <= dev.budget);
  CHECK(dev.inflight_q3 == 0);  // Q3 batches share one index slot
  dev.inflight_bytes += fp;
  ++dev.inflight_q3;
  ++dev.q3_batches;
  CHECK(dev.inflight_flat + dev.inflight_q3 <= dev.slots);
  ++s->launches;
  FakeStream* st = reinterpret_cast<FakeStream*>(stream);
  st->done_ms = std::max(st->done_ms, now_ms()) + LAUNCH_MS;
  st->pending_bytes = fp;
  st->pending_family = 1;
  return GDB_OK;
}

int q3_batch_results(Q3Slot*, const std::vector<gdb_q3_params>& reqs,
                     std::vector<gdb_q3_result>* out) {
  out->assign(reqs.size(), gdb_q3_result{});
  for (size_t i = 0; i < reqs.size(); ++i) {
    (*out)[i].groups = q3_code(reqs[i]);
    (*out)[i].nrows = reqs[i].top_k;
  }
  return GDB_OK;
}

}  // namespace gdb

// ---------------------------------------------------------------- plain pipelines
namespace {
void serial_enter(int64_t tag) {
  std::lock_guard<std::mutex> g(dev.mu);
  ++dev.serial_active;
  CHECK(dev.serial_active == 1);
  dev.serial_order.push_back(tag);
}
void serial_leave() {
  std::this_thread::sleep_for(std::chrono::microseconds(30));
  std::lock_guard<std::mutex> g(dev.mu);
  --dev.serial_active;
}
}  // namespace

extern "C" {
int gdb_q6(gdb_ctx*, const gdb_table*, const gdb_q6_params* p, gdb_q6_result* out, gdb_timing*) {
  serial_enter(p->date_lo);
  out->revenue = q6_code(*p);
  out->rows = 1;
  serial_leave();
  return p->qty_lt == BAD_QTY ? gdb::fail(GDB_E_HIP, "fake pipeline failure") : GDB_OK;
}
int gdb_q1(gdb_ctx*, const gdb_table*, const gdb_q1_params* p, gdb_q1_result* out, gdb_timing*) {
  serial_enter(p->cutoff);
  memset(out, 0, sizeof(*out));
  out->nrows = 1;
  out->rows[0].sum_qty = q1_code(*p);
  out->rows[0].count = 1;
  serial_leave();
  return GDB_OK;
}
int gdb_q3(gdb_ctx*, const gdb_table*, const gdb_table*, const gdb_table*, const gdb_q3_params* p,
           gdb_q3_result* out, gdb_timing*) {
  serial_enter(p->date);
  memset(out, 0, sizeof(*out));
  out->groups = q3_code(*p);
  out->nrows = p->top_k;
  serial_leave();
  return GDB_OK;
}
}  // extern "C"

// ---------------------------------------------------------------- scenarios
namespace {

struct Tables {
  gdb_table flat, cust, ord, li;
  Tables() {
    auto col = [](const char* n, int type) {
      gdb::Column c;
      c.name = n;
      c.type = type;
      c.minv = 0;
      c.maxv = 100;
      return c;
    };
    flat.nrows = 1000;
    flat.cols = {col("l_extendedprice", GDB_I64), col("l_discount", GDB_I32), col("l_tax", GDB_I32)};
  }
};

gdb_request make(int kind, int tag, int32_t qty = 24) {
  gdb_request r;
  memset(&r, 0, sizeof(r));
  r.kind = kind;
  if (kind == GDB_REQ_Q6) r.q6 = gdb_q6_params{tag, tag + 365, 4, 6, qty};
  else if (kind == GDB_REQ_Q1) r.q1 = gdb_q1_params{tag};
  else r.q3 = gdb_q3_params{tag % 5, tag, 10};
  return r;
}

bool matches(const gdb_request& q, const gdb_response& r) {
  if (r.kind != q.kind) return false;
  if (q.kind == GDB_REQ_Q6) return r.q6.revenue == q6_code(q.q6);
  if (q.kind == GDB_REQ_Q1) return r.q1.nrows == 1 && r.q1.rows[0].sum_qty == q1_code(q.q1);
  return r.q3.groups == q3_code(q.q3) && r.q3.nrows == q.q3.top_k;
}

// Submits `per_thread` requests from each of `threads` threads while one
// thread polls; returns id -> (request, response).
struct Run {
  std::map<uint64_t, gdb_request> req;
  std::map<uint64_t, gdb_response> resp;
  int duplicates = 0;
  std::vector<std::vector<uint64_t>> ids_by_thread;
};

Run drive(gdb_server* S, int threads, int per_thread, int (*kind_of)(int), int (*qty_of)(int)) {
  Run run;
  std::mutex mu;
  std::atomic<int> submitted{0};
  const int total = threads * per_thread;
  run.ids_by_thread.resize(threads);
  std::thread poller([&] {
    std::vector<gdb_response> buf(64);
    const double deadline = now_ms() + 60000;
    int got = 0;
    while (got < total && now_ms() < deadline) {
      const int n = gdb_poll(S, buf.data(), (int)buf.size(), 5);
      std::lock_guard<std::mutex> g(mu);
      for (int i = 0; i < n; ++i) {
        if (run.resp.count(buf[i].id)) ++run.duplicates;
        run.resp[buf[i].id] = buf[i];
      }
      got += n;
    }
  });
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t) {
    th.emplace_back([&, t] {
      std::mt19937 rng(1234 + t);
      for (int i = 0; i < per_thread; ++i) {
        const int tag = t * per_thread + i + 1;  // unique per request, > 0
        const gdb_request r = make(kind_of(tag), tag, qty_of ? qty_of(tag) : 24);
        uint64_t id = 0;
        const int rc = gdb_submit(S, &r, &id);
        CHECK(rc == GDB_OK);
        {
          std::lock_guard<std::mutex> g(mu);
          run.req[id] = r;
          run.ids_by_thread[t].push_back(id);
        }
        ++submitted;
        if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
      }
    });
  }
  for (auto& x : th) x.join();
  poller.join();
  return run;
}

int tag_of(const gdb_request& r) {
  return r.kind == GDB_REQ_Q6 ? r.q6.date_lo : r.kind == GDB_REQ_Q1 ? r.q1.cutoff : r.q3.date;
}

// Every request answered once, with its own result.
void check_answers(const Run& run, int total, bool all_ok) {
  CHECK(run.duplicates == 0);
  CHECK((int)run.req.size() == total);
  CHECK((int)run.resp.size() == total);
  for (const auto& kv : run.req) {
    auto it = run.resp.find(kv.first);
    CHECK(it != run.resp.end());
    if (it == run.resp.end()) continue;
    if (all_ok) CHECK(it->second.status == GDB_OK);
    if (it->second.status == GDB_OK) CHECK(matches(kv.second, it->second));
  }
  for (const auto& t : run.ids_by_thread)
    for (size_t i = 1; i < t.size(); ++i) CHECK(t[i] > t[i - 1]);
}

// Flat batches take the oldest pending flat requests: every request of a batch
// was submitted before every request of the next one.
void check_flat_fifo(const std::vector<std::vector<int32_t>>& batches, const Run& run) {
  std::map<int64_t, uint64_t> id_of_tag;
  for (const auto& kv : run.req)
    if (kv.second.kind != GDB_REQ_Q3) id_of_tag[tag_of(kv.second)] = kv.first;
  uint64_t prev_max = 0;
  for (const auto& b : batches) {
    uint64_t lo = UINT64_MAX, hi = 0;
    for (int32_t tag : b) {
      auto it = id_of_tag.find(tag);
      CHECK(it != id_of_tag.end());
      if (it == id_of_tag.end()) continue;
      lo = std::min(lo, it->second);
      hi = std::max(hi, it->second);
    }
    CHECK(lo > prev_max);
    prev_max = hi;
  }
}

// Q3 requests are launched in submission (id) order.
void check_fifo(const std::vector<int64_t>& order, const Run& run, int kind) {
  std::map<int64_t, uint64_t> id_of_tag;
  for (const auto& kv : run.req)
    if (kv.second.kind == kind) id_of_tag[kind == GDB_REQ_Q3 ? q3_code(kv.second.q3) : tag_of(kv.second)] = kv.first;
  uint64_t last = 0;
  for (int64_t tag : order) {
    auto it = id_of_tag.find(tag);
    CHECK(it != id_of_tag.end());
    if (it == id_of_tag.end()) continue;
    CHECK(it->second > last);
    last = it->second;
  }
}

int mixed_kind(int tag) { return tag % 10 < 5 ? GDB_REQ_Q6 : tag % 10 < 8 ? GDB_REQ_Q1 : GDB_REQ_Q3; }
int flat_kind(int tag) { return tag % 3 ? GDB_REQ_Q6 : GDB_REQ_Q1; }
int bad_every_37(int tag) { return tag % 37 == 0 ? BAD_QTY : 24; }

gdb_serve_cfg config(Tables& t, int mode, int max_batch, int slots, uint64_t budget, bool q3) {
  gdb_serve_cfg c;
  memset(&c, 0, sizeof(c));
  c.lineitem_flat = &t.flat;
  if (q3) {
    c.customer = &t.cust;
    c.orders = &t.ord;
    c.lineitem = &t.li;
  }
  c.mode = mode;
  c.max_batch = max_batch;
  c.slots = slots;
  c.hbm_budget = budget;
  return c;
}

void batched_mix(gdb_ctx* ctx, Tables& t) {
  reset_device(0, 16, 2);
  gdb_serve_cfg c = config(t, GDB_SERVE_BATCHED, 16, 2, 0, true);
  gdb_server* S = nullptr;
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  const Run run = drive(S, 8, 400, mixed_kind, nullptr);
  check_answers(run, 3200, true);
  int maxb = 0;
  for (const auto& kv : run.resp) maxb = std::max(maxb, kv.second.batch_size);
  CHECK(maxb >= 1 && maxb <= 16);
  gdb_serve_stats st;
  CHECK(gdb_serve_stats_get(S, &st) == GDB_OK);
  CHECK(st.submitted == 3200 && st.completed == 3200 && st.rejected == 0);
  CHECK(st.batches >= 1 && st.batches <= 3200 && st.mean_batch >= 1.0);
  CHECK(st.p50_ms <= st.p99_ms && st.p99_ms <= st.max_ms && st.qps > 0);
  CHECK(st.index_bytes == 4096);
  CHECK(gdb_serve_stop(S) == GDB_OK);
  std::lock_guard<std::mutex> g(dev.mu);
  check_fifo(dev.q3_order, run, GDB_REQ_Q3);
  check_flat_fifo(dev.flat_order, run);
  CHECK(dev.inflight_flat == 0 && dev.inflight_q3 == 0 && dev.inflight_bytes == 0);
  std::printf("batched mix: %d flat + %d Q3 batches for 3200 requests\n", dev.flat_batches,
              dev.q3_batches);
}

void budget_defers(gdb_ctx* ctx, Tables& t) {
  // Room for one batch of at most two Q6 requests: the rest wait for its scratch.
  const uint64_t budget = flat_fp(2, 4) + 100;
  reset_device(budget, 8, 4);
  gdb_serve_cfg c = config(t, GDB_SERVE_BATCHED, 8, 4, budget, true);
  gdb_server* S = nullptr;
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  // A Q3 request needs 64 KiB of scratch: refused at submission.
  gdb_request q3 = make(GDB_REQ_Q3, 5);
  uint64_t id = 0;
  CHECK(gdb_submit(S, &q3, &id) == GDB_E_ADMISSION);
  // A Q1 request alone (6 cells) does not fit either.
  gdb_request q1 = make(GDB_REQ_Q1, 5);
  CHECK(gdb_submit(S, &q1, &id) == GDB_E_ADMISSION);
  const Run run = drive(S, 4, 150, [](int) { return (int)GDB_REQ_Q6; }, nullptr);
  check_answers(run, 600, true);
  gdb_serve_stats st;
  CHECK(gdb_serve_stats_get(S, &st) == GDB_OK);
  CHECK(st.completed == 600 && st.rejected == 2);
  CHECK(st.inflight_peak_bytes <= budget);
  CHECK(gdb_serve_stop(S) == GDB_OK);
  std::printf("budget: 600 deferred-then-served requests, peak %llu of %llu B\n",
              (unsigned long long)st.inflight_peak_bytes, (unsigned long long)budget);
}

void launch_failures(gdb_ctx* ctx, Tables& t) {
  reset_device(0, 8, 2);
  gdb_serve_cfg c = config(t, GDB_SERVE_BATCHED, 8, 2, 0, true);
  gdb_server* S = nullptr;
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  // A Q6 request with qty_lt BAD_QTY fails its whole batch; the server goes on.
  const Run run = drive(S, 4, 200, [](int tag) { return tag % 4 == 0 ? (int)GDB_REQ_Q3 : (int)GDB_REQ_Q6; },
                        bad_every_37);
  check_answers(run, 800, false);
  int bad = 0, collateral = 0;
  for (const auto& kv : run.resp) {
    const gdb_request& q = run.req.at(kv.first);
    const bool is_bad = q.kind == GDB_REQ_Q6 && q.q6.qty_lt == BAD_QTY;
    bad += is_bad;
    if (is_bad) CHECK(kv.second.status == GDB_E_HIP);
    else if (kv.second.status != GDB_OK) ++collateral;
    CHECK(kv.second.status == GDB_OK || kv.second.status == GDB_E_HIP);
  }
  CHECK(bad > 0 && collateral <= bad * 7);  // batch-mates of a bad request, at most
  std::vector<uint64_t> ids;
  for (int i = 0; i < 50; ++i) {
    gdb_request r = make(GDB_REQ_Q3, 100000 + i);
    if (i % 10 == 3) r.q3.date = BAD_Q3_DATE;
    uint64_t id = 0;
    CHECK(gdb_submit(S, &r, &id) == GDB_OK);
    ids.push_back(id);
  }
  std::vector<gdb_response> buf(64);
  std::map<uint64_t, gdb_response> got;
  const double deadline = now_ms() + 30000;
  while (got.size() < ids.size() && now_ms() < deadline) {
    const int n = gdb_poll(S, buf.data(), (int)buf.size(), 5);
    for (int i = 0; i < n; ++i) got[buf[i].id] = buf[i];
  }
  CHECK(got.size() == ids.size());
  int failed = 0;
  for (const auto& kv : got) failed += kv.second.status == GDB_E_HIP;
  CHECK(failed >= 5);  // at least the bad requests themselves
  for (const auto& kv : got)
    CHECK(kv.second.status == GDB_OK || kv.second.status == GDB_E_HIP);
  gdb_serve_stats st;
  CHECK(gdb_serve_stats_get(S, &st) == GDB_OK);
  CHECK(st.completed + st.rejected == 850);
  CHECK(gdb_serve_stop(S) == GDB_OK);
  std::printf("launch failures: %d bad Q6 (+%d batch-mates), %d of 50 Q3 failed with their batch\n",
              bad, collateral, failed);
}

void serial_mode(gdb_ctx* ctx, Tables& t) {
  reset_device(0, 1, 1);
  gdb_serve_cfg c = config(t, GDB_SERVE_SERIAL, 1, 1, 0, true);
  gdb_server* S = nullptr;
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  const Run run = drive(S, 3, 150, mixed_kind, bad_every_37);
  check_answers(run, 450, false);
  for (const auto& kv : run.resp) {
    CHECK(kv.second.batch_size == 1);
    const gdb_request& q = run.req.at(kv.first);
    const bool bad = q.kind == GDB_REQ_Q6 && q.q6.qty_lt == BAD_QTY;
    CHECK(kv.second.status == (bad ? GDB_E_HIP : GDB_OK));
  }
  CHECK(gdb_serve_stop(S) == GDB_OK);
  // one pipeline call at a time, in submission order
  std::map<int64_t, uint64_t> id_of_tag;
  for (const auto& kv : run.req) id_of_tag[tag_of(kv.second)] = kv.first;
  std::lock_guard<std::mutex> g(dev.mu);
  CHECK(dev.serial_order.size() == 450);
  uint64_t last = 0;
  for (int64_t tag : dev.serial_order) {
    CHECK(id_of_tag.count(tag) && id_of_tag[tag] > last);
    last = id_of_tag[tag];
  }
  std::printf("serial: 450 requests one at a time in order\n");
}

void loadgen(gdb_ctx* ctx, Tables& t) {
  reset_device(0, 64, 2);
  gdb_serve_cfg c = config(t, GDB_SERVE_BATCHED, 64, 2, 0, true);
  gdb_server* S = nullptr;
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  gdb_serve_stats st;
  CHECK(gdb_serve_loadgen(S, 20000, 0.1, 7, 0.2, &st) == GDB_OK);
  CHECK(st.submitted == 2000 && st.completed == 2000 && st.rejected == 0);
  CHECK(st.mean_batch >= 1.0);
  CHECK(gdb_serve_stop(S) == GDB_OK);
  std::printf("loadgen: 2000 requests, mean batch %.1f, p99 %.3f ms\n", st.mean_batch, st.p99_ms);
}

void misuse(gdb_ctx* ctx, Tables& t) {
  reset_device(0, 8, 1);
  gdb_server* S = nullptr;
  gdb_serve_cfg c = config(t, GDB_SERVE_BATCHED, 0, 1, 0, false);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_E_INVALID && !S);
  c = config(t, GDB_SERVE_BATCHED, 65, 1, 0, false);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_E_INVALID);
  c = config(t, GDB_SERVE_BATCHED, 8, 5, 0, false);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_E_INVALID);
  c = config(t, 7, 8, 1, 0, false);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_E_INVALID);
  c = config(t, GDB_SERVE_BATCHED, 8, 1, 0, false);
  c.lineitem_flat = nullptr;
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_E_INVALID);
  CHECK(gdb_serve_start(nullptr, &c, &S) == GDB_E_INVALID);
  // A slot that cannot get its scratch fails the start and frees what it made.
  {
    std::lock_guard<std::mutex> g(dev.mu);
    dev.fail_slot_init = true;
  }
  c = config(t, GDB_SERVE_BATCHED, 8, 3, 0, true);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_E_NOMEM && !S);
  reset_device(0, 8, 1);
  // Requests the server cannot take.
  c = config(t, GDB_SERVE_BATCHED, 8, 1, 0, false);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  uint64_t id = 0;
  gdb_request r = make(GDB_REQ_Q3, 1);
  CHECK(gdb_submit(S, &r, &id) == GDB_E_INVALID);  // no Q3 tables
  r = make(GDB_REQ_Q6, 1);
  r.kind = 9;
  CHECK(gdb_submit(S, &r, &id) == GDB_E_INVALID);
  CHECK(gdb_submit(S, nullptr, &id) == GDB_E_INVALID);
  gdb_response out[4];
  CHECK(gdb_poll(S, out, 4, 1) == 0);  // nothing submitted: times out empty
  CHECK(gdb_poll(S, nullptr, 4, 0) == GDB_E_INVALID);
  CHECK(gdb_serve_stop(S) == GDB_OK);
  c = config(t, GDB_SERVE_BATCHED, 8, 1, 0, true);
  CHECK(gdb_serve_start(ctx, &c, &S) == GDB_OK);
  if (!S) return;
  r = make(GDB_REQ_Q3, 1);
  r.q3.top_k = 0;
  CHECK(gdb_submit(S, &r, &id) == GDB_E_INVALID);
  r.q3.top_k = 101;
  CHECK(gdb_submit(S, &r, &id) == GDB_E_INVALID);
  CHECK(gdb_serve_stop(S) == GDB_OK);
  CHECK(gdb_serve_stop(nullptr) == GDB_E_INVALID);
  std::printf("misuse: rejected as documented\n");
}

}  // namespace

int main() {
  gdb_ctx ctx;
  (void)hipStreamCreateWithFlags(&ctx.stream, 0);
  Tables t;
  misuse(&ctx, t);
  batched_mix(&ctx, t);
  budget_defers(&ctx, t);
  launch_failures(&ctx, t);
  serial_mode(&ctx, t);
  loadgen(&ctx, t);
  (void)hipStreamDestroy(ctx.stream);
  ctx.stream = nullptr;
  if (g_failures) {
    std::fprintf(stderr, "%d checks failed\n", g_failures);
    return 1;
  }
  std::printf("serve sanitize: all checks passed\n");
  return 0;
}
