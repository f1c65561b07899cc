/*
 * oracle.c — CPU restatement of the reference's Q6, Q1, Q3, Q9, Q13 and the
 * selection / aggregation / join primitives, in exact integer arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker and the timed CPU
 * baseline, never part of the engine.
 *
 * Each function follows the reference's SQL and the reference kernels'
 * predicate forms, cited per function. Where the reference deviates from its
 * own SQL (float32 money, per-row rounding in q1.metal:84-89, the 256-bin cap
 * of q13.metal:145) the SQL wins.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- threads */
typedef void (*range_fn)(void* arg, uint64_t lo, uint64_t hi, int tid);
typedef struct {
  range_fn fn;
  void* arg;
  uint64_t lo, hi;
  int tid;
} job_t;

static void* job_main(void* p) {
  job_t* j = (job_t*)p;
  j->fn(j->arg, j->lo, j->hi, j->tid);
  return NULL;
}

/* Runs fn over [0, n) split into `nthreads` contiguous ranges; range t gets tid t. */
static int par_for(uint64_t n, int nthreads, range_fn fn, void* arg) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (n < 4096) nthreads = 1;
  if (nthreads == 1) {
    fn(arg, 0, n, 0);
    return 1;
  }
  pthread_t th[256];
  job_t jobs[256];
  uint64_t per = (n + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
  for (int t = 0; t < nthreads; ++t) {
    uint64_t lo = (uint64_t)t * per, hi = lo + per;
    if (lo > n) lo = n;
    if (hi > n) hi = n;
    jobs[t] = (job_t){fn, arg, lo, hi, t};
    pthread_create(&th[t], NULL, job_main, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return nthreads;
}

/* ---------------------------------------------------------------- strings */
static int field_len(const uint8_t* s, int width) {
  int n = 0;
  while (n < width && s[n] != 0) ++n;
  return n;
}

static int find_from(const uint8_t* s, int len, int from, const char* pat) {
  int m = (int)strlen(pat);
  for (int i = from; i + m <= len; ++i)
    if (memcmp(s + i, pat, (size_t)m) == 0) return i;
  return -1;
}

/* SQL `s LIKE '%pattern%'` on a NUL-padded fixed field (q9.metal:61-83 tests
 * the same substring with a SWAR pre-filter). */
int oracle_contains(const uint8_t* s, int width, const char* pattern) {
  return find_from(s, field_len(s, width), 0, pattern) >= 0;
}

/* SQL `s LIKE '%w1%w2%'`: the leftmost w1, then w2 starting at or after its
 * end (q13.metal:43-70). */
int oracle_like_two(const uint8_t* s, int width, const char* w1, const char* w2) {
  int len = field_len(s, width);
  int a = find_from(s, len, 0, w1);
  if (a < 0) return 0;
  return find_from(s, len, a + (int)strlen(w1), w2) >= 0;
}

/* Year of a day number, day 0 = 1992-01-01, by the days-from-civil inverse
 * (H. Hinnant, "chrono-Compatible Low-Level Date Algorithms"). Stands in for
 * EXTRACT(YEAR FROM o_orderdate) of sql/q9.sql. */
int32_t oracle_year_of_day(int32_t day) {
  int64_t z = (int64_t)day + 8035 /* 1992-01-01 - 1970-01-01 */ + 719468;
  int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  int64_t doe = z - era * 146097;
  int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  int64_t y = yoe + era * 400;
  int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  int64_t mp = (5 * doy + 2) / 153;
  int64_t m = mp < 10 ? mp + 3 : mp - 9;
  return (int32_t)(y + (m <= 2));
}

/* --------------------------------------------------------------------- Q6 */
/* sql/q6.sql:5-13; predicate forms of kernels/q6.metal:37-41 (>= / <,
 * BETWEEN inclusive, quantity <). */
typedef struct {
  const int32_t *ship, *disc, *qty;
  const int64_t* price;
  int32_t dlo, dhi, clo, chi, qlt;
  int64_t rev[256];
  uint64_t rows[256];
} q6_arg;

static void q6_range(void* p, uint64_t lo, uint64_t hi, int tid) {
  q6_arg* a = (q6_arg*)p;
  int64_t s = 0;
  uint64_t c = 0;
  for (uint64_t i = lo; i < hi; ++i) {
    if (a->ship[i] >= a->dlo && a->ship[i] < a->dhi && a->disc[i] >= a->clo &&
        a->disc[i] <= a->chi && a->qty[i] < a->qlt) {
      s += a->price[i] * (int64_t)a->disc[i];
      ++c;
    }
  }
  a->rev[tid] = s;
  a->rows[tid] = c;
}

void oracle_q6(uint64_t n, const int32_t* shipdate, const int32_t* discount,
               const int32_t* quantity, const int64_t* extprice, int32_t date_lo,
               int32_t date_hi, int32_t disc_lo, int32_t disc_hi, int32_t qty_lt, int nthreads,
               int64_t* revenue, uint64_t* rows) {
  q6_arg* a = (q6_arg*)calloc(1, sizeof(q6_arg));
  a->ship = shipdate; a->disc = discount; a->qty = quantity; a->price = extprice;
  a->dlo = date_lo; a->dhi = date_hi; a->clo = disc_lo; a->chi = disc_hi; a->qlt = qty_lt;
  int T = par_for(n, nthreads, q6_range, a);
  int64_t s = 0;
  uint64_t c = 0;
  for (int t = 0; t < T; ++t) { s += a->rev[t]; c += a->rows[t]; }
  *revenue = s;
  *rows = c;
  free(a);
}

/* --------------------------------------------------------------------- Q1 */
/* sql/q1.sql:5-25; group codes of kernels/q1.metal:27-32. sum_disc_price and
 * sum_charge are the exact decimal products of the SQL, not the per-row
 * rounded values of kernels/q1.metal:88-89. */
typedef struct {
  const int32_t *ship, *qty, *disc, *tax;
  const int64_t* price;
  const uint8_t *rf, *ls;
  int32_t cutoff;
  int64_t acc[256][30];
  __int128 charge[256][6];
} q1_arg;

static int rf_code(uint8_t c) { return c == 'A' ? 0 : c == 'N' ? 1 : c == 'R' ? 2 : -1; }
static int ls_code(uint8_t c) { return c == 'F' ? 0 : c == 'O' ? 1 : -1; }

static void q1_range(void* p, uint64_t lo, uint64_t hi, int tid) {
  q1_arg* a = (q1_arg*)p;
  int64_t acc[30] = {0};
  __int128 ch[6] = {0};
  for (uint64_t i = lo; i < hi; ++i) {
    if (a->ship[i] > a->cutoff) continue;
    int r = rf_code(a->rf[i]), l = ls_code(a->ls[i]);
    if (r < 0 || l < 0) continue;
    int g = r * 2 + l;
    int64_t dp = a->price[i] * (int64_t)(100 - a->disc[i]);
    acc[g * 5 + 0] += a->qty[i];
    acc[g * 5 + 1] += a->price[i];
    acc[g * 5 + 2] += dp;
    acc[g * 5 + 3] += a->disc[i];
    acc[g * 5 + 4] += 1;
    ch[g] += (__int128)dp * (int64_t)(100 + a->tax[i]);
  }
  memcpy(a->acc[tid], acc, sizeof(acc));
  memcpy(a->charge[tid], ch, sizeof(ch));
}

int oracle_q1(uint64_t n, const int32_t* shipdate, const int32_t* quantity,
              const int32_t* discount, const int32_t* tax, const int64_t* extprice,
              const uint8_t* returnflag, const uint8_t* linestatus, int32_t cutoff,
              int nthreads, int64_t* out, uint64_t* charge_lo, int64_t* charge_hi) {
  /* The sums of price and price*(100-disc) are int64 (the charge is int128):
   * refuse (-1) a table where rows x max|price| x (100 + max|disc|) reaches
   * 2^63 instead of returning a wrapped sum. */
  unsigned __int128 mp = 0, md = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const unsigned __int128 p = extprice[i] < 0 ? (unsigned __int128)(-(__int128)extprice[i])
                                                : (unsigned __int128)extprice[i];
    const unsigned __int128 d = discount[i] < 0 ? (unsigned __int128)(-(__int128)discount[i])
                                                : (unsigned __int128)discount[i];
    if (p > mp) mp = p;
    if (d > md) md = d;
  }
  if ((unsigned __int128)n * mp * (100 + md) >= ((unsigned __int128)1 << 63)) return -1;
  q1_arg* a = (q1_arg*)calloc(1, sizeof(q1_arg));
  a->ship = shipdate; a->qty = quantity; a->disc = discount; a->tax = tax;
  a->price = extprice; a->rf = returnflag; a->ls = linestatus; a->cutoff = cutoff;
  int T = par_for(n, nthreads, q1_range, a);
  memset(out, 0, 30 * sizeof(int64_t));
  __int128 ch[6] = {0};
  for (int t = 0; t < T; ++t) {
    for (int k = 0; k < 30; ++k) out[k] += a->acc[t][k];
    for (int g = 0; g < 6; ++g) ch[g] += a->charge[t][g];
  }
  for (int g = 0; g < 6; ++g) {
    charge_lo[g] = (uint64_t)ch[g];
    charge_hi[g] = (int64_t)(ch[g] >> 64);
  }
  free(a);
  return 0;
}

/* -------------------------------------------------------------- key index */
/* Unique int32 keys (primary keys: custkey, orderkey, partkey, suppkey) of the
 * rows that pass a filter -> an int32 value per key (a row number, a nation,
 * a year). Any key layout: a direct array over [min, max] of the kept keys
 * when that span is at most 4 x rows + 2^20 values, else the (key, value)
 * pairs sorted by key and searched by bisection; both give the same answers,
 * the direct array keeps benchmark-size runs fast. -1 = no such key. */
typedef struct { int32_t key, val; } kv_t;
typedef struct {
  int32_t lo;
  uint64_t span;   /* direct: values over [lo, lo + span) */
  int32_t* direct;
  kv_t* pairs;     /* sorted, when not direct */
  uint64_t n;
} keymap;

static int kv_cmp(const void* x, const void* y) {
  const kv_t *a = (const kv_t*)x, *b = (const kv_t*)y;
  return (a->key > b->key) - (a->key < b->key);
}

/* keep: NULL = every row; vals: NULL = 1 for every kept key. */
static void keymap_build(keymap* m, uint64_t n, const int32_t* keys, const int32_t* vals,
                         const uint8_t* keep) {
  memset(m, 0, sizeof(*m));
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  uint64_t kept = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (keep && !keep[i]) continue;
    if (keys[i] < lo) lo = keys[i];
    if (keys[i] > hi) hi = keys[i];
    ++kept;
  }
  m->n = kept;
  if (!kept) return;
  const uint64_t span = (uint64_t)(hi - lo) + 1;
  if (span <= 4 * kept + (1u << 20)) {
    m->lo = (int32_t)lo;
    m->span = span;
    m->direct = (int32_t*)malloc(span * sizeof(int32_t));
    memset(m->direct, 0xff, span * sizeof(int32_t));
    for (uint64_t i = 0; i < n; ++i)
      if (!keep || keep[i]) m->direct[(int64_t)keys[i] - lo] = vals ? vals[i] : 1;
    return;
  }
  m->pairs = (kv_t*)malloc(kept * sizeof(kv_t));
  uint64_t j = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (!keep || keep[i]) m->pairs[j++] = (kv_t){keys[i], vals ? vals[i] : 1};
  qsort(m->pairs, kept, sizeof(kv_t), kv_cmp);
}

static int32_t keymap_get(const keymap* m, int32_t key) {
  if (m->direct) {
    const int64_t d = (int64_t)key - m->lo;
    return d >= 0 && (uint64_t)d < m->span ? m->direct[d] : -1;
  }
  uint64_t lo = 0, hi = m->n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (m->pairs[mid].key < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < m->n && m->pairs[lo].key == key ? m->pairs[lo].val : -1;
}

static void keymap_free(keymap* m) {
  free(m->direct);
  free(m->pairs);
}

/* 0..n-1 as int32 row numbers. */
static int32_t* row_numbers(uint64_t n) {
  int32_t* r = (int32_t*)calloc(n ? n : 1, sizeof(int32_t));
  for (uint64_t i = 0; i < n; ++i) r[i] = (int32_t)i;
  return r;
}

/* --------------------------------------------------------------------- Q3 */
/* sql/q3.sql:4-26; filters of kernels/q3.metal:52 (segment), :70 (orderdate <
 * D), :167 (shipdate > D); order of src/infra.h:547-551 plus orderkey asc as
 * the final tie-break. Orderkeys are unique in orders (primary key). */
typedef struct {
  const int32_t *lok, *lship, *ldisc;
  const int64_t* lprice;
  const keymap* omap;  /* orderkey of a qualifying order -> its orders row */
  int32_t date;
  int64_t* rev;        /* per orders row */
  uint8_t* hit;
} q3_arg;

static void q3_range(void* p, uint64_t lo, uint64_t hi, int tid) {
  (void)tid;
  q3_arg* a = (q3_arg*)p;
  for (uint64_t i = lo; i < hi; ++i) {
    if (a->lship[i] <= a->date) continue;
    const int32_t r = keymap_get(a->omap, a->lok[i]);
    if (r < 0) continue;
    __atomic_fetch_add(&a->rev[r], a->lprice[i] * (int64_t)(100 - a->ldisc[i]), __ATOMIC_RELAXED);
    a->hit[r] = 1;
  }
}

typedef struct { int32_t key, date, prio; int64_t rev; } q3_row;

static int q3_cmp(const void* x, const void* y) {
  const q3_row *a = (const q3_row*)x, *b = (const q3_row*)y;
  if (a->rev != b->rev) return a->rev > b->rev ? -1 : 1;
  if (a->date != b->date) return a->date < b->date ? -1 : 1;
  return (a->key > b->key) - (a->key < b->key);
}

int oracle_q3(uint64_t nc, const int32_t* c_custkey, const uint8_t* c_segment, uint64_t no,
              const int32_t* o_orderkey, const int32_t* o_custkey, const int32_t* o_orderdate,
              const int32_t* o_shippriority, uint64_t nl, const int32_t* l_orderkey,
              const int32_t* l_shipdate, const int32_t* l_discount, const int64_t* l_extprice,
              int32_t segment, int32_t date, int k, int nthreads, int32_t* out_orderkey,
              int32_t* out_orderdate, int32_t* out_shippriority, int64_t* out_revenue,
              int64_t* groups) {
  uint8_t* keep = (uint8_t*)malloc((nc > no ? nc : no) + 1);
  for (uint64_t i = 0; i < nc; ++i) keep[i] = c_segment[i] == (uint8_t)segment;
  keymap cseg, omap;
  keymap_build(&cseg, nc, c_custkey, NULL, keep);
  for (uint64_t i = 0; i < no; ++i)
    keep[i] = o_orderdate[i] < date && keymap_get(&cseg, o_custkey[i]) >= 0;
  int32_t* orow = row_numbers(no);
  keymap_build(&omap, no, o_orderkey, orow, keep);
  q3_arg a = {l_orderkey, l_shipdate, l_discount, l_extprice, &omap, date,
              (int64_t*)calloc(no + 1, sizeof(int64_t)), (uint8_t*)calloc(no + 1, 1)};
  par_for(nl, nthreads, q3_range, &a);
  int64_t ng = 0;
  for (uint64_t r = 0; r < no; ++r) ng += a.hit[r];
  q3_row* rows = (q3_row*)malloc((size_t)(ng ? ng : 1) * sizeof(q3_row));
  int64_t m = 0;
  for (uint64_t r = 0; r < no; ++r)
    if (a.hit[r]) rows[m++] = (q3_row){o_orderkey[r], o_orderdate[r], o_shippriority[r], a.rev[r]};
  qsort(rows, (size_t)m, sizeof(q3_row), q3_cmp);
  int w = (int)(m < k ? m : k);
  for (int i = 0; i < w; ++i) {
    out_orderkey[i] = rows[i].key;
    out_orderdate[i] = rows[i].date;
    out_shippriority[i] = rows[i].prio;
    out_revenue[i] = rows[i].rev;
  }
  *groups = ng;
  free(rows); free(a.rev); free(a.hit); free(orow); free(keep);
  keymap_free(&omap);
  keymap_free(&cseg);
  return w;
}

/* --------------------------------------------------------------------- Q9 */
/* sql/q9.sql; join chain of kernels/q9.metal:228-284 (part filter ->
 * supplier nation -> partsupp (partkey, suppkey) -> orders year); amount =
 * extprice*(1-discount) - supplycost*quantity in exact fixed point; order by
 * nation name asc, year desc (the SQL's ORDER BY nation, o_year DESC).
 * (partkey, suppkey) is unique in partsupp and orderkey in orders. */
#define Q9_YEARS 64
typedef struct {
  const int32_t *lpk, *lsk, *lok, *lqty, *ldisc;
  const int64_t* lprice;
  const keymap* green;   /* partkey of a part whose name has the pattern -> 1 */
  const keymap* snat;    /* suppkey -> nation */
  const keymap* oyear;   /* orderkey -> year */
  uint64_t* ht_key; int64_t* ht_cost; uint64_t ht_mask;
  int64_t (*acc)[25 * Q9_YEARS];
  uint8_t (*seen)[25 * Q9_YEARS];
} q9_arg;

static uint64_t q9_hash(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
  return k ^ (k >> 33);
}

static void q9_range(void* p, uint64_t lo, uint64_t hi, int tid) {
  q9_arg* a = (q9_arg*)p;
  int64_t* acc = a->acc[tid];
  uint8_t* seen = a->seen[tid];
  for (uint64_t i = lo; i < hi; ++i) {
    const int32_t pk = a->lpk[i], sk = a->lsk[i];
    if (keymap_get(a->green, pk) < 0) continue;
    const int32_t nat = keymap_get(a->snat, sk);
    const int32_t year = keymap_get(a->oyear, a->lok[i]);
    if (nat < 0 || year < 0) continue;
    uint64_t key = ((uint64_t)(uint32_t)pk << 32) | (uint32_t)sk;
    uint64_t h = q9_hash(key) & a->ht_mask;
    int64_t cost = -1;
    while (a->ht_key[h] != UINT64_MAX) {
      if (a->ht_key[h] == key) { cost = a->ht_cost[h]; break; }
      h = (h + 1) & a->ht_mask;
    }
    if (cost < 0) continue;
    int64_t amount = a->lprice[i] * (int64_t)(100 - a->ldisc[i]) - cost * (int64_t)a->lqty[i] * 100;
    int slot = nat * Q9_YEARS + (year - 1970);
    acc[slot] += amount;
    seen[slot] = 1;
  }
}

static const char* const NATION[25] = {
    "ALGERIA", "ARGENTINA", "BRAZIL", "CANADA", "EGYPT", "ETHIOPIA", "FRANCE",
    "GERMANY", "INDIA", "INDONESIA", "IRAN", "IRAQ", "JAPAN", "JORDAN", "KENYA",
    "MOROCCO", "MOZAMBIQUE", "PERU", "CHINA", "ROMANIA", "SAUDI ARABIA", "VIETNAM",
    "RUSSIA", "UNITED KINGDOM", "UNITED STATES"};

typedef struct { int32_t nation, year; int64_t profit; } q9_row;
static int q9_cmp(const void* x, const void* y) {
  const q9_row *a = (const q9_row*)x, *b = (const q9_row*)y;
  int c = strcmp(NATION[a->nation], NATION[b->nation]);
  if (c) return c;
  return (b->year > a->year) - (b->year < a->year);
}

int oracle_q9(uint64_t np, const int32_t* p_partkey, const uint8_t* p_name, int name_width,
              uint64_t ns, const int32_t* s_suppkey, const int32_t* s_nationkey, uint64_t nps,
              const int32_t* ps_partkey, const int32_t* ps_suppkey, const int64_t* ps_supplycost,
              uint64_t no, const int32_t* o_orderkey, const int32_t* o_orderdate, uint64_t nl,
              const int32_t* l_partkey, const int32_t* l_suppkey, const int32_t* l_orderkey,
              const int32_t* l_quantity, const int32_t* l_discount, const int64_t* l_extprice,
              const char* pattern, int nthreads, int32_t* out_nation, int32_t* out_year,
              int64_t* out_profit, int max_rows) {
  uint8_t* named = (uint8_t*)malloc(np + 1);
  for (uint64_t i = 0; i < np; ++i)
    named[i] = (uint8_t)oracle_contains(p_name + i * (uint64_t)name_width, name_width, pattern);
  int32_t* years = (int32_t*)malloc((no + 1) * sizeof(int32_t));
  for (uint64_t i = 0; i < no; ++i) years[i] = oracle_year_of_day(o_orderdate[i]);
  keymap green, snat, oyear;
  keymap_build(&green, np, p_partkey, NULL, named);
  keymap_build(&snat, ns, s_suppkey, s_nationkey, NULL);
  keymap_build(&oyear, no, o_orderkey, years, NULL);
  uint64_t cap = 16;
  while (cap < 2 * nps + 16) cap <<= 1;
  uint64_t* hk = (uint64_t*)malloc(cap * sizeof(uint64_t));
  int64_t* hc = (int64_t*)malloc(cap * sizeof(int64_t));
  memset(hk, 0xff, cap * sizeof(uint64_t));
  for (uint64_t i = 0; i < nps; ++i) {
    int32_t pk = ps_partkey[i];
    if (keymap_get(&green, pk) < 0) continue;
    uint64_t key = ((uint64_t)(uint32_t)pk << 32) | (uint32_t)ps_suppkey[i];
    uint64_t h = q9_hash(key) & (cap - 1);
    while (hk[h] != UINT64_MAX && hk[h] != key) h = (h + 1) & (cap - 1);
    hk[h] = key;
    hc[h] = ps_supplycost[i];
  }
  int T = nthreads < 1 ? 1 : (nthreads > 256 ? 256 : nthreads);
  q9_arg a = {l_partkey, l_suppkey, l_orderkey, l_quantity, l_discount, l_extprice, &green, &snat,
              &oyear, hk, hc, cap - 1,
              calloc((size_t)T, sizeof(int64_t[25 * Q9_YEARS])),
              calloc((size_t)T, sizeof(uint8_t[25 * Q9_YEARS]))};
  par_for(nl, T, q9_range, &a);
  q9_row rows[25 * Q9_YEARS];
  int m = 0;
  for (int s = 0; s < 25 * Q9_YEARS; ++s) {
    int64_t tot = 0;
    int any = 0;
    for (int t = 0; t < T; ++t) { tot += a.acc[t][s]; any |= a.seen[t][s]; }
    if (any) rows[m++] = (q9_row){s / Q9_YEARS, 1970 + s % Q9_YEARS, tot};
  }
  qsort(rows, (size_t)m, sizeof(q9_row), q9_cmp);
  int w = m < max_rows ? m : max_rows;
  for (int i = 0; i < w; ++i) {
    out_nation[i] = rows[i].nation;
    out_year[i] = rows[i].year;
    out_profit[i] = rows[i].profit;
  }
  free(a.acc); free(a.seen); free(hk); free(hc); free(years); free(named);
  keymap_free(&green);
  keymap_free(&snat);
  keymap_free(&oyear);
  return w;
}

/* -------------------------------------------------------------------- Q13 */
/* sql/q13.sql; pattern test of kernels/q13.metal:43-70; counting of
 * :121-126 without its dense-custkey assumption; zero-order customers in
 * c_count 0 (the LEFT OUTER JOIN, :153-158); no bin cap. */
typedef struct {
  const int32_t* ock;
  const uint8_t* com;
  int width;
  const char *w1, *w2;
  const keymap* crow;  /* custkey -> customer row */
  uint32_t* cnt;       /* per customer row */
} q13_arg;

static void q13_range(void* p, uint64_t lo, uint64_t hi, int tid) {
  (void)tid;
  q13_arg* a = (q13_arg*)p;
  for (uint64_t i = lo; i < hi; ++i) {
    const int32_t r = keymap_get(a->crow, a->ock[i]);
    if (r < 0) continue;
    if (oracle_like_two(a->com + i * (uint64_t)a->width, a->width, a->w1, a->w2)) continue;
    __atomic_fetch_add(&a->cnt[r], 1u, __ATOMIC_RELAXED);
  }
}

typedef struct { int32_t c; int64_t d; } q13_row;
static int q13_cmp(const void* x, const void* y) {
  const q13_row *a = (const q13_row*)x, *b = (const q13_row*)y;
  if (a->d != b->d) return a->d > b->d ? -1 : 1;
  return (b->c > a->c) - (b->c < a->c);
}

int oracle_q13(uint64_t nc, const int32_t* c_custkey, uint64_t no, const int32_t* o_custkey,
               const uint8_t* o_comment, int comment_width, const char* word1, const char* word2,
               int nthreads, int32_t* out_count, int64_t* out_custdist, int max_rows) {
  int32_t* rows_of = row_numbers(nc);
  keymap crow;
  keymap_build(&crow, nc, c_custkey, rows_of, NULL);
  q13_arg a = {o_custkey, o_comment, comment_width, word1, word2, &crow,
               (uint32_t*)calloc(nc ? nc : 1, sizeof(uint32_t))};
  par_for(no, nthreads, q13_range, &a);
  uint32_t maxc = 0;
  for (uint64_t i = 0; i < nc; ++i) if (a.cnt[i] > maxc) maxc = a.cnt[i];
  int64_t* hist = (int64_t*)calloc((size_t)maxc + 1, sizeof(int64_t));
  for (uint64_t i = 0; i < nc; ++i) hist[a.cnt[i]]++;
  q13_row* rows = (q13_row*)malloc(((size_t)maxc + 1) * sizeof(q13_row));
  int m = 0;
  for (uint32_t c = 0; c <= maxc; ++c) if (hist[c]) rows[m++] = (q13_row){(int32_t)c, hist[c]};
  qsort(rows, (size_t)m, sizeof(q13_row), q13_cmp);
  int w = m < max_rows ? m : max_rows;
  for (int i = 0; i < w; ++i) { out_count[i] = rows[i].c; out_custdist[i] = rows[i].d; }
  free(rows); free(hist); free(a.cnt); free(rows_of);
  keymap_free(&crow);
  return w;
}

/* ------------------------------------------------------------- primitives */
/* selection_kernel, kernels/microbenchmarks.metal:160-171 (`col < value`). */
uint64_t oracle_select_lt(uint64_t n, const int32_t* col, int32_t value, uint64_t* bitmap) {
  uint64_t c = 0;
  if (bitmap) memset(bitmap, 0, ((n + 63) / 64) * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) {
    if (col[i] < value) {
      ++c;
      if (bitmap) bitmap[i >> 6] |= 1ULL << (i & 63);
    }
  }
  return c;
}

/* sum_kernel_stage1/2, kernels/microbenchmarks.metal:177-218, exact. */
int64_t oracle_sum_i32(uint64_t n, const int32_t* col) {
  int64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) s += col[i];
  return s;
}
int64_t oracle_sum_i64(uint64_t n, const int64_t* col) {
  int64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) s += col[i];
  return s;
}

/* hash_join_build/probe, kernels/microbenchmarks.metal:224-293: number of
 * (probe, build) pairs with equal keys, by sort + binary search. */
static int cmp_i32(const void* x, const void* y) {
  int32_t a = *(const int32_t*)x, b = *(const int32_t*)y;
  return (a > b) - (a < b);
}
uint64_t oracle_join_count(uint64_t nb, const int32_t* build, uint64_t np, const int32_t* probe) {
  int32_t* s = (int32_t*)malloc((nb ? nb : 1) * sizeof(int32_t));
  memcpy(s, build, nb * sizeof(int32_t));
  qsort(s, nb, sizeof(int32_t), cmp_i32);
  uint64_t c = 0;
  for (uint64_t i = 0; i < np; ++i) {
    uint64_t lo = 0, hi = nb;
    while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (s[mid] < probe[i]) lo = mid + 1; else hi = mid; }
    uint64_t a = lo;
    hi = nb;
    while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (s[mid] <= probe[i]) lo = mid + 1; else hi = mid; }
    c += lo - a;
  }
  free(s);
  return c;
}
