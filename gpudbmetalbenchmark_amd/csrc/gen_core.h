// Seeded, counter-based synthetic table generator: the value of every cell is a
// pure function of (seed, stream, key), so tables can be generated in any order,
// in parallel, or chunk by chunk (the streamed 600M-row fact table) and always
// come out identical.
//
// The reference loads TPC-H dbgen .tbl files (src/infra.h:262-361); dbgen needs
// the network here, so the tables are synthesised with the distributions the
// benchmark configurations fix:
//   flat fact  : shipdate U[0,2525], quantity U[1,50], discount U[0,10] (hundredths),
//                extprice = quantity * retail, retail U[90000,209900] (cents),
//                tax U[0,8], receipt = ship + U[1,30],
//                linestatus 'F' if shipdate < 1263 else 'O',
//                returnflag 'R'/'A' 50/50 if receipt < 1263 else 'N'.
//   customer   : custkey = i+1 (dense), segment U{0..4}.
//   orders     : orderkey = TPC-H sparse key (8 of every 32), custkey U[1,C],
//                orderdate U[0,2405], shippriority 0, 80-byte comment.
//   joined fact: 1..7 rows per order, shipdate = orderdate + U[1,121],
//                partkey U[1,P], suppkey = one of the part's 4 suppliers.
//   part       : partkey = i+1, 55-byte name of 5 distinct words of 92.
//   partsupp   : 4 rows per part, supplycost U[100,100000] (cents).
//   supplier   : suppkey = i+1, nationkey U[0,24].
// Day 0 is 1992-01-01.
#pragma once
#include <stdint.h>
#include <string.h>

namespace gdbgen {

enum Stream : uint32_t {
  LI_SHIP = 1, LI_QTY = 2, LI_DISC = 3, LI_RETAIL = 4, LI_TAX = 5, LI_RECEIPT = 6, LI_RFLAG = 7,
  C_SEG = 10,
  O_CUST = 20, O_DATE = 21, O_NLINES = 22, O_CLEN = 23, O_CWORD = 24,
  J_SHIP = 30, J_QTY = 31, J_DISC = 32, J_RETAIL = 33, J_PART = 34, J_SUPPIDX = 35,
  P_NAME = 40,
  PS_COST = 50,
  S_NATION = 60,
};

static const uint64_t GAMMA = 0x9E3779B97F4A7C15ULL;

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
static inline uint64_t mix64(uint64_t z) {
  z += GAMMA;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t stream_base(uint64_t seed, uint32_t stream) {
  return mix64(seed ^ ((uint64_t)stream << 32));
}
// k-th output of the SplitMix64 sequence that starts at `base`.
static inline uint64_t draw(uint64_t base, uint64_t key) { return mix64(base + key * GAMMA); }
// Uniform integer in [lo, hi] by Lemire's multiply-shift on the top 32 bits.
static inline int64_t uni(uint64_t r, int64_t lo, int64_t hi) {
  uint64_t span = (uint64_t)(hi - lo + 1);
  return lo + (int64_t)(((r >> 32) * span) >> 32);
}

// Calendar and domain constants.
static const int32_t SHIP_MAX = 2525;        // last ship day of the flat fact table
static const int32_t STATUS_DAY = 1263;      // 1995-06-17
static const int32_t ORDERDATE_MAX = 2405;   // 1998-08-02
static const int32_t RETAIL_LO = 90000, RETAIL_HI = 209900;
static const int COMMENT_W = 80;
static const int PNAME_W = 55;
static const int NSEGMENTS = 5;
static const int NNATIONS = 25;

static inline int32_t orderkey_of(uint64_t i) {
  return (int32_t)(((i >> 3) << 5) + (i & 7) + 1);
}
static inline int32_t lines_of_order(uint64_t seed, uint64_t order_index) {
  return (int32_t)uni(draw(stream_base(seed, O_NLINES), order_index), 1, 7);
}
static inline int32_t orderdate_of(uint64_t seed, uint64_t order_index) {
  return (int32_t)uni(draw(stream_base(seed, O_DATE), order_index), 0, ORDERDATE_MAX);
}
// j-th (0..3) supplier of part p (1-based) among S suppliers; distinct for S >= 4.
static inline int32_t supplier_of(int64_t p, int j, int64_t S) {
  return (int32_t)(((p - 1) + (int64_t)j * (S / 4)) % S + 1);
}

// The 92 colour words of the TPC-H P_NAME grammar (TPC-H spec 4.2.3).
static const char* const PNAME_WORDS[92] = {
  "almond", "antique", "aquamarine", "azure", "beige", "bisque", "black", "blanched",
  "blue", "blush", "brown", "burlywood", "burnished", "chartreuse", "chiffon", "chocolate",
  "coral", "cornflower", "cornsilk", "cream", "cyan", "dark", "deep", "dim",
  "dodger", "drab", "firebrick", "floral", "forest", "frosted", "gainsboro", "ghost",
  "goldenrod", "green", "grey", "honeydew", "hot", "indian", "ivory", "khaki",
  "lace", "lavender", "lawn", "lemon", "light", "lime", "linen", "magenta",
  "maroon", "medium", "metallic", "midnight", "mint", "misty", "moccasin", "navajo",
  "navy", "olive", "orange", "orchid", "pale", "papaya", "peach", "peru",
  "pink", "plum", "powder", "puff", "purple", "red", "rose", "rosy",
  "royal", "saddle", "salmon", "sandy", "seashell", "sienna", "sky", "slate",
  "smoke", "snow", "spring", "steel", "tan", "thistle", "tomato", "turquoise",
  "violet", "wheat", "white", "yellow"};

// Vocabulary of the order comments: 48 words, so "special" and "requests" each
// occur with probability 1/48 per word and about 1 comment in 100 matches
// '%special%requests%'.
static const char* const COMMENT_WORDS[48] = {
  "furiously", "carefully", "quickly", "slyly", "blithely", "regular", "final", "ironic",
  "pending", "express", "bold", "even", "unusual", "silent", "daring", "fluffy",
  "deposits", "accounts", "packages", "requests", "special", "theodolites", "instructions", "foxes",
  "pinto", "beans", "asymptotes", "dependencies", "excuses", "platelets", "ideas", "courts",
  "dolphins", "sleep", "wake", "haggle", "nag", "use", "boost", "affix",
  "detect", "integrate", "cajole", "among", "across", "above", "after", "according"};
static const int COMMENT_NWORDS = 48;
static const int COMMENT_MAXWORDS = 16;

// Writes the 80-byte, NUL-padded comment of order i: up to 16 words joined by
// single spaces and cut at a length drawn from U[20,80] (a cut may split a word,
// and a comment of length 80 has no terminating NUL).
static inline void order_comment(uint64_t seed, uint64_t i, uint8_t* out) {
  char buf[16 * 16];
  int len = 0;
  const uint64_t wb = stream_base(seed, O_CWORD);
  for (int w = 0; w < COMMENT_MAXWORDS; ++w) {
    const char* word = COMMENT_WORDS[uni(draw(wb, i * COMMENT_MAXWORDS + w), 0, COMMENT_NWORDS - 1)];
    if (w) buf[len++] = ' ';
    int wl = (int)strlen(word);
    memcpy(buf + len, word, wl);
    len += wl;
  }
  int cut = (int)uni(draw(stream_base(seed, O_CLEN), i), 20, COMMENT_W);
  if (cut > len) cut = len;
  memset(out, 0, COMMENT_W);
  memcpy(out, buf, cut);
}

// Writes the 55-byte, NUL-padded name of part index i (partkey i+1): 5 distinct
// words drawn in order, a repeat being redrawn from the next key.
static inline void part_name(uint64_t seed, uint64_t i, uint8_t* out) {
  const uint64_t nb = stream_base(seed, P_NAME);
  int picked[5];
  uint64_t k = 0;
  for (int w = 0; w < 5; ++w) {
    for (;;) {
      int cand = (int)uni(draw(nb, i * 64 + k), 0, 91);
      ++k;
      bool dup = false;
      for (int q = 0; q < w; ++q) dup |= (picked[q] == cand);
      if (!dup) { picked[w] = cand; break; }
    }
  }
  memset(out, 0, PNAME_W);
  int len = 0;
  for (int w = 0; w < 5; ++w) {
    if (w) out[len++] = ' ';
    int wl = (int)strlen(PNAME_WORDS[picked[w]]);
    memcpy(out + len, PNAME_WORDS[picked[w]], wl);
    len += wl;
  }
}

}  // namespace gdbgen
