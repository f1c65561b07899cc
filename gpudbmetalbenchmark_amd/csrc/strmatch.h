// Substring search over fixed-width byte fields staged in LDS.
//
// A field is a NUL-padded byte string (its length is the offset of the first
// NUL, or the full width). `find_from` returns the first offset >= from where
// the pattern occurs entirely inside the string, or -1. Candidates are found
// 4 bytes at a time: exact per-byte equality flags for the pattern's first two
// bytes (SWAR, no false positives), combined across dword boundaries, then
// each candidate is verified byte by byte from LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace gdbk {

struct Pattern {
  uint8_t b[32];
  int len;  // 1..31
};

// Bit 7 of byte i set iff byte i of x == 0 (exact).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
  const uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
  return ~t & 0x80808080u;
}
__device__ __forceinline__ uint32_t eq_bytes(uint32_t x, uint32_t bcast) { return zero_bytes(x ^ bcast); }

// Length of the NUL-padded string in s[0, width) (width a multiple of 4,
// s 4-byte aligned in LDS).
__device__ __forceinline__ int field_len(const uint8_t* s, int width) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(s);
  for (int i = 0; i < width / 4; ++i) {
    const uint32_t z = zero_bytes(w[i]);
    if (z) return 4 * i + (__builtin_ctz(z) >> 3);
  }
  return width;
}

// First occurrence of p in s[from, len) (s 4-byte aligned, readable up to
// round_up(len, 4) + 4 bytes), or -1.
__device__ __forceinline__ int find_from(const uint8_t* s, int len, int from, const Pattern& p) {
  if (p.len <= 0 || from + p.len > len) return p.len <= 0 ? from : -1;
  const uint32_t* w = reinterpret_cast<const uint32_t*>(s);
  const uint32_t c0 = 0x01010101u * p.b[0];
  const uint32_t c1 = 0x01010101u * (p.len > 1 ? p.b[1] : p.b[0]);
  const int last = len - p.len;  // last admissible start
  const int nw = (last >> 2) + 1;
  uint32_t cur = w[from >> 2];
  for (int i = from >> 2; i < nw; ++i) {
    const uint32_t nxt = w[i + 1];
    uint32_t m = eq_bytes(cur, c0);
    if (p.len > 1) {
      // second byte of a match starting at byte j of this dword: byte j+1 of (cur:nxt)
      const uint32_t shifted = __builtin_amdgcn_alignbyte(nxt, cur, 1);
      m &= eq_bytes(shifted, c1);
    }
    while (m) {
      const int j = 4 * i + (__builtin_ctz(m) >> 3);
      m &= m - 1;
      if (j < from) continue;
      if (j > last) return -1;
      bool ok = true;
      for (int q = 2; q < p.len; ++q) {
        if (s[j + q] != p.b[q]) {
          ok = false;
          break;
        }
      }
      if (ok) return j;
    }
    cur = nxt;
  }
  return -1;
}

}  // namespace gdbk

namespace gdbk {

// ---------------------------------------------------------------------------
// Register-resident variant: a lane holds its whole field in NW dwords plus two
// zero pad dwords, so the scan uses only compile-time register indices; the
// rare candidates are verified with whole-dword compares.
struct PatternW {
  uint32_t w[8];      // pattern bytes packed little-endian, zero padded
  uint32_t last_mask; // valid-byte mask of the last dword
  int nw;             // dwords holding the pattern
  int len;            // 1..31
  uint32_t c0, c1;    // first / second byte broadcast to 4 lanes
  uint32_t m0;        // valid-byte mask of the first dword (~0 when len >= 4)
};

// Host: packs a 1..31-byte word for the register-resident matchers.
inline bool make_pattern(const char* s, PatternW* p) {
  const size_t n = strnlen(s, 32);
  if (n == 0 || n > 31) return false;
  memset(p, 0, sizeof(*p));
  uint8_t b[32] = {0};
  memcpy(b, s, n);
  memcpy(p->w, b, 32);
  p->len = (int)n;
  p->nw = (int)(n + 3) / 4;
  const int rem = (int)n - 4 * (p->nw - 1);
  p->last_mask = rem == 4 ? 0xFFFFFFFFu : ((1u << (8 * rem)) - 1u);
  p->c0 = 0x01010101u * b[0];
  p->c1 = 0x01010101u * (n > 1 ? b[1] : b[0]);
  p->m0 = n >= 4 ? 0xFFFFFFFFu : p->last_mask;
  return true;
}

// Value of d[idx] for a run-time idx (select chain over compile-time indices).
template <int N>
__device__ __forceinline__ uint32_t pick(const uint32_t* d, int idx) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) r = (idx == k) ? d[k] : r;
  return r;
}

// True if the pattern occurs at byte j (j + len <= bytes held in d).
template <int N>
__device__ __forceinline__ bool verify_at(const uint32_t* d, int j, const PatternW& p) {
  const int i0 = j >> 2, o = j & 3;
  uint32_t lo = pick<N>(d, i0);
  for (int c = 0; c < p.nw; ++c) {
    const uint32_t hi = pick<N>(d, i0 + c + 1);
    const uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, o);
    const uint32_t m = (c == p.nw - 1) ? p.last_mask : 0xFFFFFFFFu;
    if ((v & m) != p.w[c]) return false;
    lo = hi;
  }
  return true;
}

// Length of the NUL-padded string held in d[0, NW).
template <int NW>
__device__ __forceinline__ int reg_len(const uint32_t* d) {
  int len = NW * 4;
#pragma unroll
  for (int i = NW - 1; i >= 0; --i) {
    const uint32_t z = zero_bytes(d[i]);
    if (z) len = 4 * i + (__builtin_ctz(z) >> 3);
  }
  return len;
}

// First occurrence of p at an offset >= from entirely inside [0, len), or -1.
template <int NW>
__device__ __forceinline__ int reg_find(const uint32_t* d, int len, int from, const PatternW& p) {
  const int last = len - p.len;
  if (from > last) return -1;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (4 * i > last) break;
    if (4 * i + 3 < from) continue;
    uint32_t m = eq_bytes(d[i], p.c0);
    if (p.len > 1) m &= eq_bytes(__builtin_amdgcn_alignbyte(d[i + 1], d[i], 1), p.c1);
    while (m) {
      const int j = 4 * i + (__builtin_ctz(m) >> 3);
      m &= m - 1;
      if (j < from) continue;
      if (j > last) return -1;
      if (verify_at<NW + 2>(d, j, p)) return j;
    }
  }
  return -1;
}

// True if the pattern's dwords 1 .. nw-1 follow at byte 4i+O (dword 0 was
// already compared); I and O are compile-time, so every index is too.
template <int I, int O>
__device__ __forceinline__ bool rest_at(const uint32_t* d, const PatternW& p) {
#pragma unroll
  for (int c = 1; c < 8; ++c) {
    if (c >= p.nw) break;
    const uint32_t y = O == 0 ? d[I + c] : __builtin_amdgcn_alignbyte(d[I + c + 1], d[I + c], O);
    const uint32_t m = (c == p.nw - 1) ? p.last_mask : 0xFFFFFFFFu;
    if ((y & m) != p.w[c]) return false;
  }
  return true;
}

template <int I, int NW, bool FULL0>
__device__ __forceinline__ void like_two_dword(const uint32_t* d, const PatternW& a,
                                               const PatternW& b, int& pa, int& pb) {
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const int pos = 4 * I + o;
    const uint32_t x = o == 0 ? d[I] : __builtin_amdgcn_alignbyte(d[I + 1], d[I], o);
    // a: its first occurrence
    if ((FULL0 ? x : (x & a.m0)) == a.w[0]) {
      bool hit;
      switch (o) {
        case 0: hit = rest_at<I, 0>(d, a); break;
        case 1: hit = rest_at<I, 1>(d, a); break;
        case 2: hit = rest_at<I, 2>(d, a); break;
        default: hit = rest_at<I, 3>(d, a); break;
      }
      if (hit && pa < 0) pa = pos;
    }
    // b: its first occurrence starting at or after the end of the first a
    if ((FULL0 ? x : (x & b.m0)) == b.w[0]) {
      if (pb < 0 && pa >= 0 && pos >= pa + a.len) {
        bool hit;
        switch (o) {
          case 0: hit = rest_at<I, 0>(d, b); break;
          case 1: hit = rest_at<I, 1>(d, b); break;
          case 2: hit = rest_at<I, 2>(d, b); break;
          default: hit = rest_at<I, 3>(d, b); break;
        }
        if (hit) pb = pos;
      }
    }
  }
}

template <int I, int NW, bool FULL0>
struct LikeTwoLoop {
  __device__ __forceinline__ static void run(const uint32_t* d, const PatternW& a,
                                             const PatternW& b, int& pa, int& pb) {
    like_two_dword<I, NW, FULL0>(d, a, b, pa, pb);
    LikeTwoLoop<I + 1, NW, FULL0>::run(d, a, b, pa, pb);
  }
};
template <int NW, bool FULL0>
struct LikeTwoLoop<NW, NW, FULL0> {
  __device__ __forceinline__ static void run(const uint32_t*, const PatternW&, const PatternW&,
                                             int&, int&) {}
};

// SQL `s LIKE '%a%b%'` for the NUL-padded string held in d[0, NW) (d holds
// at least 8 zero dwords after it): the leftmost a inside the string, then a b
// starting at or after its end, also inside the string. Every byte offset is
// tested by whole-dword compares of the pattern's first four bytes (one
// alignbyte shared by both patterns); the rest of a pattern is compared only
// where those match. Patterns hold no NUL byte, so an occurrence lies inside
// the string iff it starts before the first NUL: occurrences are found over
// the whole field and only the chosen b is checked against the length. FULL0:
// both patterns are at least 4 bytes long (no mask on the first dword).
template <int NW, bool FULL0>
__device__ __forceinline__ bool like_two(const uint32_t* d, const PatternW& a, const PatternW& b) {
  int pa = -1, pb = -1;
  LikeTwoLoop<0, NW, FULL0>::run(d, a, b, pa, pb);
  if (pb < 0) return false;
  return pb < reg_len<NW>(d);
}

// ---------------------------------------------------------------------------
// Branch-free `s LIKE '%a%b%'` for patterns of compile-time lengths LA, LB in
// 4..8 (every TPC-H Q13 substitution word has 7 or 8 bytes). A pattern of 4..8
// bytes occurs at p iff its first four bytes match at p and its last four at
// p + L - 4, so each byte offset costs one alignbyte (shared) and two compares
// per pattern, with every index compile-time. The compares become lane masks:
// an occurrence of a at p is folded into `seen` at offset p + LA, so a b found
// at q with `seen` set starts at or after the end of some a. A b at q inside
// dword i lies before the first NUL iff dwords 0..i hold no NUL byte (the
// bytes of b from q to the end of dword i are not NUL, so a NUL in dword i
// would precede q); an a before such a b is then inside the string too.
__device__ __forceinline__ bool has_zero_byte(uint32_t v) {
  return ((v - 0x01010101u) & ~v & 0x80808080u) != 0u;
}

template <int NW>
__device__ __forceinline__ uint32_t bytes_at(const uint32_t* d, int p) {
  return (p & 3) == 0 ? d[p >> 2] : __builtin_amdgcn_alignbyte(d[(p >> 2) + 1], d[p >> 2], p & 3);
}

template <int NW, int LA, int LB>
__device__ __forceinline__ bool like_two_fixed(const uint32_t* d, uint32_t a_head, uint32_t a_tail,
                                               uint32_t b_head, uint32_t b_tail) {
  static_assert(LA >= 4 && LA <= 8 && LB >= 4 && LB <= 8, "pattern lengths 4..8");
  constexpr int NB = NW * 4;
  bool hit_a[NB];
  bool seen = false, matched = false, alive = true;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const bool alive_next = alive && !has_zero_byte(d[i]);
    bool found = false;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      const int p = 4 * i + o;
      const uint32_t x = bytes_at<NW>(d, p);
      hit_a[p] = false;
      if (p + LA <= NB) hit_a[p] = (x == a_head) & (bytes_at<NW>(d, p + LA - 4) == a_tail);
      if (p >= LA) seen = seen | hit_a[p - LA];
      if (p + LB <= NB) found = found | (seen & (x == b_head) & (bytes_at<NW>(d, p + LB - 4) == b_tail));
    }
    matched = matched | (found & alive_next);
    alive = alive_next;
  }
  return matched;
}

}  // namespace gdbk
