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
