// Block-level top-k over (revenue desc, date asc, key asc) entries in LDS.
//
// A block streams candidate entries into an LDS buffer; whenever the buffer
// could overflow on the next round it is bitonic-sorted best-first and cut
// back to the k best. The final sort leaves the block's k best at the front.
// Used by Q3 (reference: CPU partial_sort in src/infra.h:544-551, moved to
// the GPU) in two stages: every block of a grid over the aggregation table,
// then one block over the per-block winners.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gdbk {

struct TopEntry {
  int64_t rev;
  int32_t date;
  int32_t key;
  int32_t prio;
  int32_t pad;
};

__device__ __forceinline__ TopEntry worst_entry() {
  return TopEntry{INT64_MIN, INT32_MAX, INT32_MAX, 0, 0};
}

// True when a ranks before b.
__device__ __forceinline__ bool better(const TopEntry& a, const TopEntry& b) {
  if (a.rev != b.rev) return a.rev > b.rev;
  if (a.date != b.date) return a.date < b.date;
  return a.key < b.key;
}

// Sorts buf[0, n) best-first; buf must have room for the next power of two
// (padding slots are overwritten with worst entries). Called by all threads.
template <int BUF>
__device__ void bitonic_sort_best_first(TopEntry* buf, int n) {
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = n + threadIdx.x; i < P; i += blockDim.x) buf[i] = worst_entry();
  __syncthreads();
  // Each step compares P/2 pairs (i, i + stride); pair p puts a zero bit at
  // `stride` into p to get i, so every thread has work.
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int p = threadIdx.x; p < P / 2; p += blockDim.x) {
        const int i = ((p & ~(stride - 1)) << 1) | (p & (stride - 1));
        const int j = i + stride;
        const bool up = (i & size) == 0;  // this half sorts best-first
        const TopEntry a = buf[i], b = buf[j];
        if (up ? better(b, a) : better(a, b)) {
          buf[i] = b;
          buf[j] = a;
        }
      }
      __syncthreads();
    }
  }
}

// Running top-k state in LDS. `push` must be called by every thread of the
// block (a multiple of 64 threads) once per round with at most one candidate
// each (valid = false for none); rounds of blockDim candidates never overflow
// while k + 2*blockDim <= BUF. A wave reserves its slots with one LDS atomic.
//
// With `thr` (an LDS slot) set, every sort that leaves k entries records the k-th
// best in *thr, and later candidates that do not rank before it are dropped at
// once: they can no longer reach the top k, and the buffer stops refilling.
template <int BUF>
struct BlockTopK {
  TopEntry* buf;
  int* n;
  int k;
  TopEntry* thr = nullptr;   // running k-th best (valid while *have)
  int* have = nullptr;

  __device__ void init() {
    if (threadIdx.x == 0) {
      *n = 0;
      if (have) *have = 0;
    }
    __syncthreads();
  }
  __device__ void push(bool valid, const TopEntry& e) {
    if (valid && have && *have && !better(e, *thr)) valid = false;
    const unsigned long long m = __ballot(valid);
    if (m) {
      const int lane = threadIdx.x & 63, leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(n, __popcll(m));
      base = __shfl(base, leader, 64);
      if (valid) buf[base + __popcll(m & ((1ull << lane) - 1))] = e;
    }
    __syncthreads();
    const int cur = *n;
    __syncthreads();  // every thread has read n before any thread pushes again
    if (cur > BUF - (int)blockDim.x) {
      bitonic_sort_best_first<BUF>(buf, cur);
      if (threadIdx.x == 0) {
        *n = min(*n, k);
        if (have && k > 0 && cur >= k) {
          *thr = buf[k - 1];
          *have = 1;
        }
      }
      __syncthreads();
    }
  }
  // Sorts and writes the k best (worst-padded) to out[0, k).
  __device__ void finish(TopEntry* __restrict__ out) {
    const int m = *n;
    bitonic_sort_best_first<BUF>(buf, m);
    for (int i = threadIdx.x; i < k; i += blockDim.x) out[i] = i < m ? buf[i] : worst_entry();
  }
};

}  // namespace gdbk
