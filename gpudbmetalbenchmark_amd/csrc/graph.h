// Launch sequences captured once per launch shape and replayed as HIP graphs
// (no counterpart in the reference, which encodes every command buffer anew).
//
// A template's kernels read everything that varies between replays from
// device descriptors the caller uploads before the launch, so a captured
// sequence is valid for every batch of the same shape. Captured sequences
// hold kernels only: uploads and downloads stay outside the graph, and no
// memset node is ever captured (clears are fill kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <map>

#include "internal.h"

namespace gdb {

struct GraphCache {
  std::map<uint64_t, hipGraphExec_t> execs;
  uint64_t captures = 0, replays = 0;
};

// Replays the graph of `key`, capturing `enqueue(stream)` first if needed.
template <class F>
int graph_launch(GraphCache* gc, uint64_t key, hipStream_t st, F&& enqueue) {
  auto it = gc->execs.find(key);
  if (it == gc->execs.end()) {
    GDB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != GDB_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc != GDB_OK ? rc : fail(GDB_E_HIP, "graph capture failed");
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    GDB_HIP(ie);
    it = gc->execs.emplace(key, ex).first;
    ++gc->captures;
  }
  GDB_HIP(hipGraphLaunch(it->second, st));
  ++gc->replays;
  return GDB_OK;
}

inline void graph_cache_clear(GraphCache* gc) {
  for (auto& kv : gc->execs) (void)hipGraphExecDestroy(kv.second);
  gc->execs.clear();
}

}  // namespace gdb
