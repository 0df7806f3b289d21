// graph.hpp -- replayable PCG chunks (HIP graphs).
//
// A chunk is a fixed sequence of launches: c PCG iterations and the status
// launch.  Its launches read their iteration relative to the chunk start from
// device memory (Args::kctl), so one captured graph replays any chunk of the
// same length, parity of its first iteration and launch arguments.  Replay
// runs the same kernels with the same arguments as direct launches: results
// are bit-identical.  Captured graphs hold no memset or memcpy nodes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "pgn_internal.hpp"

namespace pgn {

template <class Args>
struct ChunkGraphs {
  struct Entry {
    int c = 0, parity = 0;
    Args key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Entry> v;
  // chunks seen once: a graph is captured only for a chunk seen before, so
  // a launch sequence that never repeats (a batch whose active set keeps
  // changing) does not pay for captures it never replays
  struct Seen {
    int c = 0, parity = 0;
    Args key;
  };
  std::vector<Seen> seen;
  static constexpr size_t MAX_ENTRIES = 8;

  ChunkGraphs() = default;
  ChunkGraphs(const ChunkGraphs&) = delete;
  ChunkGraphs& operator=(const ChunkGraphs&) = delete;
  ~ChunkGraphs() { clear(); }
  void clear() {
    for (auto& e : v)
      if (e.exec) (void)hipGraphExecDestroy(e.exec);
    v.clear();
  }
  void reset() {
    clear();
    seen.clear();
  }

  // The executable graph of this chunk, captured from `launches(stream)` the
  // second time the chunk comes; null the first time (launch directly).
  // `a` must be zero-initialised before its fields were set, so that equal
  // arguments compare equal byte for byte.
  template <class F>
  hipGraphExec_t get(hipStream_t s, int c, int parity, const Args& a, F&& launches) {
    for (auto& e : v)
      if (e.c == c && e.parity == parity && std::memcmp(&e.key, &a, sizeof(Args)) == 0)
        return e.exec;
    bool again = false;
    for (auto& e : seen)
      again |= e.c == c && e.parity == parity && std::memcmp(&e.key, &a, sizeof(Args)) == 0;
    if (!again) {
      if (seen.size() >= 4 * MAX_ENTRIES) seen.clear();
      Seen n;
      n.c = c;
      n.parity = parity;
      n.key = a;
      seen.push_back(n);
      return nullptr;
    }
    if (v.size() >= MAX_ENTRIES) clear();
    PGN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      launches(s);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    PGN_HIP(hipGetLastError());
    PGN_HIP(ec);
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PGN_HIP(ei);
    Entry e;
    e.c = c;
    e.parity = parity;
    e.key = a;
    e.exec = ex;
    v.push_back(e);
    return ex;
  }
};

// Single models with more nodes than this queue their PCG chunks as direct
// launches by default: the GPU, not the host, then bounds a chunk, and the
// replayed graph was the slower of the two (per solve, replay against direct:
// 25.1 / 25.7 ms with 4,961 nodes, 48.5 / 48.8 ms with 35,721, 78.1 / 76.8
// ms with 116,281, 115.5 / 114.3 ms with 270,641).
constexpr long GRAPH_AUTO_MAX_NODES = 65536;

// PGN_GRAPH in the environment: 0 launches every PCG chunk directly, 1
// replays graphs for every single model, 2 for fused batches as well (whose
// segmented launch arguments change as instances converge, so that captures
// seldom repeat: 51 against 44 ms on 64 small beams); unset (-1), single
// models of at most GRAPH_AUTO_MAX_NODES nodes replay
inline int graphs_level() {
  const char* e = std::getenv("PGN_GRAPH");
  return e ? std::atoi(e) : -1;
}

}  // namespace pgn
