// pgn_internal.hpp -- shared host-side types of libpgn (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pgn.h"

namespace pgn {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PGN_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::pgn::Error(PGN_ERR_HIP, std::string(#call) + ": " +             \
                                          hipGetErrorString(e_) + " at " +    \
                                          __FILE__ + ":" +                    \
                                          std::to_string(__LINE__));          \
  } while (0)

#define PGN_CHECK(cond, msg)                                   \
  do {                                                         \
    if (!(cond)) throw ::pgn::Error(PGN_ERR_INVALID, (msg)); \
  } while (0)

// Device allocation owned by a model or context.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    reset();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(PGN_ERR_NOMEM, "hipMalloc of " +
                                     std::to_string(count * sizeof(T)) +
                                     " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
  }
  // keeps the allocation (and its contents) when the size already matches
  void ensure(size_t count) {
    if (count != n || (count && !p)) alloc(count);
  }
  size_t bytes() const { return n * sizeof(T); }
  void upload(const T* h, size_t count, hipStream_t s) {
    if (count != n) alloc(count);
    if (count)
      PGN_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count)
      PGN_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// total size of a set of device buffers
template <class... B>
size_t sum_bytes(const B&... b) {
  return (size_t(0) + ... + b.bytes());
}

}  // namespace pgn

struct pgn_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};
