// Shared host-side helpers of the C-ABI: error reporting and workspace carving.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/deltamic.h"

namespace dmic {

void set_error(const char* fmt, ...);

#define DMIC_HIP_CHECK(expr)                                                  \
  do {                                                                        \
    hipError_t err__ = (expr);                                                \
    if (err__ != hipSuccess) {                                                \
      ::dmic::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(err__), \
                        __FILE__, __LINE__);                                  \
      return DMIC_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

#define DMIC_LAUNCH_CHECK() DMIC_HIP_CHECK(hipGetLastError())

#define DMIC_REQUIRE(cond, code, ...) \
  do {                                \
    if (!(cond)) {                    \
      ::dmic::set_error(__VA_ARGS__); \
      return (code);                  \
    }                                 \
  } while (0)

// Carves 256-byte aligned regions out of a caller-provided workspace.  The
// same sequence of take() calls sizes the workspace when base == nullptr.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
  size_t bytes() const { return (off + 255) & ~size_t(255); }
};

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace dmic
