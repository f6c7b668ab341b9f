// Host-side helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

#define EXA_ARG(cond, msg)                        \
  do {                                            \
    if (!(cond)) {                                \
      exa_set_error("invalid argument: %s", msg); \
      return EXA_ERR_ARG;                         \
    }                                             \
  } while (0)

#define EXA_TRY(expr)              \
  do {                             \
    int r__ = (expr);              \
    if (r__ != EXA_OK) return r__; \
  } while (0)

// Device buffer that grows on demand and frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return EXA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    EXA_CHECK(hipMalloc(&p, bytes));
    n = bytes;
    return EXA_OK;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

static inline unsigned long long exa_big_bits() {
  const double big = 1.0e300;
  unsigned long long b;
  __builtin_memcpy(&b, &big, sizeof(b));
  return b;
}
