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

// Device buffer that grows on demand and frees itself.  Move-only: a copy
// would free the same allocation twice.
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  void swap(DevBuf& o) noexcept {
    void* tp = p;
    size_t tn = n;
    p = o.p;
    n = o.n;
    o.p = tp;
    o.n = tn;
  }
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

// A stable time step that is 0 (or not a finite positive number) means a cell
// held NaN entries or a negative pressure: the kernels' dt reduction maps
// every such candidate to 0 so that it wins the global minimum.
static inline int exa_check_dt(double dt, const char* what) {
  if (dt > 0.0 && dt <= 1.7976931348623157e308) return EXA_OK;
  exa_set_error("%s: the stable time step is %g -- the state holds non-finite values or a negative pressure", what,
                dt);
  return EXA_ERR_NUMERICS;
}

// Takes the device address of the allocation-free AMR staging area on the
// current device (amr_capi.hip); called by exa_init once per device.
int exa_amr_noalloc_setup(int dev);
