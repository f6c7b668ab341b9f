// Small bookkeeping kernels of the device grid and the ensemble (dt
// minimum, simulated time, Picard sweep totals).  Included by one
// translation unit only.
#pragma once
#include "aderdg_kernels.h"

namespace exa {

// dt bookkeeping between steps: dt_cur <- min over cells, time += dt used.
__global__ void k_dt_advance(double* dt_cur, unsigned long long* dt_next, double* time) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *time += *dt_cur;
    *dt_cur = __longlong_as_double((long long)*dt_next);
    *dt_next = (unsigned long long)__double_as_longlong(1.0e300);
  }
}

__global__ void k_dt_reset(unsigned long long* dt_next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dt_next = (unsigned long long)__double_as_longlong(1.0e300);
}

__global__ void k_dt_take(double* dt_cur, const unsigned long long* dt_next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dt_cur = __longlong_as_double((long long)*dt_next);
}

// Ensemble dt bookkeeping, one thread per instance.
__global__ void k_ens_dt_advance(InstParams* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i].time += p[i].dt_cur;
  p[i].dt_cur = __longlong_as_double((long long)p[i].dt_next);
  p[i].dt_next = (unsigned long long)__double_as_longlong(1.0e300);
}

__global__ void k_ens_dt_reset(InstParams* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i].dt_next = (unsigned long long)__double_as_longlong(1.0e300);
}

__global__ void k_ens_dt_take(InstParams* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    p[i].dt_cur = __longlong_as_double((long long)p[i].dt_next);
    p[i].dt_next = (unsigned long long)__double_as_longlong(1.0e300);
  }
}

__global__ void k_sum_iters(const int* __restrict__ iters, int n, unsigned long long* out) {
  unsigned long long acc = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += iters[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

}  // namespace exa
