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

// Fused time stepping (the reference's fused algorithmic steps, global time
// stepping).  The reference predicts with an optimistic time step before the
// step's minimum is known and, after the traversal
// (TimeStepSizeComputation.cpp:144-187, then startNewTimeStep,
// ADERDGSolver.cpp:556-570): used = the predictor's dt (the stable one if
// 0); used > stable: the corrector and predictor take w * stable and the
// predictor is recomputed (Runner.cpp:1114-1120); otherwise the corrector
// keeps used and the next optimistic dt is 0.5 (used + w * stable).  Here the
// minimum is known before the next predictor launch, so the predictor runs
// once with the final dt: each step is predictor + corrector with dt_cur, and
// this kernel picks the next dt_cur and optimistic dt_pred from the new
// state's stable dt, counting the steps whose optimistic dt was unstable
// (where the reference reran the predictor).
__global__ void k_dt_advance_fused(double* dt_cur, unsigned long long* dt_next, double* time, double* dt_pred,
                                   unsigned long long* reruns, double w) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double stable = __longlong_as_double((long long)*dt_next);
    double used = *dt_pred;
    if (used == 0.0) used = stable;
    *time += *dt_cur;
    if (used > stable) {
      *dt_cur = w * stable;
      *dt_pred = w * stable;
      *reruns += 1;
    } else {
      *dt_cur = used;
      *dt_pred = 0.5 * (used + w * stable);
    }
    *dt_next = (unsigned long long)__double_as_longlong(1.0e300);
  }
}

// Start of fused time stepping from a stable dt in dt_cur: the time step data
// the reference has after its first time-step-size computation with a zero
// predictor dt (used = stable, 0.5 (stable + w stable)) and the
// initialisation sweep (FusedTimeSteppingInitialisation.cpp:80-100: the
// corrector takes the predictor's dt).
__global__ void k_dt_fused_start(double* dt_cur, double* dt_pred, double w) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double s = *dt_cur;
    const double d = 0.5 * (s + w * s);
    *dt_cur = d;
    *dt_pred = d;
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
