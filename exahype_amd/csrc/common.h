// Shared device-side definitions: error handling, DG tables in constant
// memory, small wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/exahype_amd.h"

#define EXA_MAX_ORDER 9
#define EXA_NB 10

// Status codes returned by every C-ABI entry point (see include/exahype_amd.h).
#define EXA_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e__ = (expr);                                                         \
    if (e__ != hipSuccess) {                                                         \
      exa_set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e__)); \
      return EXA_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

void exa_set_error(const char* fmt, ...);

// DG operator tables of one order, padded to 10.  Filled once per device
// from the literal reference tables (dg_tables_data.h).
struct DGTables {
  double w[EXA_NB];          // Gauss-Legendre weights on [0,1]
  double xi[EXA_NB];         // Gauss-Legendre nodes on [0,1]
  double F0[EXA_NB];         // basis at t=0 (time flux)
  double FL[EXA_NB];         // basis at 0 (left extrapolation)
  double FR[EXA_NB];         // basis at 1 (right extrapolation)
  double Kxi[EXA_NB][EXA_NB];  // stiffness matrix Kxi[k][m]
  double iK1[EXA_NB][EXA_NB];  // inverse time matrix iK1[t'][t]
  double dudx[EXA_NB][EXA_NB]; // derivative of the basis dudx[l][n] = phi_n'(xi_l) (NCP gradients)
  double eq[EXA_NB][EXA_NB];   // equidistantGridProjector1d[node][j] = phi_node(j/N) (plotting)
};

// one copy per translation unit; every unit that launches kernels reading the
// tables uploads them at exa_init
static __constant__ DGTables c_tab[EXA_MAX_ORDER + 1];

// The same tables rounded to FP32, for the low-precision predictor sweeps.
struct DGTablesF {
  float w[EXA_NB];
  float F0[EXA_NB];
  float Kxi[EXA_NB][EXA_NB];
  float iK1[EXA_NB][EXA_NB];
};

static __constant__ DGTablesF c_tabf[EXA_MAX_ORDER + 1];

// PNPM factors of the CFL condition, kernels/aderdg/generic/c/3d/stableTimeStepSize.cpph:34-35
__host__ __device__ inline double exa_pnpm(int N) {
  const double t[10] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};
  return t[N];
}

// A time-step candidate that is not a finite positive number -- from a state
// with NaN entries or negative pressure (sqrt of a negative sound speed
// squared) -- becomes 0.0.  Zero then wins every minimum (comparisons with
// NaN are false, so a NaN would be dropped or kept depending on the order of
// the reduction), and the host reads a next time step of 0 as a blown-up
// state instead of stepping on with a finite-looking dt.
__device__ inline double exa_dt_sane(double v) { return (v > 0.0 && v <= 1.7976931348623157e308) ? v : 0.0; }

// positive doubles order like their bit patterns: min via integer atomics
// The minimum only decreases: a value not below the current one skips the
// atomic, so the hundreds of thousands of per-cell candidates of a step cost
// a plain load each and only the few improving ones contend for the address.
__device__ inline void atomic_min_pos_double(unsigned long long* addr, double v) {
  v = exa_dt_sane(v);
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  if (b < __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(addr, b);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ inline T wave_min(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// Block-wide sum of one double per thread; every thread gets the result.
// scratch: >= (blockDim.x/64) doubles of LDS.
__device__ inline double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += scratch[i];
  return s;
}

__device__ inline double block_min(double v, double* scratch) {
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double s = scratch[0];
  for (int i = 1; i < nw; ++i) s = scratch[i] < s ? scratch[i] : s;
  return s;
}

constexpr int round_up64(int x) { return (x + 63) / 64 * 64; }
