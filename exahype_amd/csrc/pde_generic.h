// User PDEs of the generic (flux / source / non-conservative-product) kernel
// arms, resolved statically like the reference's SolverType template
// argument.  Each PDE declares which arms it uses -- the useFlux, useSource
// and useNCP template arguments of spaceTimePredictorNonlinear.cpph:42,
// riemannSolverNonlinear.cpph:37 and volumeIntegralNonlinear.cpph:26 -- and
// provides the pointwise functions of the solver plugin surface
// (E/solvers/UserSolverInterface.h:42-176): flux, eigenvalues,
// nonConservativeProduct, algebraicSource.  FP64 only.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#include "pde_euler.h"

namespace exa {

// Run-time parameters of the compiled PDEs.
struct PdeParams {
  double gamma;   // ratio of specific heats (Euler)
  double adv[3];  // advection velocity (linear advection)
};

// A point source of the linear scheme (the usePointSource arm of
// spaceTimePredictorLinear.cpph:78-102, 287-302) as the linear predictor
// receives it.  mode 0: none; 1: the forces PSi [cell][B+1][z][y][x][v] that
// pointSource computed, read from `psi`; 2: evaluated per node on a regular
// grid, cell (ix, iy, iz) centred at off + (i + 1/2) h with x fastest, at the
// time stored after the step size (dtp[2]).  The user pointSource is a Ricker
// wavelet force amp[v] (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2, at x0.
struct PointSrc {
  int mode;
  const double* psi;
  double off[3];
  int nx, ny;
  double x0[3], amp[5], f0, t0;
};

// 3D Euler (flux only): the existing Euler functions behind this interface.
struct PdeEuler {
  static constexpr bool FLUX = true, SRC = false, NCP = false;
  __device__ static inline void flux(const double* Q, double F[3][5], const PdeParams& p) {
    EulerPDE::flux<double>(Q, F, p.gamma);
  }
  __device__ static inline double max_abs_eigenvalue(const double* Q, int d, const PdeParams& p) {
    return EulerPDE::max_abs_eigenvalue(Q, d, p.gamma);
  }
  __device__ static inline void max_abs_eigenvalues3(const double* Q, const PdeParams& p, double m[3]) {
    EulerPDE::max_abs_eigenvalues3(Q, p.gamma, m);
  }
  __device__ static inline void source(const double*, double S[5], const PdeParams&) {
    for (int v = 0; v < 5; ++v) S[v] = 0.0;
  }
  __device__ static inline void ncp(const double*, const double[3][5], double out[5], const PdeParams&) {
    for (int v = 0; v < 5; ++v) out[v] = 0.0;
  }
};

// |lambda| of a PDE whose eigenvalues are a_d for every variable, accumulated
// as stableTimeStepSize.cpph:43-48 does (max(|lambda|, m) from 0).
__device__ inline double advection_speed(double a) {
  const double s = fabs(a);
  return 0.0 < s ? s : 0.0;
}

// Linear advection of 5 variables in flux form: F_d = a_d Q, eigenvalues a_d.
struct PdeAdvection {
  static constexpr bool FLUX = true, SRC = false, NCP = false;
  __device__ static inline void flux(const double* Q, double F[3][5], const PdeParams& p) {
    for (int v = 0; v < 5; ++v) {
      F[0][v] = p.adv[0] * Q[v];
      F[1][v] = p.adv[1] * Q[v];
      F[2][v] = p.adv[2] * Q[v];
    }
  }
  __device__ static inline double max_abs_eigenvalue(const double*, int d, const PdeParams& p) {
    return advection_speed(p.adv[d]);
  }
  __device__ static inline void max_abs_eigenvalues3(const double*, const PdeParams& p, double m[3]) {
    for (int d = 0; d < 3; ++d) m[d] = advection_speed(p.adv[d]);
  }
  __device__ static inline void source(const double*, double S[5], const PdeParams&) {
    for (int v = 0; v < 5; ++v) S[v] = 0.0;
  }
  __device__ static inline void ncp(const double*, const double[3][5], double out[5], const PdeParams&) {
    for (int v = 0; v < 5; ++v) out[v] = 0.0;
  }
};

// The same advection in non-conservative form: no flux, B_d = a_d I,
// nonConservativeProduct = a_x dQ/dx + a_y dQ/dy + a_z dQ/dz.
struct PdeAdvectionNcp {
  static constexpr bool FLUX = false, SRC = false, NCP = true;
  __device__ static inline void flux(const double*, double F[3][5], const PdeParams&) {
    for (int d = 0; d < 3; ++d)
      for (int v = 0; v < 5; ++v) F[d][v] = 0.0;
  }
  __device__ static inline double max_abs_eigenvalue(const double*, int d, const PdeParams& p) {
    return advection_speed(p.adv[d]);
  }
  __device__ static inline void max_abs_eigenvalues3(const double*, const PdeParams& p, double m[3]) {
    for (int d = 0; d < 3; ++d) m[d] = advection_speed(p.adv[d]);
  }
  __device__ static inline void source(const double*, double S[5], const PdeParams&) {
    for (int v = 0; v < 5; ++v) S[v] = 0.0;
  }
  __device__ static inline void ncp(const double*, const double G[3][5], double out[5], const PdeParams& p) {
    for (int v = 0; v < 5; ++v) out[v] = p.adv[0] * G[0][v] + p.adv[1] * G[1][v] + p.adv[2] * G[2][v];
  }
  // the linear scheme's product: a dim x nVar tensor B_d dQ/dx_d
  // (spaceTimePredictorLinear.cpph:28-30), and coefficientMatrix B_d = a_d I
  __device__ static inline void ncp_linear(const double*, const double G[3][5], double T[3][5], const PdeParams& p) {
    for (int d = 0; d < 3; ++d)
      for (int v = 0; v < 5; ++v) T[d][v] = p.adv[d] * G[d][v];
  }
  __device__ static inline void coefficient_matrix(const double*, int d, double Bn[5][5], const PdeParams& p) {
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) Bn[i][j] = i == j ? p.adv[d] : 0.0;
  }
};

// 3D Euler with the test non-conservative product of the reference's Euler
// kernel test (E/tests/kernels/c/GenericEulerKernelTest3D.cpp:84-106,
// _setNcpAndMatrixBToZero == false): BgradQ = (Q0, Q3, 3, gradQ[0][0], 0.7).
struct PdeEulerNcpTest {
  static constexpr bool FLUX = true, SRC = false, NCP = true;
  __device__ static inline void flux(const double* Q, double F[3][5], const PdeParams& p) {
    EulerPDE::flux<double>(Q, F, p.gamma);
  }
  __device__ static inline double max_abs_eigenvalue(const double* Q, int d, const PdeParams& p) {
    return EulerPDE::max_abs_eigenvalue(Q, d, p.gamma);
  }
  __device__ static inline void max_abs_eigenvalues3(const double* Q, const PdeParams& p, double m[3]) {
    EulerPDE::max_abs_eigenvalues3(Q, p.gamma, m);
  }
  __device__ static inline void source(const double*, double S[5], const PdeParams&) {
    for (int v = 0; v < 5; ++v) S[v] = 0.0;
  }
  __device__ static inline void ncp(const double* Q, const double G[3][5], double out[5], const PdeParams&) {
    out[0] = Q[0];
    out[1] = Q[3];
    out[2] = 3.0;
    out[3] = G[0][0];
    out[4] = 0.7;
  }
  // in the linear scheme the test product fills the first row of the tensor
  // (it writes nVar values); the coefficient matrix of the test is zero
  // (GenericEulerKernelTest3D.cpp:108-131)
  __device__ static inline void ncp_linear(const double* Q, const double G[3][5], double T[3][5],
                                           const PdeParams& p) {
    ncp(Q, G, T[0], p);
    for (int v = 0; v < 5; ++v) T[1][v] = T[2][v] = 0.0;
  }
  __device__ static inline void coefficient_matrix(const double*, int, double Bn[5][5], const PdeParams&) {
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) Bn[i][j] = 0.0;
  }
};

// kernels::fusedSource (kernels/fusedMethods.cpph:17-32): S = algebraicSource
// (or 0) minus the non-conservative product.
template <class PDE>
__device__ inline void fused_source(const double* Q, const double G[3][5], double S[5], const PdeParams& p) {
  PDE::source(Q, S, p);
  if (PDE::NCP) {
    double n[5];
    PDE::ncp(Q, G, n, p);
    for (int v = 0; v < 5; ++v) S[v] -= n[v];
  }
}

}  // namespace exa
