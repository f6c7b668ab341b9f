// 3D compressible Euler as a user PDE compiled into the kernels: the
// pointwise flux() / eigenvalues() of the reference's solver plugin surface
// (E/tests/kernels/c/GenericEulerKernelTest3D.cpp:26-82), resolved
// statically like the reference's SolverType template argument.  Templated
// on the arithmetic type so the mixed-precision predictor can evaluate
// fluxes in FP32.  The grouping of every expression follows the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#include "../../include/exahype_amd.h"

struct EulerPDE {
  static constexpr int NVar = 5;

  // F[d][v] = flux in direction d
  template <typename T>
  __host__ __device__ static inline void flux(const T* Q, T F[3][5], T gamma) {
    const T irho = T(1.0) / Q[0];
    const T p = (gamma - T(1)) * (Q[4] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    F[0][0] = Q[1];
    F[0][1] = irho * Q[1] * Q[1] + p;
    F[0][2] = irho * Q[1] * Q[2];
    F[0][3] = irho * Q[1] * Q[3];
    F[0][4] = irho * Q[1] * (Q[4] + p);
    F[1][0] = Q[2];
    F[1][1] = irho * Q[2] * Q[1];
    F[1][2] = irho * Q[2] * Q[2] + p;
    F[1][3] = irho * Q[2] * Q[3];
    F[1][4] = irho * Q[2] * (Q[4] + p);
    F[2][0] = Q[3];
    F[2][1] = irho * Q[3] * Q[1];
    F[2][2] = irho * Q[3] * Q[2];
    F[2][3] = irho * Q[3] * Q[3] + p;
    F[2][4] = irho * Q[3] * (Q[4] + p);
  }

  template <typename T>
  __host__ __device__ static inline void eigenvalues(const T* Q, int d, T* lambda, T gamma) {
    const T irho = T(1.0) / Q[0];
    const T p = (gamma - T(1)) * (Q[4] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    const T u_n = Q[d + 1] * irho;
    const T c = sqrt(gamma * p * irho);
    lambda[0] = u_n - c;
    lambda[1] = u_n;
    lambda[2] = u_n;
    lambda[3] = u_n;
    lambda[4] = u_n + c;
  }

  // flux in direction d only (same expressions as flux(); p on the diagonal)
  __host__ __device__ static inline void flux_dir(const double* Q, int d, double* f, double gamma) {
    const double irho = 1.0 / Q[0];
    const double p = (gamma - 1) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    const double qd = Q[1 + d];
    f[0] = qd;
    f[1] = irho * qd * Q[1];
    f[2] = irho * qd * Q[2];
    f[3] = irho * qd * Q[3];
    f[1 + d] = f[1 + d] + p;
    f[4] = irho * qd * (Q[4] + p);
  }

  // max_v |lambda_v| accumulated as in stableTimeStepSize.cpph:43-48
  // (std::max(fabs(lambda), maxEigenvalue) starting from 0).
  __host__ __device__ static inline double max_abs_eigenvalue(const double* Q, int d, double gamma) {
    double lam[5];
    eigenvalues<double>(Q, d, lam, gamma);
    double m = 0.0;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double a = fabs(lam[v]);
      m = (a < m) ? m : a;  // std::max(a, m): a NaN wave speed propagates
    }
    return m;
  }

  // max_abs_eigenvalue for d = 0, 1, 2 at once: 1/rho, p and c are the same
  // values eigenvalues() computes for every direction, evaluated once.
  __host__ __device__ static inline void max_abs_eigenvalues3(const double* Q, double gamma, double m[3]) {
    const double irho = 1.0 / Q[0];
    const double p = (gamma - 1.0) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    const double c = sqrt(gamma * p * irho);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double u_n = Q[d + 1] * irho;
      const double lam[5] = {u_n - c, u_n, u_n, u_n, u_n + c};
      double mm = 0.0;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double a = fabs(lam[v]);
        mm = (a < mm) ? mm : a;
      }
      m[d] = mm;
    }
  }
};

// Pointwise initial conditions (the adjustPointSolution of the configured
// problems).  Mirrors the restatement used to check it.
__host__ __device__ inline void exa_ic_point(const exa_ic& ic, double gamma, const double x[3], double t, double* Q) {
  const double pi = 3.14159265358979323846;
  double rho, u0, u1, u2, p;
  if (ic.kind == 2) {
    for (int v = 0; v < 5; ++v) Q[v] = ic.q0[v];
    return;
  }
  if (ic.kind == 3) {  // sine modes of the linear-advection PDEs, advected with uinf
    double s = 0.0;
    for (int m = 0; m < ic.nmodes; ++m)
      s += ic.a[m] * sin(2.0 * pi * (ic.k[m][0] * (x[0] - ic.uinf[0] * t) + ic.k[m][1] * (x[1] - ic.uinf[1] * t) +
                                     ic.k[m][2] * (x[2] - ic.uinf[2] * t)) + ic.phi[m]);
    for (int v = 0; v < 5; ++v) Q[v] = 1.0 + (v + 1) * s;
    return;
  }
  if (ic.kind == 5) {
    const double d0 = x[0] - ic.xc[0], d1 = x[1] - ic.xc[1], d2 = x[2] - ic.xc[2];
    rho = ic.rho_inf + ic.amp * exp(-(d0 * d0 + d1 * d1 + d2 * d2) / ic.L);
    p = ic.p_inf;
    u0 = ic.uinf[0];
    u1 = ic.uinf[1];
    u2 = ic.uinf[2];
  } else if (ic.kind == 4) {
    double s = 0.0;
    for (int m = 0; m < ic.nmodes; ++m)
      s += ic.a[m] * sin(2.0 * pi * (ic.k[m][0] * (x[0] - ic.uinf[0] * t) + ic.k[m][1] * (x[1] - ic.uinf[1] * t) +
                                     ic.k[m][2] * (x[2] - ic.uinf[2] * t)) + ic.phi[m]);
    rho = ic.rho_inf + ic.amp * s;
    p = ic.p_inf;
    u0 = ic.uinf[0];
    u1 = ic.uinf[1];
    u2 = ic.uinf[2];
  } else if (ic.kind == 1) {
    const double d0 = x[0] - ic.xc[0], d1 = x[1] - ic.xc[1], d2 = x[2] - ic.xc[2];
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    rho = r < ic.r0 ? ic.rho_in : ic.rho_out;
    p = r < ic.r0 ? ic.p_in : ic.p_out;
    u0 = u1 = u2 = 0.0;
  } else {
    double xs = x[0] - ic.uinf[0] * t, ys = x[1] - ic.uinf[1] * t;
    xs = xs - floor(xs);
    ys = ys - floor(ys);
    const double rx = (xs - ic.xc[0]) / ic.L, ry = (ys - ic.xc[1]) / ic.L;
    const double r2 = rx * rx + ry * ry;
    const double e1 = exp(0.5 * (1.0 - r2));
    const double du = ic.eps / (2.0 * pi) * e1;
    const double dT = -(gamma - 1.0) * ic.eps * ic.eps / (8.0 * gamma * pi * pi) * exp(1.0 - r2);
    const double T = 1.0 + dT;
    rho = ic.rho_inf * pow(T, 1.0 / (gamma - 1.0));
    p = ic.p_inf * pow(T, gamma / (gamma - 1.0));
    u0 = ic.uinf[0] - du * ry;
    u1 = ic.uinf[1] + du * rx;
    u2 = ic.uinf[2];
    double s = 0.0;
    for (int m = 0; m < ic.nmodes; ++m)
      s += ic.a[m] * sin(2.0 * pi * (ic.k[m][0] * x[0] + ic.k[m][1] * x[1] + ic.k[m][2] * x[2]) + ic.phi[m]);
    rho += ic.amp * s;
  }
  Q[0] = rho;
  Q[1] = rho * u0;
  Q[2] = rho * u1;
  Q[3] = rho * u2;
  Q[4] = p / (gamma - 1.0) + 0.5 * rho * (u0 * u0 + u1 * u1 + u2 * u2);
}

// The EulerFV demonstrator's user PDE (2D, variables rho, j[3], E;
// Demonstrators/EulerFV/MyEulerSolver.cpp:51-85): p = (g-1)(E - ((0.5/rho) j).j),
// F[d][1+i] = (1/rho)(j_i j_d) + p delta_id as the generated Fluxes::j(Matrix)
// writes it (Toolkit/src/eu/exahype/variables/Variables.java:629-636).
struct EulerFV2DPDE {
  static constexpr int NVar = 5;
  __host__ __device__ static inline double pressure(const double* Q, double irho, double gamma) {
    const double h = 0.5 * irho;
    const double jj = (h * Q[1]) * Q[1] + (h * Q[2]) * Q[2] + (h * Q[3]) * Q[3];
    return (gamma - 1) * (Q[4] - jj);
  }
  __host__ __device__ static inline void flux_dir(const double* Q, int d, double* f, double gamma) {
    const double irho = 1. / Q[0];
    const double p = pressure(Q, irho, gamma);
    f[0] = Q[1 + d];
    for (int i = 0; i < 3; ++i) f[1 + i] = irho * (Q[1 + i] * Q[1 + d]) + p * (i == d ? 1.0 : 0.0);
    f[4] = (irho * (Q[4] + p)) * Q[1 + d];
  }
  __host__ __device__ static inline double max_abs_eigenvalue(const double* Q, int d, double gamma) {
    const double irho = 1. / Q[0];
    const double p = pressure(Q, irho, gamma);
    const double u_n = Q[d + 1] * irho;
    const double c = sqrt(gamma * p * irho);
    const double lam[5] = {u_n - c, u_n, u_n, u_n, u_n + c};
    double m = fabs(lam[0]);
    for (int v = 1; v < 5; ++v) m = (m < fabs(lam[v])) ? fabs(lam[v]) : m;
    return m;
  }
};
