// Two PDEs written by a "project" against include/exahype_amd_pde.h, for
// the registration tests (built by tests/user_pde/build.py into
// tests/user_pde/_build/libuser_pdes.so, linked against the library):
//
// * UserAdvection: linear advection in flux form with the velocity
//   (0.7, -0.4, 0.3) -- the arithmetic of the library's EXA_PDE_ADVECTION
//   with that velocity -- whose initial state is the sine-mode state of
//   exa_ic kind 3 (one mode) and whose boundary values are that state
//   advected to the face node (with its flux);
// * UserEuler: 3D Euler written as the reference's Euler kernel test writes
//   it (E/tests/kernels/c/GenericEulerKernelTest3D.cpp:26-82), initial state
//   an entropy wave, reflecting walls as boundary values.
#include "exahype_amd_pde.h"

namespace {

constexpr double kA[3] = {0.7, -0.4, 0.3};
constexpr double kPi = 3.14159265358979323846;

struct UserAdvection {
  static constexpr int NumberOfVariables = 5;
  static constexpr bool useFlux = true, useSource = false, useNCP = false;
  __device__ static void flux(const double* Q, double** F) {
    for (int v = 0; v < 5; ++v) {
      F[0][v] = kA[0] * Q[v];
      F[1][v] = kA[1] * Q[v];
      F[2][v] = kA[2] * Q[v];
    }
  }
  __device__ static void eigenvalues(const double*, int d, double* lambda) {
    for (int v = 0; v < 5; ++v) lambda[v] = kA[d];
  }
  __device__ static void algebraicSource(const double*, double*) {}
  __device__ static void nonConservativeProduct(const double*, const double*, double*) {}
  // Q_v = 1 + (v+1) a sin(2 pi k.(x - u t) + phi), k = (1, 2, 1), a = 0.1, phi = 0.4
  __device__ static void state(const double* x, double t, double* Q) {
    double s = 0.0;
    s += 0.1 * sin(2.0 * kPi * (1.0 * (x[0] - kA[0] * t) + 2.0 * (x[1] - kA[1] * t) + 1.0 * (x[2] - kA[2] * t)) + 0.4);
    for (int v = 0; v < 5; ++v) Q[v] = 1.0 + (v + 1) * s;
  }
  __device__ static void adjustPointSolution(const double* x, double, double t, double, double* Q) { state(x, t, Q); }
  __device__ static void boundaryValues(const double* x, double t, double, int, int d, const double*, const double*,
                                        double* fluxOut, double* stateOut) {
    state(x, t, stateOut);
    for (int v = 0; v < 5; ++v) fluxOut[v] = kA[d] * stateOut[v];
  }
};

constexpr double kGamma = 1.4;

struct UserEuler {
  static constexpr int NumberOfVariables = 5;
  static constexpr bool useFlux = true, useSource = false, useNCP = false;
  __device__ static void flux(const double* Q, double** F) {
    const double irho = 1.0 / Q[0];
    const double p = (kGamma - 1) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    for (int d = 0; d < 3; ++d) {
      F[d][0] = Q[1 + d];
      for (int i = 0; i < 3; ++i) F[d][1 + i] = irho * Q[1 + d] * Q[1 + i] + (i == d ? p : 0.0);
      F[d][4] = irho * Q[1 + d] * (Q[4] + p);
    }
  }
  __device__ static void eigenvalues(const double* Q, int d, double* lambda) {
    const double irho = 1.0 / Q[0];
    const double p = (kGamma - 1) * (Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
    const double u_n = Q[d + 1] * irho;
    const double c = sqrt(kGamma * p * irho);
    lambda[0] = u_n - c;
    lambda[1] = u_n;
    lambda[2] = u_n;
    lambda[3] = u_n;
    lambda[4] = u_n + c;
  }
  __device__ static void algebraicSource(const double*, double*) {}
  __device__ static void nonConservativeProduct(const double*, const double*, double*) {}
  // entropy wave: rho = 1 + 0.2 sin(2 pi k.(x - u t) + 0.4), k = (1, 2, 1), u = (0.7, -0.4, 0.3), p = 1
  __device__ static void adjustPointSolution(const double* x, double, double t, double, double* Q) {
    const double u[3] = {0.7, -0.4, 0.3};
    const double s = sin(2.0 * kPi * (1.0 * (x[0] - u[0] * t) + 2.0 * (x[1] - u[1] * t) + 1.0 * (x[2] - u[2] * t)) + 0.4);
    const double rho = 1.0 + 0.2 * s;
    Q[0] = rho;
    Q[1] = rho * u[0];
    Q[2] = rho * u[1];
    Q[3] = rho * u[2];
    Q[4] = 1.0 / (kGamma - 1.0) + 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  }
  // reflecting wall: the inside state with its normal momentum negated
  __device__ static void boundaryValues(const double*, double, double, int, int d, const double*,
                                        const double* stateIn, double* fluxOut, double* stateOut) {
    for (int v = 0; v < 5; ++v) stateOut[v] = stateIn[v];
    stateOut[1 + d] = -stateIn[1 + d];
    double F[3][5];
    double* rows[3] = {F[0], F[1], F[2]};
    flux(stateOut, rows);
    for (int v = 0; v < 5; ++v) fluxOut[v] = F[d][v];
  }
};

}  // namespace

// ids[0]: UserAdvection, ids[1]: UserEuler
extern "C" int user_pdes_register(int* ids) {
  int r = exa::register_pde<UserAdvection>("user_advection", &ids[0]);
  if (r == EXA_OK) r = exa::register_pde<UserEuler>("user_euler", &ids[1]);
  return r;
}
