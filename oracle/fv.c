/*
 * ORACLE -- test infrastructure only (see oracle.h).  Plain-C restatement of
 * the reference's finite-volume kernels (Rusanov, Godunov, MUSCL-Hancock,
 * FV CFL time step) in 2D and 3D, plus the EulerFV demonstrator PDE.
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

/* EulerFV demonstrator PDE (2D, rho, j[3], E), Demonstrators/EulerFV/MyEulerSolver.cpp:51-85.
 * p = (g-1)(E - ((0.5*irho)*j).j); F[d][1+i] = irho*(j_i j_d) + p*delta_id (generated
 * Fluxes::j(Matrix), Toolkit/src/eu/exahype/variables/Variables.java:629-636). */
static double eulerfv_pressure(double g, const double* Q, double irho) {
  const double h = 0.5 * irho;
  const double jj = (h * Q[1]) * Q[1] + (h * Q[2]) * Q[2] + (h * Q[3]) * Q[3];
  return (g - 1) * (Q[4] - jj);
}

static void fv_flux(const orc_cfg* c, int dim, const double* Q, double* F) {
  if (c->pde != 2) { orc_flux(c, Q, F); return; }
  const double irho = 1. / Q[0];
  const double p = eulerfv_pressure(c->gamma, Q, irho);
  for (int d = 0; d < dim; ++d) {
    double* f = F + d * c->nvar;
    f[0] = Q[1 + d];
    for (int i = 0; i < 3; ++i) f[1 + i] = irho * (Q[1 + i] * Q[1 + d]) + p * (i == d ? 1.0 : 0.0);
    f[4] = (irho * (Q[4] + p)) * Q[1 + d];
  }
}

static void fv_eigenvalues(const orc_cfg* c, const double* Q, int d, double* lambda) {
  if (c->pde != 2) { orc_eigenvalues(c, Q, d, lambda); return; }
  const double GAMMA = c->gamma;
  const double irho = 1. / Q[0];
  const double p = eulerfv_pressure(GAMMA, Q, irho);
  const double u_n = Q[d + 1] * irho;
  const double cs = sqrt(GAMMA * p * irho);
  lambda[0] = u_n - cs;
  lambda[1] = u_n;
  lambda[2] = u_n;
  lambda[3] = u_n;
  lambda[4] = u_n + cs;
}

/* rusanov<useNCP=false,useFlux=true>, kernels/finitevolumes/riemannsolvers/c/rusanov.cpph:24-83 (2D), :98-156 (3D) */
double orc_fv_rusanov(const orc_cfg* c, int dim, double* fnL, double* fnR,
                      const double* qL, const double* qR, int normalNonZero) {
  const int nv = c->nvar;
  double sL[16] = {0.0}, sR[16] = {0.0};
  fv_eigenvalues(c, qL, normalNonZero, sL);
  fv_eigenvalues(c, qR, normalNonZero, sR);
  double s_max = -1.0;
  for (int i = 0; i < nv; i++) { const double a = fabs(sL[i]); s_max = (a < s_max) ? s_max : a; }
  for (int i = 0; i < nv; i++) { const double a = fabs(sR[i]); s_max = (a < s_max) ? s_max : a; }
  double FL[3 * 16] = {0.0}, FR[3 * 16] = {0.0};
  fv_flux(c, dim, qL, FL);
  fv_flux(c, dim, qR, FR);
  for (int i = 0; i < nv; i++) {
    fnL[i] = 0.5 * s_max * (qL[i] - qR[i]);
    fnL[i] += 0.5 * (FL[normalNonZero * nv + i] + FR[normalNonZero * nv + i]);
    fnR[i] = fnL[i];
  }
  return s_max;
}

/* commons::stableTimeStepSize, kernels/finitevolumes/commons/c/commons.cpph:26-67 (2D), :410-463 (3D) */
double orc_fv_stable_dt(const orc_cfg* c, int dim, int patch, int ghost,
                        const double* luh, const double* dx, double cfl) {
  const int nd = c->nvar + c->npar, nv = c->nvar, S = patch + 2 * ghost;
  const int b = ghost, e = ghost + patch;
  double lambda[16];
  double dt_max_stable = DBL_MAX;
  const int zb = dim == 3 ? b : 0, ze = dim == 3 ? e : 1;
  for (int i = zb; i < ze; i++)
    for (int j = b; j < e; j++)
      for (int k = b; k < e; k++) {
        const double* q = luh + ((dim == 3 ? (i * S + j) * S + k : j * S + k)) * nd;
        for (int d = 0; d < dim; ++d) {
          fv_eigenvalues(c, q, d, lambda);
          double smax = fabs(lambda[0]);
          for (int v = 1; v < nv; ++v) if (smax < fabs(lambda[v])) smax = fabs(lambda[v]);
          const double cand = cfl / dim * (dx[d] / patch) / smax;
          dt_max_stable = (cand < dt_max_stable) ? cand : dt_max_stable;
        }
      }
  return dt_max_stable;
}

/* godunov::solutionUpdate<useSource=false>, kernels/finitevolumes/godunov/c/2d/godunov.cpph:33-120,
 * kernels/finitevolumes/godunov/c/3d/godunov.cpph:35-143.  luh_new must hold a copy of luh. */
double orc_godunov_update(const orc_cfg* c, int dim, int patch, double* luh_new,
                          const double* luh, const double* dx, double dt, double cfl) {
  const int nd = c->nvar + c->npar, nv = c->nvar, S = patch + 2;
  const int b = 1, e = 1 + patch;
  double h[3];
  for (int d = 0; d < dim; ++d) h[d] = dx[d] / patch;
  double fL[16], fR[16];
  double dt_max = DBL_MAX;
  if (dim == 2) {
#define I2(y, x) (((y) * S + (x)) * nd)
    for (int j = b; j < e; j++)
      for (int k = b - 1; k < e; k++) {
        const double s = orc_fv_rusanov(c, 2, fL, fR, luh + I2(j, k), luh + I2(j, k + 1), 0);
        const double cand = cfl / 2 * h[0] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[I2(j, k) + l] -= dt / h[0] * fL[l];
          luh_new[I2(j, k + 1) + l] += dt / h[0] * fR[l];
        }
      }
    for (int j = b - 1; j < e; j++)
      for (int k = b; k < e; k++) {
        const double s = orc_fv_rusanov(c, 2, fL, fR, luh + I2(j, k), luh + I2(j + 1, k), 1);
        const double cand = cfl / 2 * h[1] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[I2(j, k) + l] -= dt / h[1] * fL[l];
          luh_new[I2(j + 1, k) + l] += dt / h[1] * fR[l];
        }
      }
#undef I2
    return dt_max;
  }
#define I3(z, y, x) ((((z) * S + (y)) * S + (x)) * nd)
  for (int i = b; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b - 1; k < e; k++) {
        const double s = orc_fv_rusanov(c, 3, fL, fR, luh + I3(i, j, k), luh + I3(i, j, k + 1), 0);
        const double cand = cfl / 3 * h[0] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[I3(i, j, k) + l] -= dt / h[0] * fL[l];
          luh_new[I3(i, j, k + 1) + l] += dt / h[0] * fR[l];
        }
      }
  for (int i = b; i < e; i++)
    for (int j = b - 1; j < e; j++)
      for (int k = b; k < e; k++) {
        const double s = orc_fv_rusanov(c, 3, fL, fR, luh + I3(i, j, k), luh + I3(i, j + 1, k), 1);
        const double cand = cfl / 3 * h[1] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[I3(i, j, k) + l] -= dt / h[1] * fL[l];
          luh_new[I3(i, j + 1, k) + l] += dt / h[1] * fR[l];
        }
      }
  for (int i = b - 1; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b; k < e; k++) {
        const double s = orc_fv_rusanov(c, 3, fL, fR, luh + I3(i, j, k), luh + I3(i + 1, j, k), 2);
        const double cand = cfl / 3 * h[2] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[I3(i, j, k) + l] -= dt / h[2] * fL[l];
          luh_new[I3(i + 1, j, k) + l] += dt / h[2] * fR[l];
        }
      }
#undef I3
  return dt_max;
}

/* minmod, kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph:27-40 */
static double minmod(double a, double b) {
  if (a * b < 0.0) return 0.0;
  return (fabs(a) < fabs(b)) ? a : b;
}

/* musclhancock::solutionUpdate<useSource=false,useNCP=false,useFlux=true>,
 * kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph:50-309 (ghost width 2).
 * Boundary states of the first ghost layer are extrapolated but not evolved
 * by the half step (:156-213 loops over interior cells only). */
double orc_musclhancock_update(const orc_cfg* c, int patch, double* luh_new,
                               const double* luh, const double* dx, double dt, double cfl) {
  const int nd = c->nvar + c->npar, nv = c->nvar, G = 2, S = patch + 2 * G;
  const int b = G, e = G + patch;
  const double h[3] = {dx[0] / patch, dx[1] / patch, dx[2] / patch};
  const size_t nvp = (size_t)S * S * S * nv;
  double* slope = (double*)calloc(nvp * 3, sizeof(double));
  double* w = (double*)calloc(nvp * 6, sizeof(double));
  double* slopet = (double*)calloc((size_t)S * S * S * nd, sizeof(double));
  double *wLx = w, *wRx = w + nvp, *wLy = w + 2 * nvp, *wRy = w + 3 * nvp, *wLz = w + 4 * nvp, *wRz = w + 5 * nvp;
#define IX(z, y, x) ((((z) * S + (y)) * S + (x)) * nd)
#define IW(z, y, x) ((((z) * S + (y)) * S + (x)) * nv)
#define IS(z, y, x, d) (((((z) * S + (y)) * S + (x)) * 3 + (d)) * nv)
  for (int i = b; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b - 1; k < e + 1; k++)
        for (int l = 0; l < nv; l++)
          slope[IS(i, j, k, 0) + l] = minmod(luh[IX(i, j, k + 1) + l] - luh[IX(i, j, k) + l],
                                             luh[IX(i, j, k) + l] - luh[IX(i, j, k - 1) + l]) / h[0];
  for (int i = b; i < e; i++)
    for (int j = b - 1; j < e + 1; j++)
      for (int k = b; k < e; k++)
        for (int l = 0; l < nv; l++)
          slope[IS(i, j, k, 1) + l] = minmod(luh[IX(i, j + 1, k) + l] - luh[IX(i, j, k) + l],
                                             luh[IX(i, j, k) + l] - luh[IX(i, j - 1, k) + l]) / h[1];
  for (int i = b - 1; i < e + 1; i++)
    for (int j = b; j < e; j++)
      for (int k = b; k < e; k++)
        for (int l = 0; l < nv; l++)
          slope[IS(i, j, k, 2) + l] = minmod(luh[IX(i + 1, j, k) + l] - luh[IX(i, j, k) + l],
                                             luh[IX(i, j, k) + l] - luh[IX(i - 1, j, k) + l]) / h[2];
  for (int i = b; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b - 1; k < e + 1; k++)
        for (int l = 0; l < nv; l++) {
          wLx[IW(i, j, k) + l] = luh[IX(i, j, k) + l] - 0.5 * h[0] * slope[IS(i, j, k, 0) + l];
          wRx[IW(i, j, k) + l] = luh[IX(i, j, k) + l] + 0.5 * h[0] * slope[IS(i, j, k, 0) + l];
        }
  for (int i = b; i < e; i++)
    for (int j = b - 1; j < e + 1; j++)
      for (int k = b; k < e; k++)
        for (int l = 0; l < nv; l++) {
          wLy[IW(i, j, k) + l] = luh[IX(i, j, k) + l] - 0.5 * h[1] * slope[IS(i, j, k, 1) + l];
          wRy[IW(i, j, k) + l] = luh[IX(i, j, k) + l] + 0.5 * h[1] * slope[IS(i, j, k, 1) + l];
        }
  for (int i = b - 1; i < e + 1; i++)
    for (int j = b; j < e; j++)
      for (int k = b; k < e; k++)
        for (int l = 0; l < nv; l++) {
          wLz[IW(i, j, k) + l] = luh[IX(i, j, k) + l] - 0.5 * h[2] * slope[IS(i, j, k, 2) + l];
          wRz[IW(i, j, k) + l] = luh[IX(i, j, k) + l] + 0.5 * h[2] * slope[IS(i, j, k, 2) + l];
        }
  /* half-step evolution of the boundary states (:156-213) */
  double F[3 * 16];
  for (int i = b; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b; k < e; k++) {
        double* Qt = slopet + IX(i, j, k);
        for (int l = 0; l < nv; ++l) Qt[l] = 0.0;
        orc_flux(c, wLx + IW(i, j, k), F);
        for (int l = 0; l < nv; l++) Qt[l] += F[0 * nv + l] / h[0];
        orc_flux(c, wRx + IW(i, j, k), F);
        for (int l = 0; l < nv; l++) Qt[l] -= F[0 * nv + l] / h[0];
        orc_flux(c, wLy + IW(i, j, k), F);
        for (int l = 0; l < nv; l++) Qt[l] += F[1 * nv + l] / h[1];
        orc_flux(c, wRy + IW(i, j, k), F);
        for (int l = 0; l < nv; l++) Qt[l] -= F[1 * nv + l] / h[1];
        orc_flux(c, wLz + IW(i, j, k), F);
        for (int l = 0; l < nv; l++) Qt[l] += F[2 * nv + l] / h[2];
        orc_flux(c, wRz + IW(i, j, k), F);
        for (int l = 0; l < nv; l++) Qt[l] -= F[2 * nv + l] / h[2];
        for (int l = 0; l < nv; l++) {
          wLx[IW(i, j, k) + l] += 0.5 * dt * Qt[l];
          wRx[IW(i, j, k) + l] += 0.5 * dt * Qt[l];
          wLy[IW(i, j, k) + l] += 0.5 * dt * Qt[l];
          wRy[IW(i, j, k) + l] += 0.5 * dt * Qt[l];
          wLz[IW(i, j, k) + l] += 0.5 * dt * Qt[l];
          wRz[IW(i, j, k) + l] += 0.5 * dt * Qt[l];
        }
      }
  /* Riemann problems (:215-281) */
  double fL[16], fR[16];
  double dt_max = DBL_MAX;
  for (int i = b; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b - 1; k < e; k++) {
        const double s = orc_fv_rusanov(c, 3, fL, fR, wRx + IW(i, j, k), wLx + IW(i, j, k + 1), 0);
        const double cand = cfl / 3 * h[0] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[IX(i, j, k) + l] -= dt / h[0] * fL[l];
          luh_new[IX(i, j, k + 1) + l] += dt / h[0] * fR[l];
        }
      }
  for (int i = b; i < e; i++)
    for (int j = b - 1; j < e; j++)
      for (int k = b; k < e; k++) {
        const double s = orc_fv_rusanov(c, 3, fL, fR, wRy + IW(i, j, k), wLy + IW(i, j + 1, k), 1);
        const double cand = cfl / 3 * h[1] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[IX(i, j, k) + l] -= dt / h[1] * fL[l];
          luh_new[IX(i, j + 1, k) + l] += dt / h[1] * fR[l];
        }
      }
  for (int i = b - 1; i < e; i++)
    for (int j = b; j < e; j++)
      for (int k = b; k < e; k++) {
        const double s = orc_fv_rusanov(c, 3, fL, fR, wRz + IW(i, j, k), wLz + IW(i + 1, j, k), 2);
        const double cand = cfl / 3 * h[2] / s;
        dt_max = (cand < dt_max) ? cand : dt_max;
        for (int l = 0; l < nv; ++l) {
          luh_new[IX(i, j, k) + l] -= dt / h[2] * fL[l];
          luh_new[IX(i + 1, j, k) + l] += dt / h[2] * fR[l];
        }
      }
#undef IX
#undef IW
#undef IS
  free(slope);
  free(w);
  free(slopet);
  return dt_max;
}
