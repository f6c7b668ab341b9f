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

/* ---- periodic grid of FV patches (FiniteVolumesSolver step, §3.4 of the
 * survey: ghost filling E/solvers/FiniteVolumesSolver.cpp:659-687, update
 * :519-583, new dt :443-466) ---- */

/* ghostLayerFilling, kernels/finitevolumes/commons/c/commons.cpph:337-402 (2D), :876-982 (3D):
 * face ghost layers of width G copied from the neighbour's interior. */
static void fv_fill_boundaries(const orc_cfg* c, const orc_fv_grid* g, double dt);

static void fv_fill_ghosts(const orc_cfg* c, const orc_fv_grid* g, double dt) {
  const int nd = c->nvar + c->npar, P = g->patch, G = g->ghost, S = P + 2 * G;
  const int dim = g->dim;
  const size_t psz = (size_t)(dim == 3 ? S * S * S : S * S) * nd;
  const int nz = dim == 3 ? g->nz : 1;
  for (int pz = 0; pz < nz; ++pz)
    for (int py = 0; py < g->ny; ++py)
      for (int px = 0; px < g->nx; ++px) {
        double* me = g->luh + (((size_t)pz * g->ny + py) * g->nx + px) * psz;
        for (int d = 0; d < dim; ++d)
          for (int side = 0; side < 2; ++side) {
            int qx = px, qy = py, qz = pz;
            const int step = side ? 1 : -1;
            const int id = d == 0 ? px : (d == 1 ? py : pz), n = d == 0 ? g->nx : (d == 1 ? g->ny : nz);
            if (!g->periodic[d] && id == (side ? n - 1 : 0)) continue; /* a domain face: fv_fill_boundaries */
            if (d == 0) qx = (px + step + g->nx) % g->nx;
            if (d == 1) qy = (py + step + g->ny) % g->ny;
            if (d == 2) qz = (pz + step + nz) % nz;
            const double* nb = g->luh + (((size_t)qz * g->ny + qy) * g->nx + qx) * psz;
            const int zb = dim == 3 ? G : 0, ze = dim == 3 ? G + P : 1;
            for (int a = zb; a < ze; ++a)
              for (int b = G; b < G + P; ++b)
                for (int k = 0; k < G; ++k) {
                  /* my ghost index along d, neighbour's interior index along d */
                  const int mine = side ? G + P + k : k;
                  const int theirs = side ? G + k : P + k; /* patchBegin+k | patchEnd-G+k */
                  int ci[3], ni[3]; /* (x,y,z) */
                  if (dim == 2) {
                    /* 2D: a unused; (y,x) with b the transverse index */
                    if (d == 0) { ci[0] = mine; ci[1] = b; ni[0] = theirs; ni[1] = b; }
                    else { ci[0] = b; ci[1] = mine; ni[0] = b; ni[1] = theirs; }
                    memcpy(me + ((size_t)ci[1] * S + ci[0]) * nd, nb + ((size_t)ni[1] * S + ni[0]) * nd,
                           nd * sizeof(double));
                  } else {
                    if (d == 0) { ci[0] = mine; ci[1] = b; ci[2] = a; ni[0] = theirs; ni[1] = b; ni[2] = a; }
                    else if (d == 1) { ci[0] = b; ci[1] = mine; ci[2] = a; ni[0] = b; ni[1] = theirs; ni[2] = a; }
                    else { ci[0] = b; ci[1] = a; ci[2] = mine; ni[0] = b; ni[1] = a; ni[2] = theirs; }
                    memcpy(me + (((size_t)ci[2] * S + ci[1]) * S + ci[0]) * nd,
                           nb + (((size_t)ni[2] * S + ni[1]) * S + ni[0]) * nd, nd * sizeof(double));
                  }
                }
          }
      }
}

/* ---- boundary hooks of one patch (commons.cpph 2D :99-335, 3D :503-982) ---- */

/* subcell (x, y, z) of a patch of S = P + 2G subcells per axis */
static size_t fv_cell(int dim, int S, const int x[3]) {
  return dim == 3 ? ((size_t)x[2] * S + x[1]) * S + x[0] : (size_t)x[1] * S + x[0];
}

/* entry (layer k, slower transverse a, faster transverse b) of the face block
 * of direction d: 3D idx_bnd_x(P,P,G), idx_bnd_y(P,G,P), idx_bnd_z(G,P,P);
 * 2D idx_bnd_x(P,G), idx_bnd_y(G,P) */
static size_t fv_bnd(int dim, int d, int P, int G, int k, int a, int b) {
  if (dim == 2) return d == 0 ? (size_t)b * G + k : (size_t)k * P + b;
  if (d == 0) return ((size_t)a * P + b) * G + k;
  if (d == 1) return ((size_t)a * G + k) * P + b;
  return ((size_t)k * P + a) * P + b;
}

/* the transverse axes of direction d: o1 faster, o2 slower (3D) */
static void fv_axes(int dim, int d, int* o1, int* o2) {
  if (dim == 2) { *o1 = 1 - d; *o2 = 2; return; }
  *o1 = d == 0 ? 1 : 0;
  *o2 = d == 2 ? 1 : 2;
}

static int fv_face_dir(int dim, const int pos[3]) {
  for (int d = 0; d < dim; ++d)
    if (pos[d] != 0) return d;
  return -1;
}

void orc_fv_boundary_layer_extraction(const orc_cfg* c, int dim, int P, int G, double* luhbnd, const double* luh,
                                      const int pos[3]) {
  const int nd = c->nvar + c->npar, S = P + 2 * G, d = fv_face_dir(dim, pos);
  int o1, o2;
  if (d < 0) return;
  fv_axes(dim, d, &o1, &o2);
  const int na = dim == 3 ? P : 1;
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < P; ++b)
      for (int k = 0; k < G; ++k) {
        int x[3] = {0, 0, 0};
        x[o1] = G + b;
        if (dim == 3) x[o2] = G + a;
        /* + side: patchEnd-G+k; - side: patchBegin+k, except that the 3D y
         * and z faces read layer patchBegin for every k (commons.cpph:561, :591) */
        x[d] = pos[d] > 0 ? P + k : (dim == 3 && d > 0 ? G : G + k);
        memcpy(luhbnd + fv_bnd(dim, d, P, G, k, a, b) * nd, luh + fv_cell(dim, S, x) * nd, nd * sizeof(double));
      }
}

void orc_fv_ghost_layer_filling_at_boundary(const orc_cfg* c, int dim, int P, int G, double* luh, const double* luhbnd,
                                            const int pos[3]) {
  const int nd = c->nvar + c->npar, S = P + 2 * G, d = fv_face_dir(dim, pos);
  int o1, o2;
  if (d < 0) return;
  fv_axes(dim, d, &o1, &o2);
  const int na = dim == 3 ? P : 1;
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < P; ++b)
      for (int k = 0; k < G; ++k) {
        int x[3] = {0, 0, 0};
        x[o1] = G + b;
        if (dim == 3) x[o2] = G + a;
        x[d] = pos[d] > 0 ? G + P + k : k;
        memcpy(luh + fv_cell(dim, S, x) * nd, luhbnd + fv_bnd(dim, d, P, G, k, a, b) * nd, nd * sizeof(double));
      }
}

void orc_fv_ghost_layer_filling(const orc_cfg* c, int dim, int P, int G, double* luh, const double* luhNb,
                                const int pos[3]) {
  const int nd = c->nvar + c->npar, S = P + 2 * G, d = fv_face_dir(dim, pos);
  int o1, o2;
  if (d < 0) return;
  fv_axes(dim, d, &o1, &o2);
  const int na = dim == 3 ? P : 1;
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < P; ++b)
      for (int k = 0; k < G; ++k) {
        int x[3] = {0, 0, 0}, y[3];
        x[o1] = G + b;
        if (dim == 3) x[o2] = G + a;
        memcpy(y, x, sizeof(x));
        x[d] = pos[d] > 0 ? G + P + k : k; /* my ghost */
        y[d] = pos[d] > 0 ? G + k : P + k; /* the neighbour's patchBegin+k | patchEnd-G+k */
        memcpy(luh + fv_cell(dim, S, x) * nd, luhNb + fv_cell(dim, S, y) * nd, nd * sizeof(double));
      }
}

void orc_fv_boundary_values(const orc_cfg* c, const orc_bc* bc, const double x[3], double t, double dt,
                            int faceIndex, int d, const double* Qin, double* Qout) {
  const int nd = c->nvar + c->npar;
  (void)dt;
  (void)faceIndex;
  if (bc->kind == 3) {
    orc_ic_point(c, &bc->ic, x, t, Qout);
    for (int v = c->nvar; v < nd; ++v) Qout[v] = Qin[v];
    return;
  }
  for (int v = 0; v < nd; ++v) Qout[v] = Qin[v];
  if (bc->kind == 2) Qout[1 + d] = -Qin[1 + d];
}

void orc_fv_boundary_conditions(const orc_cfg* c, const orc_bc* bc, int dim, int P, int G, double* stateOut,
                                const double* stateIn, const double centre[3], const double size[3], double t,
                                double dt, int faceIndex, int normalNonZero) {
  const int nd = c->nvar + c->npar, side = faceIndex % 2, d = (faceIndex - side) / 2;
  int o1, o2;
  fv_axes(dim, d, &o1, &o2);
  const int na = dim == 3 ? P : 1;
  for (int k = 0; k < G; ++k) {
    const int K = G - k - 1; /* mirror */
    double x[3] = {0.0, 0.0, 0.0};
    /* the outside subcell centre, from K on the lower and k on the upper side */
    x[d] = side ? centre[d] + 0.5 * size[d] * (1.0 + k * 1.0 / (2.0 * P))
                : centre[d] - 0.5 * size[d] * (1.0 + K * 1.0 / (2.0 * P));
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < P; ++b) {
        const double qa = (2.0 * a + 1.0) / (2.0 * P), qb = (2.0 * b + 1.0) / (2.0 * P);
        if (dim == 3) x[o2] = centre[o2] + size[o2] * (qa - 0.5);
        /* the 3D x+ face scales its y coordinate by size[2] (commons.cpph:780) */
        x[o1] = centre[o1] + size[(dim == 3 && d == 0 && side) ? 2 : o1] * (qb - 0.5);
        orc_fv_boundary_values(c, bc, x, t, dt, faceIndex, normalNonZero,
                               stateIn + fv_bnd(dim, d, P, G, K, a, b) * nd,
                               stateOut + fv_bnd(dim, d, P, G, k, a, b) * nd);
      }
  }
}

/* the domain faces of the closed axes, as mergeWithBoundaryData fills them
 * (E/solvers/FiniteVolumesSolver.cpp:689-728): boundaryLayerExtraction,
 * boundaryConditions at the patch's time stamp and dt, then
 * ghostLayerFillingAtBoundary */
static void fv_fill_boundaries(const orc_cfg* c, const orc_fv_grid* g, double dt) {
  const int nd = c->nvar + c->npar, P = g->patch, G = g->ghost, S = P + 2 * G, dim = g->dim;
  const size_t psz = (size_t)(dim == 3 ? S * S * S : S * S) * nd;
  const int nz = dim == 3 ? g->nz : 1;
  const double h[3] = {g->width[0] / g->nx, g->width[1] / g->ny, g->width[2] / g->nz};
  double* bin = (double*)malloc((size_t)P * P * G * nd * sizeof(double));
  double* bout = (double*)malloc((size_t)P * P * G * nd * sizeof(double));
  for (int pz = 0; pz < nz; ++pz)
    for (int py = 0; py < g->ny; ++py)
      for (int px = 0; px < g->nx; ++px) {
        double* me = g->luh + (((size_t)pz * g->ny + py) * g->nx + px) * psz;
        const int id[3] = {px, py, pz}, n[3] = {g->nx, g->ny, nz};
        const double centre[3] = {g->offset[0] + (px + 0.5) * h[0], g->offset[1] + (py + 0.5) * h[1],
                                  dim == 3 ? g->offset[2] + (pz + 0.5) * h[2] : 0.0};
        for (int d = 0; d < dim; ++d)
          for (int side = 0; side < 2; ++side) {
            if (g->periodic[d] || id[d] != (side ? n[d] - 1 : 0)) continue;
            int pos[3] = {0, 0, 0};
            pos[d] = side ? 1 : -1;
            orc_fv_boundary_layer_extraction(c, dim, P, G, bin, me, pos);
            orc_fv_boundary_conditions(c, &g->bc, dim, P, G, bout, bin, centre, h, g->t, dt, 2 * d + side, d);
            orc_fv_ghost_layer_filling_at_boundary(c, dim, P, G, me, bout, pos);
          }
      }
  free(bin);
  free(bout);
}

int orc_fv_grid_alloc(const orc_cfg* c, orc_fv_grid* g, int dim, int scheme, int patch, int nx, int ny, int nz) {
  memset(g, 0, sizeof(*g));
  g->dim = dim;
  g->scheme = scheme;
  g->patch = patch;
  g->ghost = scheme == 1 ? 2 : 1;
  g->nx = nx; g->ny = ny; g->nz = dim == 3 ? nz : 1;
  for (int d = 0; d < 3; ++d) {
    g->width[d] = 1.0;
    g->periodic[d] = 1;
  }
  const int S = patch + 2 * g->ghost, nd = c->nvar + c->npar;
  const size_t psz = (size_t)(dim == 3 ? S * S * S : S * S) * nd;
  const size_t np = (size_t)nx * ny * g->nz;
  g->luh = (double*)calloc(np * psz, sizeof(double));
  g->luh_new = (double*)calloc(np * psz, sizeof(double));
  return g->luh && g->luh_new ? 0 : -1;
}

void orc_fv_grid_free(orc_fv_grid* g) {
  free(g->luh);
  free(g->luh_new);
  memset(g, 0, sizeof(*g));
}

/* solutionAdjustment at subcell centres, kernels/finitevolumes/commons/c/commons.cpph:69-97 (2D) */
void orc_fv_grid_init(const orc_cfg* c, orc_fv_grid* g, const orc_ic* ic) {
  const int nd = c->nvar + c->npar, P = g->patch, G = g->ghost, S = P + 2 * G, dim = g->dim;
  const size_t psz = (size_t)(dim == 3 ? S * S * S : S * S) * nd;
  const double hx = g->width[0] / g->nx, hy = g->width[1] / g->ny, hz = g->width[2] / g->nz;
  for (int pz = 0; pz < g->nz; ++pz)
    for (int py = 0; py < g->ny; ++py)
      for (int px = 0; px < g->nx; ++px) {
        double* me = g->luh + (((size_t)pz * g->ny + py) * g->nx + px) * psz;
        const double cx = g->offset[0] + (px + 0.5) * hx, cy = g->offset[1] + (py + 0.5) * hy,
                     cz = g->offset[2] + (pz + 0.5) * hz;
        const int zb = dim == 3 ? G : 0, ze = dim == 3 ? G + P : 1;
        for (int k = zb; k < ze; ++k)
          for (int j = G; j < G + P; ++j)
            for (int i = G; i < G + P; ++i) {
              const double qr = (2.0 * (i - G) + 1.0) / (2.0 * P);
              const double qs = (2.0 * (j - G) + 1.0) / (2.0 * P);
              const double qt = (2.0 * (k - G) + 1.0) / (2.0 * P);
              double x[3] = {cx + hx * (qr - 0.5), cy + hy * (qs - 0.5), dim == 3 ? cz + hz * (qt - 0.5) : 0.0};
              double* Q = me + (dim == 3 ? (((size_t)k * S + j) * S + i) : ((size_t)j * S + i)) * nd;
              orc_ic_point(c, ic, x, 0.0, Q);
            }
      }
}

/* solutionAdjustment of one 3D patch at its subcell centres
 * (kernels/finitevolumes/commons/c/commons.cpph:466-500): patch of P^3
 * subcells with ghost width G, cell centre and size */
void orc_fv_patch_ic(const orc_cfg* c, const orc_ic* ic, int P, int G, const double centre[3], const double dx[3],
                     double* luh) {
  const int nd = c->nvar + c->npar, S = P + 2 * G;
  for (int k = G; k < G + P; k++) {
    const double qk = (2.0 * (k - G) + 1.0) / (2.0 * P);
    for (int j = G; j < G + P; j++) {
      const double qj = (2.0 * (j - G) + 1.0) / (2.0 * P);
      for (int i = G; i < G + P; i++) {
        const double qi = (2.0 * (i - G) + 1.0) / (2.0 * P);
        double x[3] = {centre[0] + dx[0] * (qi - 0.5), centre[1] + dx[1] * (qj - 0.5), centre[2] + dx[2] * (qk - 0.5)};
        orc_ic_point(c, ic, x, 0.0, luh + (((size_t)k * S + j) * S + i) * nd);
      }
    }
  }
}

double orc_fv_grid_stable_dt(const orc_cfg* c, const orc_fv_grid* g, int nthreads) {
  const int nd = c->nvar + c->npar, S = g->patch + 2 * g->ghost;
  const size_t psz = (size_t)(g->dim == 3 ? S * S * S : S * S) * nd;
  const int np = g->nx * g->ny * g->nz;
  const double dx[3] = {g->width[0] / g->nx, g->width[1] / g->ny, g->width[2] / g->nz};
  double dtmin = DBL_MAX;
#pragma omp parallel for num_threads(nthreads) reduction(min : dtmin) schedule(static)
  for (int p = 0; p < np; ++p) {
    const double d = orc_fv_stable_dt(c, g->dim, g->patch, g->ghost, g->luh + (size_t)p * psz, dx, c->cfl);
    if (d < dtmin) dtmin = d;
  }
  return dtmin;
}

/* one step: ghost filling, update with dt (luh_new starts as a copy of luh),
 * swap; returns the min over patches of the kernels' admissible dt through
 * *dt_admissible and the stable dt of the new state. */
double orc_fv_grid_step(const orc_cfg* c, orc_fv_grid* g, double dt, double* dt_admissible, int nthreads) {
  const int nd = c->nvar + c->npar, S = g->patch + 2 * g->ghost;
  const size_t psz = (size_t)(g->dim == 3 ? S * S * S : S * S) * nd;
  const int np = g->nx * g->ny * g->nz;
  const double dx[3] = {g->width[0] / g->nx, g->width[1] / g->ny, g->width[2] / g->nz};
  fv_fill_ghosts(c, g, dt);
  fv_fill_boundaries(c, g, dt);
  memcpy(g->luh_new, g->luh, np * psz * sizeof(double));
  double adm = DBL_MAX;
#pragma omp parallel for num_threads(nthreads) reduction(min : adm) schedule(dynamic, 4)
  for (int p = 0; p < np; ++p) {
    double a;
    if (g->scheme == 1)
      a = orc_musclhancock_update(c, g->patch, g->luh_new + (size_t)p * psz, g->luh + (size_t)p * psz, dx, dt, 0.99);
    else
      a = orc_godunov_update(c, g->dim, g->patch, g->luh_new + (size_t)p * psz, g->luh + (size_t)p * psz, dx, dt,
                             c->cfl);
    if (a < adm) adm = a;
  }
  double* t = g->luh;
  g->luh = g->luh_new;
  g->luh_new = t;
  if (dt_admissible) *dt_admissible = adm;
  g->t += dt;
  return orc_fv_grid_stable_dt(c, g, nthreads);
}
