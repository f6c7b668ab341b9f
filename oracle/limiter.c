/*
 * ORACLE -- test infrastructure only (see oracle.h).  Plain-C restatement of
 * the reference's a-posteriori subcell limiter: the DG <-> FV projection
 * matrices (kernels/LimiterProjectionMatrices.cpp:39-233), the projections
 * (kernels/limiter/generic/c/Limiter.cpp:23-112), the cell min/max over
 * Gauss-Legendre nodes, Gauss-Lobatto nodes and subcell centres (:210-368) and
 * the discrete maximum principle (:114-165).  The observables are the first
 * nobs conserved variables (identity map).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "../exahype_amd/csrc/dg_tables_data.h"

#define MAXB 10

/* BaseFunc1D, kernels/LimiterProjectionMatrices.cpp:148-159 */
static void base_func_1d(double* phi, double xi, int B) {
  const double* x = EXA_TAB_gl_nodes[B - 1];
  for (int i = 0; i < B; i++) phi[i] = 1;
  for (int m = 0; m < B; m++)
    for (int j = 0; j < B; j++) {
      if (j == m) continue;
      phi[m] = phi[m] * (xi - x[j]) / (x[m] - x[j]);
    }
}

/* matrixInverse, kernels/LimiterProjectionMatrices.cpp:162-223, including the
 * pivot search quirk: the running maximum is replaced by the signed value
 * (mlV = c[j][i], no abs) once a larger magnitude is found (:181). */
static int matrix_inverse(int n, const double* a, double* ia) {
  double c[(MAXB + 1) * 2 * (MAXB + 1)];
#define C(i, j) c[(i) * 2 * n + (j)]
  for (int i = 0; i < n; i++) {
    int j;
    for (j = 0; j < n; j++) C(i, j) = a[j * n + i];
    for (; j < 2 * n; j++) C(i, j) = 0.;
    C(i, i + n) = 1.;
  }
  for (int i = 0; i < n; i++) {
    int ml = i;
    double mlV = fabs(C(i, i));
    for (int j = i + 1; j < n; j++) {
      if (fabs(C(j, i)) > mlV) {
        ml = j;
        mlV = C(j, i);
      }
    }
    for (int k = 0; k < 2 * n; k++) {
      const double tmp = C(ml, k);
      C(ml, k) = C(i, k);
      C(i, k) = tmp;
    }
    if (C(i, i) == 0) return -1;
    const double piv = 1. / C(i, i);
    for (int k = 0; k < 2 * n; k++) C(i, k) *= piv;
    for (int j = i + 1; j < n; j++) {
      const double tmp = C(j, i);
      for (int k = 0; k < 2 * n; k++) C(j, k) -= tmp * C(i, k);
    }
  }
  for (int i = n - 1; i >= 0; i--)
    for (int j = i - 1; j >= 0; j--) {
      const double tmp = C(j, i);
      for (int k = 0; k < 2 * n; k++) C(j, k) -= tmp * C(i, k);
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) ia[j * n + i] = C(i, j + n);
#undef C
  return 0;
}

/* initLimiterProjectionMatrices, kernels/LimiterProjectionMatrices.cpp:39-140.
 * uh2lim[k*L + i] (B x L), uh2lob[j*B + i] (B x B), lim2uh[i*B + j] (L x B). */
int orc_limiter_matrices(int N, double* uh2lim, double* uh2lob, double* lim2uh) {
  const int B = N + 1, L = 2 * N + 1;
  double phi[MAXB];
  memset(uh2lim, 0, sizeof(double) * B * L);
  const double dxi = 1. / L;
  for (int i = 0; i < L; i++) {
    const double xLeft = i * dxi;
    for (int j = 0; j < B; j++) {
      const double xi = xLeft + dxi * EXA_TAB_gl_nodes[B - 1][j];
      base_func_1d(phi, xi, B);
      for (int k = 0; k < B; k++) uh2lim[k * L + i] += EXA_TAB_gl_weights[B - 1][j] * phi[k];
    }
  }
  for (int i = 0; i < B; i++) {
    base_func_1d(phi, EXA_TAB_gll_nodes[B - 1][i], B);
    for (int j = 0; j < B; j++) uh2lob[j * B + i] = phi[j];
  }
  const int M = B + 1;
  double lsqm[(MAXB + 1) * (MAXB + 1)], ilsqm[(MAXB + 1) * (MAXB + 1)];
  double lsqrhs[(2 * MAXB + 1) * (MAXB + 1)];
  for (int i = 0; i < B; i++) {
    for (int j = 0; j < B; j++) {
      lsqm[i * M + j] = 0.;
      for (int k = 0; k < L; k++) lsqm[i * M + j] += 2 * uh2lim[i * L + k] * uh2lim[j * L + k];
    }
    lsqm[i * M + B] = EXA_TAB_gl_weights[B - 1][i];
  }
  for (int i = 0; i < B; i++) lsqm[B * M + i] = -EXA_TAB_gl_weights[B - 1][i];
  lsqm[B * M + B] = 0.;
  if (matrix_inverse(M, lsqm, ilsqm) != 0) return -1;
  for (int i = 0; i < L; i++) {
    for (int j = 0; j < B; j++) lsqrhs[i * M + j] = 2 * uh2lim[j * L + i];
    lsqrhs[i * M + B] = dxi;
  }
  for (int i = 0; i < L; i++)
    for (int j = 0; j < B; j++) {
      lim2uh[i * B + j] = 0.;
      for (int k = 0; k < M; k++) lim2uh[i * B + j] += ilsqm[k * M + j] * lsqrhs[i * M + k];
    }
  return 0;
}

/* projectOnFVLimiterSpace, kernels/limiter/generic/c/Limiter.cpp:23-66 (3D);
 * lim has side L+2*ghost, only the interior is written. */
void orc_project_on_fv(int N, int nvar, int ghost, const double* luh, double* lim) {
  const int B = N + 1, L = 2 * N + 1, S = L + 2 * ghost;
  double u2l[MAXB * (2 * MAXB + 1)], u2lob[MAXB * MAXB], l2u[(2 * MAXB + 1) * MAXB];
  orc_limiter_matrices(N, u2l, u2lob, l2u);
  for (int z = ghost; z < L + ghost; z++)
    for (int y = ghost; y < L + ghost; y++)
      for (int x = ghost; x < L + ghost; x++)
        for (int v = 0; v < nvar; v++) {
          double* o = &lim[(((size_t)z * S + y) * S + x) * nvar + v];
          *o = 0;
          for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
              for (int ix = 0; ix < B; ix++)
                *o += luh[((iz * B + iy) * B + ix) * nvar + v] * u2l[iz * L + (z - ghost)] *
                      u2l[iy * L + (y - ghost)] * u2l[ix * L + (x - ghost)];
        }
}

/* projectOnDGSpace, kernels/limiter/generic/c/Limiter.cpp:69-112 (3D) */
void orc_project_on_dg(int N, int nvar, int ghost, const double* lim, double* luh) {
  const int B = N + 1, L = 2 * N + 1, S = L + 2 * ghost;
  double u2l[MAXB * (2 * MAXB + 1)], u2lob[MAXB * MAXB], l2u[(2 * MAXB + 1) * MAXB];
  orc_limiter_matrices(N, u2l, u2lob, l2u);
  for (int z = 0; z < B; z++)
    for (int y = 0; y < B; y++)
      for (int x = 0; x < B; x++)
        for (int v = 0; v < nvar; v++) {
          double* o = &luh[((z * B + y) * B + x) * nvar + v];
          *o = 0;
          for (int iz = ghost; iz < L + ghost; iz++)
            for (int iy = ghost; iy < L + ghost; iy++)
              for (int ix = ghost; ix < L + ghost; ix++)
                *o += lim[(((size_t)iz * S + iy) * S + ix) * nvar + v] * l2u[(iz - ghost) * B + z] *
                      l2u[(iy - ghost) * B + y] * l2u[(ix - ghost) * B + x];
        }
}

/* findCellLocalMinAndMax with compareWithADERDGSolutionAtGaussLobattoNodes and
 * ...AtFVSubcellCenters, kernels/limiter/generic/c/Limiter.cpp:210-368 */
void orc_cell_min_max(int N, int nvar, int nobs, const double* luh, double* mn, double* mx) {
  const int B = N + 1, L = 2 * N + 1;
  double u2l[MAXB * (2 * MAXB + 1)], u2lob[MAXB * MAXB], l2u[(2 * MAXB + 1) * MAXB];
  orc_limiter_matrices(N, u2l, u2lob, l2u);
  for (int v = 0; v < nobs; v++) {
    mn[v] = DBL_MAX;
    mx[v] = -DBL_MAX;
  }
  for (int n = 0; n < B * B * B; n++)
    for (int v = 0; v < nobs; v++) {
      const double o = luh[n * nvar + v];
      mn[v] = o < mn[v] ? o : mn[v];
      mx[v] = mx[v] < o ? o : mx[v];
    }
  double val[16];
  for (int z = 0; z < B; z++)
    for (int y = 0; y < B; y++)
      for (int x = 0; x < B; x++) {
        for (int v = 0; v < nvar; v++) {
          val[v] = 0.0;
          for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
              for (int ix = 0; ix < B; ix++)
                val[v] += luh[((iz * B + iy) * B + ix) * nvar + v] * u2lob[iz * B + z] * u2lob[iy * B + y] *
                          u2lob[ix * B + x];
        }
        for (int v = 0; v < nobs; v++) {
          mn[v] = val[v] < mn[v] ? val[v] : mn[v];
          mx[v] = mx[v] < val[v] ? val[v] : mx[v];
        }
      }
  for (int z = 0; z < L; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        for (int v = 0; v < nvar; v++) {
          val[v] = 0.0;
          for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
              for (int ix = 0; ix < B; ix++)
                val[v] += luh[((iz * B + iy) * B + ix) * nvar + v] * u2l[iz * L + z] * u2l[iy * L + y] *
                          u2l[ix * L + x];
        }
        for (int v = 0; v < nobs; v++) {
          mn[v] = val[v] < mn[v] ? val[v] : mn[v];
          mx[v] = mx[v] < val[v] ? val[v] : mx[v];
        }
      }
}

/* discreteMaximumPrincipleAndMinAndMaxSearch, kernels/limiter/generic/c/Limiter.cpp:114-165.
 * bmin/bmax: [6][nobs] face-neighbour minima/maxima; on return block 0 holds the
 * cell's own min/max and block 1 the boundary min/max, as the reference writes. */
int orc_dmp(int N, int nvar, int nobs, const double* luh, double relax, double scaling, double* bmin,
            double* bmax) {
  double lmin[16], lmax[16];
  orc_cell_min_max(N, nvar, nobs, luh, lmin, lmax);
  int ok = 1;
  for (int v = 0; v < nobs; v++) {
    double boundaryMin = bmin[v];
    for (int i = 1; i < 6; i++) boundaryMin = bmin[i * nobs + v] < boundaryMin ? bmin[i * nobs + v] : boundaryMin;
    double boundaryMax = bmax[v];
    for (int i = 1; i < 6; i++) boundaryMax = boundaryMax < bmax[i * nobs + v] ? bmax[i * nobs + v] : boundaryMax;
    double scaledDifference = (boundaryMax - boundaryMin) * scaling;
    scaledDifference = scaledDifference < relax ? relax : scaledDifference;
    if ((lmin[v] < (boundaryMin - scaledDifference)) || (lmax[v] > (boundaryMax + scaledDifference))) ok = 0;
    bmin[v] = lmin[v];
    bmax[v] = lmax[v];
    bmin[v + nobs] = boundaryMin;
    bmax[v + nobs] = boundaryMax;
  }
  return ok;
}
