/*
 * ORACLE -- test infrastructure only (see oracle.h).
 *
 * The linear ADER-DG kernel family, the `isLinear` branch of the generated
 * glue (Toolkit/src/eu/exahype/solvers/templates/
 * AbstractGenericADERDGSolverImplementation.template:55-150,188-192):
 * Cauchy-Kowalevski predictor, volume and surface integrals of the
 * fluctuation form, and the Rusanov Riemann solver on the coefficient matrix.
 * Loop and summation orders are the reference's.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "../exahype_amd/csrc/dg_tables_data.h"

#define W(n, i) EXA_TAB_gl_weights[n][i]
#define MAXB 10

/* nonConservativeProduct of the linear scheme: a dim x nVar tensor
 * (kernels/aderdg/generic/c/3d/spaceTimePredictorLinear.cpph:28-30).
 * pde 4: BgradQ[d][v] = a_d gradQ[d][v]; pde 5: row 0 is the test product of
 * E/tests/kernels/c/GenericEulerKernelTest3D.cpp:84-106 (which writes only
 * nVar values), rows 1 and 2 zero. */
void orc_ncp_linear(const orc_cfg* c, const double* Q, const double* gradQ, double* T) {
  const int nv = c->nvar;
  for (int i = 0; i < 3 * nv; ++i) T[i] = 0.0;
  if (c->pde == 4) {
    for (int d = 0; d < 3; ++d)
      for (int v = 0; v < nv; ++v) T[d * nv + v] = c->adv[d] * gradQ[d * nv + v];
  } else if (c->pde == 5) {
    orc_ncp(c, Q, gradQ, T);
  }
}

/* coefficientMatrix(Q, d, Bn), Bn[nv][nv]: pde 4: a_d I; pde 5: zero, as the
 * reference's Euler kernel test (GenericEulerKernelTest3D.cpp:108-131). */
void orc_coefficient_matrix(const orc_cfg* c, const double* Q, int d, double* Bn) {
  (void)Q;
  const int nv = c->nvar;
  for (int i = 0; i < nv * nv; ++i) Bn[i] = 0.0;
  if (c->pde == 4)
    for (int v = 0; v < nv; ++v) Bn[v * nv + v] = c->adv[d];
}

/* The Lagrange basis polynomial of GL node i of the order's nodes on [0,1],
 * from its definition l_i(xi) = prod_{m != i} (xi - xi_m) / (xi_i - xi_m).
 * The reference evaluates generated polynomial literals
 * (kernels/DGBasisFunctions.cpp:327-...), which agree to rounding. */
double orc_lagrange(int order, int i, double xi) {
  const double* n = EXA_TAB_gl_nodes[order];
  double num = 1.0, den = 1.0;
  for (int m = 0; m <= order; ++m) {
    if (m == i) continue;
    num *= xi - n[m];
    den *= n[i] - n[m];
  }
  return num / den;
}

/* The user pointSource(x, t, dt, forceVector, x0) of the plugin surface
 * (E/solvers/UserSolverInterface.h:91-95) used by the point-source arm here:
 * a Ricker wavelet force amp[v] (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2,
 * located at src_x0 (independent of the node x and of dt). */
void orc_user_point_source(const orc_cfg* c, const double x[3], double t, double dt, double* force, double x0[3]) {
  (void)x;
  (void)dt;
  const double pa = M_PI * c->src_f0 * (t - c->src_t0);
  const double a = pa * pa;
  const double r = (1.0 - 2.0 * a) * exp(-a);
  for (int v = 0; v < c->nvar; ++v) force[v] = c->src_amp[v] * r;
  for (int d = 0; d < 3; ++d) x0[d] = c->src_x0[d];
}

/* kernels::aderdg::generic::c::pointSource (kernels/aderdg/generic/c/3d/
 * pointSource.cpph:28-145): for time level n (n = 0 at t, n >= 1 at the GL
 * time node n-1 of [t, t+dt]) and every spatial node, the user force at the
 * node, projected onto the node's basis function at the source position:
 * PSi = ((l_z(r_z) l_y(r_y)) l_x(r_x)) force / (((((w_z dz) w_y) dy) w_x) dx),
 * r = (x0 - (centre - dx/2)) / dx the source in reference coordinates; zero
 * when the source lies outside the cell. */
void orc_point_source(const orc_cfg* c, double t, double dt, const double centre[3], const double dx[3],
                      double* PSi) {
  const int N = c->order, B = N + 1, nv = c->nvar;
  const double* xi = EXA_TAB_gl_nodes[N];
  const double* w = EXA_TAB_gl_weights[N];
  double x[3], x0[3], f[8];
  for (int n = 0; n < B + 1; n++) {
    const double tn = n == 0 ? t : dt * xi[n - 1] + t;
    for (int i = 0; i < B; i++) {
      x[2] = centre[2] + dx[2] * (xi[i] - 0.5);
      for (int j = 0; j < B; j++) {
        x[1] = centre[1] + dx[1] * (xi[j] - 0.5);
        for (int k = 0; k < B; k++) {
          x[0] = centre[0] + dx[0] * (xi[k] - 0.5);
          orc_user_point_source(c, x, tn, dt, f, x0);
          double r[3];
          for (int d = 0; d < 3; d++) r[d] = (x0[d] - (centre[d] - dx[d] / 2.0)) / dx[d];
          double* out = PSi + ((((size_t)n * B + i) * B + j) * B + k) * nv;
          if (r[0] >= 0.0 && r[1] >= 0.0 && r[2] >= 0.0 && r[0] <= 1.0 && r[1] <= 1.0 && r[2] <= 1.0) {
            const double b = orc_lagrange(N, i, r[2]) * orc_lagrange(N, j, r[1]) * orc_lagrange(N, k, r[0]);
            const double den = w[i] * dx[2] * w[j] * dx[1] * w[k] * dx[0];
            for (int v = 0; v < nv; v++) out[v] = b * f[v] / den;
          } else {
            for (int v = 0; v < nv; v++) out[v] = 0.0;
          }
        }
      }
    }
  }
}

/* spaceTimePredictorLinear<usePointSource, useSource=false, useFlux=false,
 * useNCP=true> (kernels/aderdg/generic/c/3d/spaceTimePredictorLinear.cpph:
 * 33-443).  lQi [B+1][z][y][x][nData] (time derivatives) and lFi
 * [B][z][y][x][3][nVar] may be NULL; lQh [z][y][x][nData]; lFh
 * [4][z][y][x][nVar] (block 3, the source, zero).  PSi: the point forces of
 * orc_point_source, or NULL (usePointSource = false). */
void orc_space_time_predictor_linear(const orc_cfg* c, double* lQbnd, double* lFbnd, double* lQi, double* lFi,
                                     double* lQh, double* lFh, const double* luh, const double dx[3], double dt,
                                     const double* PSi) {
  /* per-thread scratch, allocated on first use */
  static _Thread_local double *sQi, *sFi, *gradQ;
  if (!sQi) {
    sQi = (double*)calloc((size_t)(MAXB + 1) * MAXB * MAXB * MAXB * 8, sizeof(double));
    sFi = (double*)calloc((size_t)MAXB * MAXB * MAXB * MAXB * 3 * 8, sizeof(double));
    gradQ = (double*)calloc((size_t)MAXB * MAXB * MAXB * MAXB * 3 * 8, sizeof(double));
  }
  const int N = c->order, B = N + 1, nv = c->nvar, nd = c->nvar + c->npar;
  const int B2 = B * B, B3 = B2 * B;
  if (!lQi) lQi = sQi;
  if (!lFi) lFi = sFi;
  const double* dudx = &EXA_TAB_dudx[N][0][0];
#define DUDX(a, b) dudx[(a) * MAXB + (b)]
#define QI(t, z, y, x, m) lQi[((((t) * B + (z)) * B + (y)) * B + (x)) * nd + (m)]
#define FI(t, z, y, x, d, m) lFi[(((((t) * B + (z)) * B + (y)) * B + (x)) * 3 + (d)) * nv + (m)]
#define GQ(t, z, y, x, d, m) gradQ[(((((t) * B + (z)) * B + (y)) * B + (x)) * 3 + (d)) * nv + (m)]
#define QH(z, y, x, m) lQh[(((z) * B + (y)) * B + (x)) * nd + (m)]
#define FH(d, z, y, x, m) lFh[((((d) * B + (z)) * B + (y)) * B + (x)) * nv + (m)]
  memset(lQi, 0, sizeof(double) * (size_t)(B + 1) * B3 * nd);
  /* the zeroth time derivative is the initial condition (:60-73) */
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) memcpy(&QI(0, i, j, k, 0), luh + ((i * B + j) * B + k) * nd, sizeof(double) * nd);
#define PS(n, z, y, x, m) PSi[((((n) * B + (z)) * B + (y)) * B + (x)) * nv + (m)]
  if (PSi) {
    /* the time integral of the force enters the zeroth derivative (:78-91),
     * the force at t the first (:93-102) */
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < nv; m++) {
            double sum = 0.0;
            for (int n = 1; n < B + 1; n++) sum = sum + W(N, n - 1) * PS(n, j, k, l, m);
            QI(0, j, k, l, m) += dt * sum;
          }
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int n = 0; n < nv; n++) QI(1, j, k, l, n) += PS(0, j, k, l, n);
  }
  memset(gradQ, 0, sizeof(double) * (size_t)B3 * B * 3 * nv);
  /* Cauchy-Kowalevski (:110-222) */
  for (int i = 0; i < B; i++) {
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < nv; m++)
            for (int n = 0; n < B; n++) GQ(i, j, k, l, 0, m) += 1.0 / dx[0] * QI(i, j, k, n, m) * DUDX(l, n);
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < nv; m++)
            for (int n = 0; n < B; n++) GQ(i, j, l, k, 1, m) += 1.0 / dx[1] * QI(i, j, n, k, m) * DUDX(l, n);
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < nv; m++)
            for (int n = 0; n < B; n++) GQ(i, l, j, k, 2, m) += 1.0 / dx[2] * QI(i, n, j, k, m) * DUDX(l, n);
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++) orc_ncp_linear(c, &QI(i, j, k, l, 0), &GQ(i, j, k, l, 0, 0), &FI(i, j, k, l, 0, 0));
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < 3; m++)
            for (int n = 0; n < nv; n++) QI(i + 1, j, k, l, n) -= FI(i, j, k, l, m, n);
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int n = nv; n < nd; n++) QI(i + 1, j, k, l, n) = QI(i, j, k, l, n);
  }
  /* time averages of the Taylor series (:223-285) */
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < nd; l++) QH(i, j, k, l) = QI(0, i, j, k, l);
  double dtavFac = 0.5 * dt;
  for (int i = 1; i < B; i++) {
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < nv; m++) QH(j, k, l, m) += dtavFac * QI(i, j, k, l, m);
    dtavFac *= dt / (i + 2);
  }
  memset(lFh + 3 * B3 * nv, 0, sizeof(double) * B3 * nv);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < nv; m++) FH(i, j, k, l, m) = FI(0, j, k, l, i, m);
  dtavFac = 0.5 * dt;
  for (int i = 1; i < B; i++) {
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < B; m++)
            for (int n = 0; n < nv; n++) FH(j, k, l, m, n) += dtavFac * FI(i, k, l, m, j, n);
    dtavFac *= dt / (i + 2);
  }
  if (PSi) {
    /* a third of the force's time integral in every direction's block (:287-302) */
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < B; m++)
            for (int n = 0; n < nv; n++) {
              double sum = 0.0;
              for (int n0 = 1; n0 < B + 1; n0++) sum = sum + W(N, n0 - 1) * PS(n0, k, l, m, n);
              FH(j, k, l, m, n) += sum / 3;
            }
  }
#undef PS
  /* boundary extrapolation (:305-440) */
  memset(lQbnd, 0, sizeof(double) * 6 * B2 * nd);
  memset(lFbnd, 0, sizeof(double) * 6 * B2 * nv);
  const double* FL = EXA_TAB_FLCoeff[N];
  const double* FR = EXA_TAB_FRCoeff[N];
#define QB(f, a, b, m) lQbnd[(((f) * B + (a)) * B + (b)) * nd + (m)]
#define FB(f, a, b, m) lFbnd[(((f) * B + (a)) * B + (b)) * nv + (m)]
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nd; k++)
        for (int l = 0; l < B; l++) {
          QB(0, i, j, k) += QH(i, j, l, k) * FL[l];
          QB(1, i, j, k) += QH(i, j, l, k) * FR[l];
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nv; k++)
        for (int l = 0; l < B; l++) {
          FB(0, i, j, k) += FH(0, i, j, l, k) * FL[l];
          FB(1, i, j, k) += FH(0, i, j, l, k) * FR[l];
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nd; k++)
        for (int l = 0; l < B; l++) {
          QB(2, i, j, k) += QH(i, l, j, k) * FL[l];
          QB(3, i, j, k) += QH(i, l, j, k) * FR[l];
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nv; k++)
        for (int l = 0; l < B; l++) {
          FB(2, i, j, k) += FH(1, i, l, j, k) * FL[l];
          FB(3, i, j, k) += FH(1, i, l, j, k) * FR[l];
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nd; k++)
        for (int l = 0; l < B; l++) {
          QB(4, i, j, k) += QH(l, i, j, k) * FL[l];
          QB(5, i, j, k) += QH(l, i, j, k) * FR[l];
        }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nv; k++)
        for (int l = 0; l < B; l++) {
          FB(4, i, j, k) += FH(2, l, i, j, k) * FL[l];
          FB(5, i, j, k) += FH(2, l, i, j, k) * FR[l];
        }
#undef DUDX
#undef QI
#undef FI
#undef GQ
#undef QH
#undef FH
#undef QB
#undef FB
}

/* volumeIntegralLinear (kernels/aderdg/generic/c/3d/volumeIntegralLinear.cpph:21-55):
 * lduh = -w^3 sum_d lFhi[d], lFhi [3][z][y][x][nVar]. */
void orc_volume_integral_linear(const orc_cfg* c, double* lduh, const double* lFhi, const double dx[3]) {
  (void)dx;
  const int N = c->order, B = N + 1, nv = c->nvar, B3 = B * B * B;
  memset(lduh, 0, sizeof(double) * B3 * nv);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = W(N, i) * W(N, j) * W(N, k);
        for (int l = 0; l < nv; l++)
          for (int m = 0; m < 3; m++)
            lduh[((i * B + j) * B + k) * nv + l] -= weight * lFhi[(((m * B + i) * B + j) * B + k) * nv + l];
      }
}

/* surfaceIntegralLinear (kernels/aderdg/generic/c/3d/surfaceIntegralLinear.cpph:24-90):
 * the fluctuations of both faces enter with a plus sign. */
void orc_surface_integral_linear(const orc_cfg* c, double* lduh, const double* lFbnd, const double dx[3]) {
  const int N = c->order, B = N + 1, nv = c->nvar;
  const double* FL = EXA_TAB_FLCoeff[N];
  const double* FR = EXA_TAB_FRCoeff[N];
#define IL(a, b, cc, m) ((((a) * B + (b)) * B + (cc)) * nv + (m))
#define IF(f, a, b, m) ((((f) * B + (a)) * B + (b)) * nv + (m))
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      const double weight = W(N, i) * W(N, j);
      const double updateSize = weight / dx[0];
      for (int k = 0; k < B; k++)
        for (int l = 0; l < nv; l++)
          lduh[IL(i, j, k, l)] -= (lFbnd[IF(1, i, j, l)] * FR[k] + lFbnd[IF(0, i, j, l)] * FL[k]) * updateSize;
    }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = W(N, i) * W(N, k);
        const double updateSize = weight / dx[1];
        for (int l = 0; l < nv; l++)
          lduh[IL(i, j, k, l)] -= (lFbnd[IF(3, i, k, l)] * FR[j] + lFbnd[IF(2, i, k, l)] * FL[j]) * updateSize;
      }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = W(N, j) * W(N, k);
        const double updateSize = weight / dx[2];
        for (int l = 0; l < nv; l++)
          lduh[IL(i, j, k, l)] -= (lFbnd[IF(5, j, k, l)] * FR[i] + lFbnd[IF(4, j, k, l)] * FL[i]) * updateSize;
      }
#undef IL
#undef IF
}

/* riemannSolverLinear<useNCP> (kernels/aderdg/generic/c/3d/riemannSolverLinear.cpph:24-140):
 * FL = (Bn - smax I)^T (QR - QL)/2, FR = (Bn + smax I)^T (QR - QL)/2 with Bn the
 * coefficient matrix at the mean of the face averages. */
void orc_riemann_solver_linear(const orc_cfg* c, double* FL, double* FR, const double* QL, const double* QR,
                               double dt, int direction) {
  (void)dt;
  const int N = c->order, B = N + 1, nv = c->nvar, nd = c->nvar + c->npar, B2 = B * B;
  double QavL[16] = {0.0}, QavR[16] = {0.0};
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      const double weight = W(N, i) * W(N, j);
      for (int k = 0; k < nd; k++) {
        QavL[k] += weight * QL[(i * B + j) * nd + k];
        QavR[k] += weight * QR[(i * B + j) * nd + k];
      }
    }
  double LL[16] = {0.0}, LR[16] = {0.0};
  orc_eigenvalues(c, QavL, direction, LL);
  orc_eigenvalues(c, QavR, direction, LR);
  double smax_L = fabs(LL[0]), smax_R = fabs(LR[0]);
  for (int v = 1; v < nv; ++v) {
    if (fabs(LL[v]) > smax_L) smax_L = fabs(LL[v]);
    if (fabs(LR[v]) > smax_R) smax_R = fabs(LR[v]);
  }
  const double smax = (smax_L < smax_R) ? smax_R : smax_L;
  double Qav[16];
  for (int k = 0; k < nd; ++k) Qav[k] = 0.5 * (QavL[k] + QavR[k]);
  double Bn[64], Bm[64], Bp[64];
  orc_coefficient_matrix(c, Qav, direction, Bn);
  for (int i = 0; i < nv * nv; ++i) Bm[i] = Bp[i] = Bn[i];
  for (int i = 0; i < nv; ++i) {
    Bm[i * nv + i] -= smax;
    Bp[i * nv + i] += smax;
  }
  double Q[MAXB * MAXB * 8];
  for (int i = 0; i < B2 * nd; ++i) Q[i] = 0.5 * (QR[i] - QL[i]);
  memset(FL, 0, sizeof(double) * B2 * nv);
  memset(FR, 0, sizeof(double) * B2 * nv);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < nv; k++)
        for (int l = 0; l < nv; l++) {
          FL[(i * B + j) * nv + k] += Bm[l * nv + k] * Q[(i * B + j) * nd + l];
          FR[(i * B + j) * nv + k] += Bp[l * nv + k] * Q[(i * B + j) * nd + l];
        }
}
