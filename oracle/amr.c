/*
 * ORACLE -- test infrastructure only (see oracle.h).  Plain-C restatement of
 * the reference's 3D ADER-DG level transfers between a coarse cell and one of
 * its 3^(levelDelta) descendants per axis
 * (kernels/aderdg/generic/c/3d/amrRoutines.cpph:24-493):
 *
 *   face prolongation   :24-144   face restriction   :146-258
 *   volume prolongation :260-374  volume restriction :376-493
 *
 * Single-level operators are the full (not sum-factorised) tensor products of
 * fineGridProjector1d[N][sub][coarse][fine], accumulated in the reference's
 * loop order.  Level jumps > 1 are applied one level at a time: prolongation
 * takes the base-3 digits of the sub-index from the most significant down,
 * restriction from the least significant up.  Prolongation overwrites the
 * fine array, restriction adds into the coarse array.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "../exahype_amd/csrc/dg_tables_data.h"

static double P(int N, int sub, int c, int f) { return EXA_TAB_fineGridProjector1d[N][sub][c][f]; }

/* singleLevelFaceUnknownsProlongation, amrRoutines.cpph:24-43 */
static void face_prolong_1(int N, int nd, double* fine, const double* coarse, const int sub[2]) {
  const int B = N + 1;
  for (int m2 = 0; m2 < B; m2++)
    for (int m1 = 0; m1 < B; m1++)
      for (int v = 0; v < nd; v++)
        for (int n2 = 0; n2 < B; n2++)
          for (int n1 = 0; n1 < B; n1++)
            fine[(m2 * B + m1) * nd + v] +=
                coarse[(n2 * B + n1) * nd + v] * P(N, sub[0], n1, m1) * P(N, sub[1], n2, m2);
}

/* singleLevelFaceUnknownsRestriction, amrRoutines.cpph:146-171 */
static void face_restrict_1(int N, int nd, double* coarse, const double* fine, const int sub[2]) {
  const int B = N + 1;
  const double* w = EXA_TAB_gl_weights[N];
  for (int m2 = 0; m2 < B; m2++)
    for (int m1 = 0; m1 < B; m1++)
      for (int v = 0; v < nd; v++)
        for (int n2 = 0; n2 < B; n2++)
          for (int n1 = 0; n1 < B; n1++)
            coarse[(m2 * B + m1) * nd + v] += w[n1] * w[n2] * P(N, sub[0], m1, n1) * P(N, sub[1], m2, n2) *
                                              fine[(n2 * B + n1) * nd + v] / w[m1] / w[m2] / 9.0;
}

/* singleLevelVolumeUnknownsProlongation, amrRoutines.cpph:260-287 */
static void vol_prolong_1(int N, int nd, double* fine, const double* coarse, const int sub[3]) {
  const int B = N + 1;
  for (int m3 = 0; m3 < B; m3++)
    for (int m2 = 0; m2 < B; m2++)
      for (int m1 = 0; m1 < B; m1++)
        for (int v = 0; v < nd; v++)
          for (int n3 = 0; n3 < B; n3++)
            for (int n2 = 0; n2 < B; n2++)
              for (int n1 = 0; n1 < B; n1++)
                fine[((m3 * B + m2) * B + m1) * nd + v] += coarse[((n3 * B + n2) * B + n1) * nd + v] *
                                                           P(N, sub[0], n1, m1) * P(N, sub[1], n2, m2) *
                                                           P(N, sub[2], n3, m3);
}

/* singleLevelVolumeUnknownsRestriction, amrRoutines.cpph:376-408 */
static void vol_restrict_1(int N, int nd, double* coarse, const double* fine, const int sub[3]) {
  const int B = N + 1;
  const double* w = EXA_TAB_gl_weights[N];
  for (int m3 = 0; m3 < B; m3++)
    for (int m2 = 0; m2 < B; m2++)
      for (int m1 = 0; m1 < B; m1++)
        for (int v = 0; v < nd; v++)
          for (int n3 = 0; n3 < B; n3++)
            for (int n2 = 0; n2 < B; n2++)
              for (int n1 = 0; n1 < B; n1++)
                coarse[((m3 * B + m2) * B + m1) * nd + v] +=
                    w[n1] * w[n2] * w[n3] * P(N, sub[0], m1, n1) * P(N, sub[1], m2, n2) *
                    P(N, sub[2], m3, n3) * fine[((n3 * B + n2) * B + n1) * nd + v] / w[m1] / w[m2] / w[m3] /
                    27.0;
}

static int ipow3(int e) {
  int r = 1;
  while (e-- > 0) r *= 3;
  return r;
}

/* Applies `levels` single-level prolongations to `count` doubles, digits of
 * idx[0..dim) taken most significant first (amrRoutines.cpph:75-100 and
 * :322-349).  The result lands in out. */
static void prolong_levels(int N, int nd, int dim, int levels, const int* idx, const double* in, double* out) {
  const int B = N + 1;
  const int count = (dim == 2 ? B * B : B * B * B) * nd;
  double* a = (double*)malloc(sizeof(double) * 2 * count);
  double* b = a + count;
  int prev[3] = {idx[0], idx[1], dim == 3 ? idx[2] : 0};
  const double* src = in;
  double* dst = a;
  for (int l = 1; l <= levels; l++) {
    const int sig = ipow3(levels - l);
    int sub[3];
    for (int k = 0; k < dim; k++) {
      const int cur = prev[k] % sig;
      sub[k] = (prev[k] - cur) / sig;
      prev[k] = cur;
    }
    memset(dst, 0, sizeof(double) * count);
    if (dim == 2)
      face_prolong_1(N, nd, dst, src, sub);
    else
      vol_prolong_1(N, nd, dst, src, sub);
    src = dst;
    dst = (dst == a) ? b : a;
  }
  memcpy(out, src, sizeof(double) * count);
  free(a);
}

/* Applies `levels` single-level restrictions, digits least significant first
 * (amrRoutines.cpph:203-246 and :440-481); the result is added into out. */
static void restrict_levels(int N, int nd, int dim, int levels, const int* idx, const double* in, double* out) {
  const int B = N + 1;
  const int count = (dim == 2 ? B * B : B * B * B) * nd;
  double* a = (double*)malloc(sizeof(double) * 2 * count);
  double* b = a + count;
  int cur[3] = {idx[0], idx[1], dim == 3 ? idx[2] : 0};
  const double* src = in;
  double* dst = a;
  for (int l = 1; l <= levels; l++) {
    int sub[3];
    for (int k = 0; k < dim; k++) {
      sub[k] = cur[k] % 3;
      cur[k] = (cur[k] - sub[k]) / 3;
    }
    memset(dst, 0, sizeof(double) * count);
    if (dim == 2)
      face_restrict_1(N, nd, dst, src, sub);
    else
      vol_restrict_1(N, nd, dst, src, sub);
    src = dst;
    dst = (dst == a) ? b : a;
  }
  for (int i = 0; i < count; i++) out[i] += src[i];
  free(a);
}

/* faceUnknownsProlongation, amrRoutines.cpph:45-144 */
void orc_face_unknowns_prolongation(int N, int nvar, int npar, double* lQhbndFine, double* lFhbndFine,
                                    const double* lQhbndCoarse, const double* lFhbndCoarse, int coarseLevel,
                                    int fineLevel, const int subfaceIndex[2]) {
  const int levels = fineLevel - coarseLevel;
  prolong_levels(N, nvar + npar, 2, levels, subfaceIndex, lQhbndCoarse, lQhbndFine);
  prolong_levels(N, nvar, 2, levels, subfaceIndex, lFhbndCoarse, lFhbndFine);
}

/* faceUnknownsRestriction, amrRoutines.cpph:182-258 */
void orc_face_unknowns_restriction(int N, int nvar, int npar, double* lQhbndCoarse, double* lFhbndCoarse,
                                   const double* lQhbndFine, const double* lFhbndFine, int coarseLevel,
                                   int fineLevel, const int subfaceIndex[2]) {
  const int levels = fineLevel - coarseLevel;
  restrict_levels(N, nvar + npar, 2, levels, subfaceIndex, lQhbndFine, lQhbndCoarse);
  restrict_levels(N, nvar, 2, levels, subfaceIndex, lFhbndFine, lFhbndCoarse);
}

/* volumeUnknownsProlongation, amrRoutines.cpph:289-374 */
void orc_volume_unknowns_prolongation(int N, int ndata, double* luhFine, const double* luhCoarse, int coarseLevel,
                                      int fineLevel, const int subcellIndex[3]) {
  prolong_levels(N, ndata, 3, fineLevel - coarseLevel, subcellIndex, luhCoarse, luhFine);
}

/* volumeUnknownsRestriction, amrRoutines.cpph:410-493 */
void orc_volume_unknowns_restriction(int N, int ndata, double* luhCoarse, const double* luhFine, int coarseLevel,
                                     int fineLevel, const int subcellIndex[3]) {
  restrict_levels(N, ndata, 3, fineLevel - coarseLevel, subcellIndex, luhFine, luhCoarse);
}
