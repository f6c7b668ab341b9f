// AMR level transfers between a coarse cell (or face) and one descendant
// 1..10 levels finer, in the reference's layouts [z][y][x][nData] (volumes) and
// [a][b][nData] (faces) -- kernels/aderdg/generic/c/3d/amrRoutines.cpph.
//
// The reference applies one full tensor-product operator per level.  Here the
// per-level 1D operators of each axis are first multiplied into one B x B
// matrix per axis (B^2 work per level), then the data goes through a
// sum-factorised pass per axis: DIM*B^(DIM+1) multiply-adds per variable instead
// of levels*B^(2*DIM).  Same linear map, different rounding (FP64, ~1e-15).
//
// One workgroup owns one output block.  Restrictions into the same coarse block
// form one group processed by a single workgroup in list order, so the
// accumulation coarse += R(child_1) + R(child_2) + ... happens in the caller's
// order without atomics.
#pragma once
#include <hip/hip_runtime.h>

namespace exa {
namespace amr {

constexpr int MAXB = 10;
constexpr int THREADS = 256;

// Per-order tables: P[sub][coarse][fine] = phi_coarse((sub + xi_fine)/3), w = GL weights.
struct AmrTables {
  double P[3][MAXB][MAXB];
  double w[MAXB];
};

struct Transfer {
  long long fine;    // offset (doubles) of the fine block
  long long coarse;  // offset (doubles) of the coarse block
  int levels;        // fine level - coarse level, 1..10
  int index[3];      // sub-cell / sub-face index per axis, in [0, 3^levels)
};

// T_l[o][i] for one level and one axis: prolongation out = fine, in = coarse;
// restriction out = coarse, in = fine (L2 projection, weight ratio / 3).
__device__ inline double level_op(const AmrTables& t, bool prolong, int sub, int o, int i) {
  return prolong ? t.P[sub][i][o] : t.w[i] * t.P[sub][o][i] / (t.w[o] * 3.0);
}

// A = T_L ... T_1 per axis; digits most significant first for prolongation,
// least significant first for restriction.
__device__ inline void build_axis_ops(const AmrTables& t, bool prolong, const Transfer& tr, int dim, int B,
                                      double* A, double* tmp) {
  const int tid = threadIdx.x;
  for (int a = 0; a < dim; ++a) {
    double* Aa = A + a * MAXB * MAXB;
    int pw = 1;
    for (int l = 1; l < tr.levels; ++l) pw *= 3;  // 3^(levels-1)
    int msd_div = pw, lsd_div = 1;
    for (int l = 1; l <= tr.levels; ++l) {
      const int div = prolong ? msd_div : lsd_div;
      const int sub = (tr.index[a] / div) % 3;
      msd_div /= 3;
      lsd_div *= 3;
      __syncthreads();
      if (tid < B * B) {
        const int o = tid / B, i = tid % B;
        double v;
        if (l == 1) {
          v = level_op(t, prolong, sub, o, i);
        } else {
          v = 0.0;
          for (int k = 0; k < B; ++k) v += level_op(t, prolong, sub, o, k) * Aa[k * MAXB + i];
        }
        tmp[o * MAXB + i] = v;
      }
      __syncthreads();
      if (tid < B * B) Aa[(tid / B) * MAXB + tid % B] = tmp[(tid / B) * MAXB + tid % B];
    }
  }
  __syncthreads();
}

// out[...o on axis...] = sum_i A[o][i] * in[...i on axis...] over a B^DIM block;
// axis 0 is the fastest index.
template <int DIM>
__device__ inline void axis_pass(const double* in, double* out, const double* A, int axis, int B) {
  int n = B * B;
  if (DIM == 3) n *= B;
  int stride = 1;
  for (int a = 0; a < axis; ++a) stride *= B;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int o = (e / stride) % B;
    const int base = e - o * stride;
    double s = 0.0;
    for (int i = 0; i < B; ++i) s += A[o * MAXB + i] * in[base + i * stride];
    out[e] = s;
  }
}

// groups[g]..groups[g+1]-1 are the transfers of workgroup g.  Prolongation
// overwrites the fine block (one transfer per group); restriction adds every
// transfer of the group into their common coarse block.
template <int DIM, bool PROLONG>
__global__ void __launch_bounds__(THREADS) k_amr_transfer(const AmrTables* __restrict__ tab,
                                                          const Transfer* __restrict__ tr, const int* __restrict__ groups,
                                                          double* __restrict__ fine, double* __restrict__ coarse, int B,
                                                          int nd) {
  __shared__ double A[3 * MAXB * MAXB];
  __shared__ double tmp[MAXB * MAXB];
  __shared__ double buf0[MAXB * MAXB * MAXB];
  __shared__ double buf1[MAXB * MAXB * MAXB];
  const int g = blockIdx.x;
  const int n = DIM == 3 ? B * B * B : B * B;
  for (int it = groups[g]; it < groups[g + 1]; ++it) {
    const Transfer t = tr[it];
    build_axis_ops(*tab, PROLONG, t, DIM, B, A, tmp);
    const double* src = PROLONG ? coarse + t.coarse : fine + t.fine;
    double* dst = PROLONG ? fine + t.fine : coarse + t.coarse;
    for (int v = 0; v < nd; ++v) {
      for (int e = threadIdx.x; e < n; e += blockDim.x) buf0[e] = src[(long long)e * nd + v];
      __syncthreads();
      axis_pass<DIM>(buf0, buf1, A, 0, B);
      __syncthreads();
      axis_pass<DIM>(buf1, buf0, A + MAXB * MAXB, 1, B);
      __syncthreads();
      const double* res = buf0;
      if (DIM == 3) {
        axis_pass<DIM>(buf0, buf1, A + 2 * MAXB * MAXB, 2, B);
        __syncthreads();
        res = buf1;
      }
      for (int e = threadIdx.x; e < n; e += blockDim.x) {
        double* o = dst + (long long)e * nd + v;
        if (PROLONG)
          *o = res[e];
        else
          *o += res[e];
      }
      __syncthreads();
    }
  }
}

}  // namespace amr
}  // namespace exa
