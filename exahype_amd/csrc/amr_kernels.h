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
constexpr int THREADS = 256;  // >= B^2 = 100: build_axis_ops gives every (o, i) its own thread

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
  if (tr.levels == 1) {  // one level: the operators are table rows, all axes in one pass
    __syncthreads();     // the previous transfer of the group is done with A
    for (int e = tid; e < dim * B * B; e += blockDim.x) {
      const int a = e / (B * B), oi = e - a * B * B, o = oi / B, i = oi - o * B;
      A[a * MAXB * MAXB + o * MAXB + i] = level_op(t, prolong, tr.index[a] % 3, o, i);
    }
    __syncthreads();
    return;
  }
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

// out[...o on axis...] = sum_i A[o][i] * in[...i on axis...] over a B^DIM block
// of nodes (axis 0 fastest) holding vc values per node (value fastest).
template <int DIM, int B>
__device__ inline void axis_pass(const double* in, double* out, const double* A, int axis, int vc) {
  constexpr int n = DIM == 3 ? B * B * B : B * B;
  const int stride = axis == 0 ? 1 : (axis == 1 ? B : B * B);
  for (int node = threadIdx.x; node < n; node += blockDim.x) {
    const int o = (node / stride) % B;
    const int base = node - o * stride;
    for (int v = 0; v < vc; ++v) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < B; ++i) s += A[o * MAXB + i] * in[(base + i * stride) * vc + v];
      out[node * vc + v] = s;
    }
  }
}

// Values per node processed together: the two LDS buffers hold B^DIM * vc
// doubles each (AMR_LDS_DOUBLES per buffer).
constexpr int AMR_LDS_DOUBLES = 8192;
__host__ __device__ inline int amr_chunk(int B, int DIM, int nd) {
  const int n = DIM == 3 ? B * B * B : B * B;
  int vc = AMR_LDS_DOUBLES / n;
  if (vc > nd) vc = nd;
  return vc < 1 ? 1 : vc;
}

// groups[g]..groups[g+1]-1 are the transfers of workgroup g.  Prolongation
// overwrites the fine block (one transfer per group); restriction adds every
// transfer of the group into their common coarse block.  Blocks are
// [node][nd] (the reference layout); a chunk of vc values per node is loaded
// into LDS in one pass and contracted axis by axis for all its values.  The
// order B is a template argument (constant index arithmetic).
template <int DIM, bool PROLONG, int B>
__global__ void __launch_bounds__(THREADS) k_amr_transfer(const AmrTables* __restrict__ tab,
                                                          const Transfer* __restrict__ tr, const int* __restrict__ groups,
                                                          double* __restrict__ fine, double* __restrict__ coarse,
                                                          int nd) {
  __shared__ double A[3 * MAXB * MAXB];
  __shared__ double tmp[MAXB * MAXB];
  extern __shared__ double sbuf[];  // 2 * n * amr_chunk(B, DIM, nd), set by the launch
  const int g = blockIdx.x;
  constexpr int n = DIM == 3 ? B * B * B : B * B;
  const int vcmax = amr_chunk(B, DIM, nd);
  double* buf0 = sbuf;
  double* buf1 = sbuf + n * vcmax;
  for (int it = groups[g]; it < groups[g + 1]; ++it) {
    const Transfer t = tr[it];
    build_axis_ops(*tab, PROLONG, t, DIM, B, A, tmp);
    const double* src = PROLONG ? coarse + t.coarse : fine + t.fine;
    double* dst = PROLONG ? fine + t.fine : coarse + t.coarse;
    for (int v0 = 0; v0 < nd; v0 += vcmax) {
      const int vc = nd - v0 < vcmax ? nd - v0 : vcmax;
      if (vc == nd) {  // whole blocks: one contiguous stream
        for (int e = threadIdx.x; e < n * nd; e += blockDim.x) buf0[e] = src[e];
      } else {
        for (int node = threadIdx.x; node < n; node += blockDim.x)
          for (int v = 0; v < vc; ++v) buf0[node * vc + v] = src[(long long)node * nd + v0 + v];
      }
      __syncthreads();
      axis_pass<DIM, B>(buf0, buf1, A, 0, vc);
      __syncthreads();
      axis_pass<DIM, B>(buf1, buf0, A + MAXB * MAXB, 1, vc);
      __syncthreads();
      const double* res = buf0;
      if (DIM == 3) {
        axis_pass<DIM, B>(buf0, buf1, A + 2 * MAXB * MAXB, 2, vc);
        __syncthreads();
        res = buf1;
      }
      if (vc == nd) {
        for (int e = threadIdx.x; e < n * nd; e += blockDim.x) {
          if (PROLONG)
            dst[e] = res[e];
          else
            dst[e] += res[e];
        }
      } else {
        for (int node = threadIdx.x; node < n; node += blockDim.x)
          for (int v = 0; v < vc; ++v) {
            double* o = dst + (long long)node * nd + v0 + v;
            if (PROLONG)
              *o = res[node * vc + v];
            else
              *o += res[node * vc + v];
          }
      }
      __syncthreads();
    }
  }
}

}  // namespace amr
}  // namespace exa
