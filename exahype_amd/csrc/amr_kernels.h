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

// Per-order tables: P[sub][coarse][fine] = phi_coarse((sub + xi_fine)/3), w = GL weights;
// Op[0][sub][o][i], Op[1][sub][o][i]: the one-level operators of level_op
// (prolongation, restriction), evaluated once on the host.
struct AmrTables {
  double P[3][MAXB][MAXB];
  double w[MAXB];
  double Op[2][3][MAXB][MAXB];
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

// One axis of a one-level transfer: pencil owners (B^(DIM-1) pencils times nd
// values per node), the axis operator read from the tables at addresses the
// whole workgroup shares (scalar loads), each pencil's B values in registers.
// The same products and sums as axis_pass.
template <int DIM, int B, int AXIS>
__device__ inline void axis_pass_op(const double* in, double* out, const double* __restrict__ Op, int nd) {
  constexpr int S = AXIS == 0 ? 1 : (AXIS == 1 ? B : B * B);  // node stride along the axis
  constexpr int NP = DIM == 3 ? B * B : B;                     // pencils
  for (int item = threadIdx.x; item < NP * nd; item += blockDim.x) {
    const int p = item / nd, v = item - p * nd;
    // the pencil's first node from the other axes' coordinates
    const int base = AXIS == 0 ? p * B : ((AXIS == 1 && DIM == 3) ? (p / B) * B * B + p % B : p);
    double q[B];
#pragma unroll
    for (int i = 0; i < B; ++i) q[i] = in[(base + i * S) * nd + v];
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < B; ++i) sum += Op[o * MAXB + i] * q[i];
      out[(base + o * S) * nd + v] = sum;
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

// Register-staged transfers: a block of at most THREADS * amr_regs values is
// loaded into registers (thread t owns values t, t + THREADS, ...), so the
// next transfer's block can be in flight while the current one is contracted:
// room for 8 values per node, at most 24 registers.
template <int DIM, int B>
constexpr int amr_regs() {
  constexpr int n = DIM == 3 ? B * B * B : B * B;
  constexpr int r = (n * 8 + THREADS - 1) / THREADS;
  return r > 24 ? 24 : r;
}

// The (group, transfer) after (g, it) in a workgroup's sequence -- groups
// blockIdx.x, blockIdx.x + gridDim.x, ... in order, each group's transfers
// in list order; false at the end.
__device__ inline bool next_transfer(const int* groups, int ngroups, int g, int it, int& g2, int& it2) {
  g2 = g;
  it2 = it + 1;
  while (it2 >= groups[g2 + 1]) {
    g2 += gridDim.x;
    if (g2 >= ngroups) return false;
    it2 = groups[g2];
  }
  return true;
}

// groups[g]..groups[g+1]-1 are the transfers of group g; workgroup b takes
// groups b, b + gridDim.x, ...  Prolongation overwrites the fine block (one
// transfer per group); restriction adds every transfer of the group into
// their common coarse block, in list order: coarse += R(child_1), then
// += R(child_2), ... (accumulated in registers, written once per group).
// Blocks are [node][nd] (the reference layout).  When every transfer of the
// launch spans one level (one_level), blocks of at most THREADS * amr_regs
// values go through registers with the next transfer's block loaded during
// the current one's contraction, and the operators come from the tables;
// otherwise (or with larger blocks) the operators of several levels are
// multiplied in LDS and blocks are processed in chunks of vc values per node.
// The order B is a template argument (constant index arithmetic).
template <int DIM, bool PROLONG, int B>
__global__ void __launch_bounds__(THREADS) k_amr_transfer(const AmrTables* __restrict__ tab,
                                                          const Transfer* __restrict__ tr, const int* __restrict__ groups,
                                                          double* __restrict__ fine, double* __restrict__ coarse,
                                                          int nd, int ngroups, int one_level) {
  __shared__ double A[3 * MAXB * MAXB];
  __shared__ double tmp[MAXB * MAXB];
  extern __shared__ double sbuf[];  // 2 * n * amr_chunk(B, DIM, nd), set by the launch
  constexpr int n = DIM == 3 ? B * B * B : B * B;
  const int vcmax = amr_chunk(B, DIM, nd);
  double* buf0 = sbuf;
  double* buf1 = sbuf + n * vcmax;
  constexpr int RMAX = amr_regs<DIM, B>();
  const int tid = threadIdx.x;
  const int nv = n * nd;
  // one-level transfers only: operators straight from the tables
  if (one_level && nv <= THREADS * RMAX) {
    int g = blockIdx.x;
    if (g >= ngroups) return;
    int it = groups[g];
    while (it >= groups[g + 1]) {  // empty groups
      g += gridDim.x;
      if (g >= ngroups) return;
      it = groups[g];
    }
    double r[RMAX], acc[RMAX];
    Transfer t = tr[it];
    {
      const double* src = PROLONG ? coarse + t.coarse : fine + t.fine;
#pragma unroll
      for (int k = 0; k < RMAX; ++k)
        if (tid + k * THREADS < nv) r[k] = src[tid + k * THREADS];
    }
    for (;;) {
      int g2, it2;
      const bool more = next_transfer(groups, ngroups, g, it, g2, it2);
      const bool first = it == groups[g];
      const bool last = !more || g2 != g;
      double* dst = PROLONG ? fine + t.fine : coarse + t.coarse;
      if (!PROLONG && first) {
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
          if (tid + k * THREADS < nv) acc[k] = dst[tid + k * THREADS];
      }
#pragma unroll
      for (int k = 0; k < RMAX; ++k)
        if (tid + k * THREADS < nv) buf0[tid + k * THREADS] = r[k];
      __syncthreads();
      Transfer t2 = t;
      if (more) {  // the next block, in flight during the contraction
        t2 = tr[it2];
        const double* src = PROLONG ? coarse + t2.coarse : fine + t2.fine;
#pragma unroll
        for (int k = 0; k < RMAX; ++k)
          if (tid + k * THREADS < nv) r[k] = src[tid + k * THREADS];
      }
      const double* Ot = &tab->Op[PROLONG ? 0 : 1][0][0][0];
      axis_pass_op<DIM, B, 0>(buf0, buf1, Ot + (t.index[0] % 3) * MAXB * MAXB, nd);
      __syncthreads();
      axis_pass_op<DIM, B, 1>(buf1, buf0, Ot + (t.index[1] % 3) * MAXB * MAXB, nd);
      __syncthreads();
      const double* res = buf0;
      if (DIM == 3) {
        axis_pass_op<DIM, B, 2>(buf0, buf1, Ot + (t.index[2] % 3) * MAXB * MAXB, nd);
        __syncthreads();
        res = buf1;
      }
#pragma unroll
      for (int k = 0; k < RMAX; ++k) {
        const int e = tid + k * THREADS;
        if (e < nv) {
          if (PROLONG) {
            dst[e] = res[e];
          } else {
            acc[k] += res[e];
            if (last) dst[e] = acc[k];
          }
        }
      }
      if (!more) return;
      __syncthreads();  // buf0 / buf1 are refilled by the next transfer
      g = g2;
      it = it2;
      t = t2;
    }
  }
  for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
    for (int it = groups[g]; it < groups[g + 1]; ++it) {
      const Transfer t = tr[it];
      build_axis_ops(*tab, PROLONG, t, DIM, B, A, tmp);
      const double* src = PROLONG ? coarse + t.coarse : fine + t.fine;
      double* dst = PROLONG ? fine + t.fine : coarse + t.coarse;
      for (int v0 = 0; v0 < nd; v0 += vcmax) {
        const int vc = nd - v0 < vcmax ? nd - v0 : vcmax;
        if (vc == nd) {  // whole blocks: one contiguous stream
          for (int e = tid; e < n * nd; e += blockDim.x) buf0[e] = src[e];
        } else {
          for (int node = tid; node < n; node += blockDim.x)
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
          for (int e = tid; e < n * nd; e += blockDim.x) {
            if (PROLONG)
              dst[e] = res[e];
            else
              dst[e] += res[e];
          }
        } else {
          for (int node = tid; node < n; node += blockDim.x)
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
}

}  // namespace amr
}  // namespace exa
