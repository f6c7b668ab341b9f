// A-posteriori subcell limiter on gfx950: DG <-> FV subcell projections and
// the discrete maximum principle (DMP).
//
// The projection matrices are built on the host with the reference's
// algorithm (kernels/LimiterProjectionMatrices.cpp:39-223): subcell averages
// of the Lagrange basis by Gauss-Legendre quadrature (uh2lim), basis values at
// the Gauss-Lobatto nodes (uh2lob), and the constrained least-squares inverse
// (lim2uh) through the same Gauss-Jordan elimination, whose pivot search keeps
// the signed value of a new pivot candidate (:181).  The same sequence of
// operations gives the same matrices bit for bit.
//
// Device kernels contract one axis at a time (sum factorisation): one
// workgroup per cell, one variable at a time, the partially contracted
// tensors of that variable in LDS ((B^2 L + B L^2 + L^3) doubles, 99 KB at
// N = 9).
#pragma once
#include <cmath>
#include <cstring>

#include "common.h"

namespace exa {
namespace lim {

constexpr int V = 5;
constexpr int MAXB = 10, MAXL = 19;

struct LimTables {
  double u2l[MAXB][MAXL];   // uh2lim[k][i]: average of phi_k over subcell i
  double u2lob[MAXB][MAXB]; // uh2lob[j][i]: phi_j at Gauss-Lobatto node i
  double l2u[MAXL][MAXB];   // lim2uh[i][j]
};

inline void base_func_1d(double* phi, double xi, int B, const double* nodes) {
  for (int i = 0; i < B; i++) phi[i] = 1;
  for (int m = 0; m < B; m++)
    for (int j = 0; j < B; j++) {
      if (j == m) continue;
      phi[m] = phi[m] * (xi - nodes[j]) / (nodes[m] - nodes[j]);
    }
}

// Gauss-Jordan inverse with the reference's pivot bookkeeping.
inline bool gauss_jordan_inverse(int n, const double* a, double* ia) {
  double c[(MAXB + 1) * 2 * (MAXB + 1)];
  auto C = [&](int i, int j) -> double& { return c[i * 2 * n + j]; };
  for (int i = 0; i < n; i++) {
    int j = 0;
    for (; j < n; j++) C(i, j) = a[j * n + i];
    for (; j < 2 * n; j++) C(i, j) = 0.;
    C(i, i + n) = 1.;
  }
  for (int i = 0; i < n; i++) {
    int ml = i;
    double mlV = std::fabs(C(i, i));
    for (int j = i + 1; j < n; j++)
      if (std::fabs(C(j, i)) > mlV) {
        ml = j;
        mlV = C(j, i);
      }
    for (int k = 0; k < 2 * n; k++) {
      const double t = C(ml, k);
      C(ml, k) = C(i, k);
      C(i, k) = t;
    }
    if (C(i, i) == 0) return false;
    const double piv = 1. / C(i, i);
    for (int k = 0; k < 2 * n; k++) C(i, k) *= piv;
    for (int j = i + 1; j < n; j++) {
      const double t = C(j, i);
      for (int k = 0; k < 2 * n; k++) C(j, k) -= t * C(i, k);
    }
  }
  for (int i = n - 1; i >= 0; i--)
    for (int j = i - 1; j >= 0; j--) {
      const double t = C(j, i);
      for (int k = 0; k < 2 * n; k++) C(j, k) -= t * C(i, k);
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) ia[j * n + i] = C(i, j + n);
  return true;
}

// gl_nodes/gl_weights/gll_nodes: the order-N rows of the quadrature tables.
inline bool build_tables(int N, const double* gl_nodes, const double* gl_weights, const double* gll_nodes,
                         LimTables& t) {
  const int B = N + 1, L = 2 * N + 1, M = B + 1;
  std::memset(&t, 0, sizeof(t));
  double phi[MAXB];
  const double dxi = 1. / L;
  for (int i = 0; i < L; i++) {
    const double xLeft = i * dxi;
    for (int j = 0; j < B; j++) {
      const double xi = xLeft + dxi * gl_nodes[j];
      base_func_1d(phi, xi, B, gl_nodes);
      for (int k = 0; k < B; k++) t.u2l[k][i] += gl_weights[j] * phi[k];
    }
  }
  for (int i = 0; i < B; i++) {
    base_func_1d(phi, gll_nodes[i], B, gl_nodes);
    for (int j = 0; j < B; j++) t.u2lob[j][i] = phi[j];
  }
  double lsqm[(MAXB + 1) * (MAXB + 1)], ilsqm[(MAXB + 1) * (MAXB + 1)], rhs[MAXL * (MAXB + 1)];
  for (int i = 0; i < B; i++) {
    for (int j = 0; j < B; j++) {
      lsqm[i * M + j] = 0.;
      for (int k = 0; k < L; k++) lsqm[i * M + j] += 2 * t.u2l[i][k] * t.u2l[j][k];
    }
    lsqm[i * M + B] = gl_weights[i];
  }
  for (int i = 0; i < B; i++) lsqm[B * M + i] = -gl_weights[i];
  lsqm[B * M + B] = 0.;
  if (!gauss_jordan_inverse(M, lsqm, ilsqm)) return false;
  for (int i = 0; i < L; i++) {
    for (int j = 0; j < B; j++) rhs[i * M + j] = 2 * t.u2l[j][i];
    rhs[i * M + B] = dxi;
  }
  for (int i = 0; i < L; i++)
    for (int j = 0; j < B; j++) {
      t.l2u[i][j] = 0.;
      for (int k = 0; k < M; k++) t.l2u[i][j] += ilsqm[k * M + j] * rhs[i * M + k];
    }
  return true;
}

// Contract axis `axis` of one variable's [n2][n1][n0] tensor (n0 fastest)
// with the matrix M[a_in][a_out]: in has extent nin along the axis, out has
// extent nout.  Threads stride over the output elements.
template <typename MAT>
__device__ inline void contract_axis(const double* in, double* out, int n0, int n1, int n2, int axis, int nin,
                                     int nout, MAT mat) {
  const int o0 = axis == 0 ? nout : n0, o1 = axis == 1 ? nout : n1, o2 = axis == 2 ? nout : n2;
  const int total = o0 * o1 * o2;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i0 = e % o0, i1 = (e / o0) % o1, i2 = e / (o0 * o1);
    double acc = 0.0;
    for (int k = 0; k < nin; ++k) {
      const int j0 = axis == 0 ? k : i0, j1 = axis == 1 ? k : i1, j2 = axis == 2 ? k : i2;
      const int oo = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
      acc += in[(j2 * n1 + j1) * n0 + j0] * mat(k, oo);
    }
    out[(i2 * o1 + i1) * o0 + i0] = acc;
  }
}

// luh [cell][v][B^3] -> lim [cell][v][S^3] interior (S = L + 2 ghost);
// projectOnFVLimiterSpace, kernels/limiter/generic/c/Limiter.cpp:23-66.
// One cell per workgroup: cell block u ([v][B^3]) -> patch block p ([v][S^3]).
__device__ inline void project_on_fv_block(const double* __restrict__ cellu, double* __restrict__ patch,
                                           const LimTables* tab, int B, int ghost, double* sm) {
  const int L = 2 * B - 1, S = L + 2 * ghost;
  double* a = sm;              // [B][B][L]
  double* b = sm + B * B * L;  // [B][L][L]
  double* c = b + B * L * L;   // [L][L][L]
  auto m = [&](int k, int o) { return tab->u2l[k][o]; };
  for (int v = 0; v < V; ++v) {
    const double* u = cellu + (size_t)v * B * B * B;
    contract_axis(u, a, B, B, B, 0, B, L, m);
    __syncthreads();
    contract_axis(a, b, L, B, B, 1, B, L, m);
    __syncthreads();
    contract_axis(b, c, L, L, B, 2, B, L, m);
    __syncthreads();
    double* o = patch + (size_t)v * S * S * S;
    for (int e = threadIdx.x; e < L * L * L; e += blockDim.x) {
      const int x = e % L, y = (e / L) % L, z = e / (L * L);
      o[((z + ghost) * S + y + ghost) * S + x + ghost] = c[e];
    }
    __syncthreads();
  }
}

// The two projections with the order a compile-time constant: all variables
// at once, pencil owners per axis with the operator entries read from the
// tables at constant offsets (scalar loads), the first axis reading the
// source block from memory and the last writing the result to memory; two
// barriers per block instead of four per variable.  The same products and
// sums in the same order as project_on_fv_block / project_on_dg_block.
template <int B>
constexpr int proj_t_lds_doubles() {
  return V * (B * B * (2 * B - 1) + B * (2 * B - 1) * (2 * B - 1));
}

template <int B>
__device__ inline void project_on_fv_block_t(const double* __restrict__ cellu, double* __restrict__ patch,
                                             const LimTables* __restrict__ tab, int ghost, double* sm) {
  constexpr int L = 2 * B - 1;
  const int S = L + 2 * ghost;
  double* a = sm;                  // [v][z][y][L]
  double* b = sm + V * B * B * L;  // [v][z][L][L]
  for (int it = threadIdx.x; it < V * B * B; it += blockDim.x) {  // x: pencils (v, z, y) of the cell
    double q[B];
#pragma unroll
    for (int k = 0; k < B; ++k) q[k] = cellu[it * B + k];
#pragma unroll
    for (int o = 0; o < L; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += q[k] * tab->u2l[k][o];
      a[it * L + o] = acc;
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < V * B * L; it += blockDim.x) {  // y: pencils (v, z, x)
    const int vz = it / L, x = it - vz * L;
    double q[B];
#pragma unroll
    for (int k = 0; k < B; ++k) q[k] = a[(vz * B + k) * L + x];
#pragma unroll
    for (int o = 0; o < L; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += q[k] * tab->u2l[k][o];
      b[(vz * L + o) * L + x] = acc;
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < V * L * L; it += blockDim.x) {  // z: pencils (v, y, x), into the patch
    const int v = it / (L * L), yx = it - v * L * L, y = yx / L, x = yx - y * L;
    double q[B];
#pragma unroll
    for (int k = 0; k < B; ++k) q[k] = b[((v * B + k) * L + y) * L + x];
    double* o_ = patch + (size_t)v * S * S * S + (size_t)(y + ghost) * S + x + ghost;
#pragma unroll
    for (int o = 0; o < L; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += q[k] * tab->u2l[k][o];
      o_[(size_t)(o + ghost) * S * S] = acc;
    }
  }
}

template <int B>
__device__ inline void project_on_dg_block_t(const double* __restrict__ patch, double* __restrict__ cellu,
                                             const LimTables* __restrict__ tab, int ghost, double* sm) {
  constexpr int L = 2 * B - 1;
  const int S = L + 2 * ghost;
  double* a = sm;                  // [v][z][y][B] over the subcells' z, y
  double* b = sm + V * L * L * B;  // [v][z][B][B]
  for (int it = threadIdx.x; it < V * L * L; it += blockDim.x) {  // x: pencils (v, z, y) of the patch interior
    const int v = it / (L * L), zy = it - v * L * L, z = zy / L, y = zy - z * L;
    const double* in = patch + (size_t)v * S * S * S + ((size_t)(z + ghost) * S + y + ghost) * S + ghost;
    double q[L];
#pragma unroll
    for (int k = 0; k < L; ++k) q[k] = in[k];
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < L; ++k) acc += q[k] * tab->l2u[k][o];
      a[it * B + o] = acc;
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < V * L * B; it += blockDim.x) {  // y: pencils (v, z, x)
    const int vz = it / B, x = it - vz * B;
    double q[L];
#pragma unroll
    for (int k = 0; k < L; ++k) q[k] = a[(vz * L + k) * B + x];
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < L; ++k) acc += q[k] * tab->l2u[k][o];
      b[(vz * B + o) * B + x] = acc;
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < V * B * B; it += blockDim.x) {  // z: pencils (v, y, x), into the cell
    const int v = it / (B * B), yx = it - v * B * B;
    double q[L];
#pragma unroll
    for (int k = 0; k < L; ++k) q[k] = b[(v * L + k) * B * B + yx];
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < L; ++k) acc += q[k] * tab->l2u[k][o];
      cellu[(size_t)v * B * B * B + o * B * B + yx] = acc;
    }
  }
}

// Block-per-cell projections with the order fixed (B <= 9: the two
// all-variable buffers must fit the LDS).
template <int B>
__global__ __launch_bounds__(256) void k_project_on_fv_t(const double* __restrict__ luh, double* __restrict__ lim,
                                                         const LimTables* __restrict__ tab, int ghost,
                                                         const int* __restrict__ cells) {
  extern __shared__ double sm[];
  constexpr int L = 2 * B - 1;
  const int S = L + 2 * ghost;
  const int cell = cells ? cells[blockIdx.x] : blockIdx.x;
  project_on_fv_block_t<B>(luh + (size_t)cell * V * B * B * B, lim + (size_t)blockIdx.x * V * S * S * S, tab, ghost,
                           sm);
}
template <int B>
__global__ __launch_bounds__(256) void k_project_on_dg_t(const double* __restrict__ lim, double* __restrict__ luh,
                                                         const LimTables* __restrict__ tab, int ghost,
                                                         const int* __restrict__ cells) {
  extern __shared__ double sm[];
  constexpr int L = 2 * B - 1;
  const int S = L + 2 * ghost;
  const int cell = cells ? cells[blockIdx.x] : blockIdx.x;
  project_on_dg_block_t<B>(lim + (size_t)blockIdx.x * V * S * S * S, luh + (size_t)cell * V * B * B * B, tab, ghost,
                           sm);
}

static __global__ void k_project_on_fv(const double* __restrict__ luh, double* __restrict__ lim, const LimTables* tab,
                                int B, int ghost, const int* __restrict__ cells) {
  extern __shared__ double sm[];
  const int L = 2 * B - 1, S = L + 2 * ghost;
  const int cell = cells ? cells[blockIdx.x] : blockIdx.x;
  project_on_fv_block(luh + (size_t)cell * V * B * B * B, lim + (size_t)blockIdx.x * V * S * S * S, tab, B, ghost,
                      sm);
}

// lim interior -> luh; projectOnDGSpace, kernels/limiter/generic/c/Limiter.cpp:69-112.
__device__ inline void project_on_dg_block(const double* __restrict__ patch, double* __restrict__ cellu,
                                           const LimTables* tab, int B, int ghost, double* sm) {
  const int L = 2 * B - 1, S = L + 2 * ghost;
  double* c = sm;              // [L][L][L]
  double* a = sm + L * L * L;  // [L][L][B]
  double* b = a + L * L * B;   // [L][B][B]
  auto m = [&](int k, int o) { return tab->l2u[k][o]; };
  for (int v = 0; v < V; ++v) {
    const double* in = patch + (size_t)v * S * S * S;
    for (int e = threadIdx.x; e < L * L * L; e += blockDim.x) {
      const int x = e % L, y = (e / L) % L, z = e / (L * L);
      c[e] = in[((z + ghost) * S + y + ghost) * S + x + ghost];
    }
    __syncthreads();
    contract_axis(c, a, L, L, L, 0, L, B, m);
    __syncthreads();
    contract_axis(a, b, B, L, L, 1, L, B, m);
    __syncthreads();
    contract_axis(b, cellu + (size_t)v * B * B * B, B, B, L, 2, L, B, m);
    __syncthreads();
  }
}

static __global__ void k_project_on_dg(const double* __restrict__ lim, double* __restrict__ luh, const LimTables* tab,
                                int B, int ghost, const int* __restrict__ cells) {
  extern __shared__ double sm[];
  const int L = 2 * B - 1, S = L + 2 * ghost;
  const int cell = cells ? cells[blockIdx.x] : blockIdx.x;
  project_on_dg_block(lim + (size_t)blockIdx.x * V * S * S * S, luh + (size_t)cell * V * B * B * B, tab, B, ghost,
                      sm);
}

// Cell min/max of the observables (the first nobs variables) over the GL
// nodes, the Gauss-Lobatto nodes and the subcell averages
// (findCellLocalMinAndMax, kernels/limiter/generic/c/Limiter.cpp:210-368).
// out: [cell][2][nobs] (min then max).
static __global__ void k_cell_min_max(const double* __restrict__ luh, double* __restrict__ out, const LimTables* tab, int B,
                               int nobs, int ncells) {
  extern __shared__ double sm[];
  const int L = 2 * B - 1, B3 = B * B * B;
  const int cell = blockIdx.x;
  double* a = sm;                // subcell chain: [B][B][L], [B][L][L], [L][L][L]
  double* b = a + B * B * L;
  double* c = b + B * L * L;
  double* g1 = c + L * L * L;    // Lobatto chain: [B][B][B] x 2
  double* g2 = g1 + B3;
  double* red = g2 + B3;         // 2 * (blockDim / 64)
  auto ml = [&](int k, int o) { return tab->u2l[k][o]; };
  auto mg = [&](int k, int o) { return tab->u2lob[k][o]; };
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  for (int v = 0; v < nobs; ++v) {
    const double* u = luh + ((size_t)cell * V + v) * B3;
    contract_axis(u, a, B, B, B, 0, B, L, ml);
    contract_axis(u, g1, B, B, B, 0, B, B, mg);
    __syncthreads();
    contract_axis(a, b, L, B, B, 1, B, L, ml);
    contract_axis(g1, g2, B, B, B, 1, B, B, mg);
    __syncthreads();
    contract_axis(b, c, L, L, B, 2, B, L, ml);
    contract_axis(g2, g1, B, B, B, 2, B, B, mg);
    __syncthreads();
    double mn = 1.0e300, mx = -1.0e300;
    for (int e = threadIdx.x; e < B3; e += blockDim.x) {
      const double p = u[e], q = g1[e];
      mn = fmin(mn, fmin(p, q));
      mx = fmax(mx, fmax(p, q));
    }
    for (int e = threadIdx.x; e < L * L * L; e += blockDim.x) {
      mn = fmin(mn, c[e]);
      mx = fmax(mx, c[e]);
    }
    mn = wave_min(mn);
    mx = -wave_min(-mx);
    if (lane == 0) {
      red[wid] = mn;
      red[nw + wid] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < nw; ++w) {
        mn = fmin(mn, red[w]);
        mx = fmax(mx, red[nw + w]);
      }
      out[((size_t)cell * 2 + 0) * nobs + v] = mn;
      out[((size_t)cell * 2 + 1) * nobs + v] = mx;
    }
    __syncthreads();
  }
}

// k_cell_min_max with the order as a template argument and one wavefront per
// (cell, observable): lanes own the x rows, then the y columns, then the z
// pencils of the three axis contractions, the last of which feeds the
// extrema without an LDS store; the tables are read through the scalar
// cache.  The same products accumulated in the same order as k_cell_min_max,
// the same out layout.  only_if: when not null, nothing is done unless *only_if != 0.
template <int B>
__global__ void __launch_bounds__(64) k_cell_min_max_w(const double* __restrict__ luh, double* __restrict__ out,
                                                       const LimTables* __restrict__ tab, int nobs,
                                                       const int* __restrict__ only_if) {
  constexpr int L = 2 * B - 1, B2 = B * B, B3 = B2 * B;
  __shared__ double a[B2 * L], b[B * L * L], g1[B3], g2[B3];
  if (only_if && !*only_if) return;
  const long long item = blockIdx.x;
  const long long cell = item / nobs;
  const int v = (int)(item % nobs);
  const double* u = luh + ((size_t)cell * V + v) * B3;
  const int lane = threadIdx.x;
  double mn = 1.0e300, mx = -1.0e300;
  // x: rows (z, y); the GL values themselves enter the extrema here
  for (int r = lane; r < B2; r += 64) {
    double x[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      x[k] = u[r * B + k];
      mn = fmin(mn, x[k]);
      mx = fmax(mx, x[k]);
    }
#pragma unroll
    for (int o = 0; o < L; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += x[k] * tab->u2l[k][o];
      a[r * L + o] = acc;
    }
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += x[k] * tab->u2lob[k][o];
      g1[r * B + o] = acc;
    }
  }
  __syncthreads();
  // y: columns (z, x) of a [z][y][L] into b [z][L][L], of g1 into g2
  for (int c = lane; c < B * L; c += 64) {
    const int z = c / L, i = c % L;
    double x[B];
#pragma unroll
    for (int k = 0; k < B; ++k) x[k] = a[(z * B + k) * L + i];
#pragma unroll
    for (int o = 0; o < L; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += x[k] * tab->u2l[k][o];
      b[(z * L + o) * L + i] = acc;
    }
  }
  for (int c = lane; c < B2; c += 64) {
    const int z = c / B, i = c % B;
    double x[B];
#pragma unroll
    for (int k = 0; k < B; ++k) x[k] = g1[(z * B + k) * B + i];
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += x[k] * tab->u2lob[k][o];
      g2[(z * B + o) * B + i] = acc;
    }
  }
  __syncthreads();
  // z: pencils (y, x); subcell averages and Lobatto values go to the extrema
  for (int c = lane; c < L * L; c += 64) {
    double x[B];
#pragma unroll
    for (int k = 0; k < B; ++k) x[k] = b[k * L * L + c];
#pragma unroll
    for (int o = 0; o < L; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += x[k] * tab->u2l[k][o];
      mn = fmin(mn, acc);
      mx = fmax(mx, acc);
    }
  }
  for (int c = lane; c < B2; c += 64) {
    double x[B];
#pragma unroll
    for (int k = 0; k < B; ++k) x[k] = g2[k * B2 + c];
#pragma unroll
    for (int o = 0; o < B; ++o) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) acc += x[k] * tab->u2lob[k][o];
      mn = fmin(mn, acc);
      mx = fmax(mx, acc);
    }
  }
  mn = wave_min(mn);
  mx = -wave_min(-mx);
  if (lane == 0) {
    out[((size_t)cell * 2 + 0) * nobs + v] = mn;
    out[((size_t)cell * 2 + 1) * nobs + v] = mx;
  }
}

// DMP of every cell against the min/max of itself and its six face
// neighbours (discreteMaximumPrincipleAndMinAndMaxSearch,
// kernels/limiter/generic/c/Limiter.cpp:114-165; the boundary min/max of a
// cell are the minima/maxima of its neighbours' cell values merged over the
// faces, E/solvers/LimitingADERDGSolver.cpp:993-1026).  ok[cell] = 1 if
// admissible.  minmax: [cell][2][nobs]; nbr: [6][ncells] neighbour ids.
static __global__ void k_dmp(const double* __restrict__ minmax, const double* __restrict__ bminmax,
                      int* __restrict__ ok, int nobs, int ncells, double relax, double scaling) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncells) return;
  int good = 1;
  for (int v = 0; v < nobs; ++v) {
    const double bmin = bminmax[((size_t)cell * 2 + 0) * nobs + v];
    const double bmax = bminmax[((size_t)cell * 2 + 1) * nobs + v];
    double sd = (bmax - bmin) * scaling;
    sd = sd < relax ? relax : sd;
    const double lmin = minmax[((size_t)cell * 2 + 0) * nobs + v];
    const double lmax = minmax[((size_t)cell * 2 + 1) * nobs + v];
    if (lmin < bmin - sd || lmax > bmax + sd) good = 0;
  }
  ok[cell] = good;
}

}  // namespace lim
}  // namespace exa
