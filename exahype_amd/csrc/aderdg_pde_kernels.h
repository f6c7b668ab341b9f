// ADER-DG kernels of the generic PDE arms (flux / algebraic source /
// non-conservative product), FP64, orders 1..7.
//
// Same device layout and summation orders as the Euler kernels
// (aderdg_kernels.h).  The predictor keeps one thread per spatial node with
// the current iterate and the next one in registers and passes one time
// slice at a time through LDS: the flux slice for the Kxi contraction and,
// for a non-conservative product, the iterate slice for the dudx gradients.
// spaceTimePredictorNonlinear.cpph:43-577 (all three arms),
// riemannSolverNonlinear.cpph:38-120 (useNCP), volumeIntegralNonlinear.cpph:26-125.
#pragma once
#include "aderdg_kernels.h"
#include "pde_generic.h"

namespace exa {

constexpr int PDE_MAX_ORDER = 7;

template <int N>
struct PdeOrd {
  // iterate slice [v][node] + flux slice [d][v][node] + reduction scratch;
  // after the sweeps the same words hold sQh [v][node] and sFh [d][v][node]
  static constexpr int LDS_DOUBLES = 7 * V * Ord<N>::B3 + 16;  // + the NCP gradients
};

template <int N, class PDE, bool DEBUG>
__global__ __launch_bounds__(Ord<N>::NT) void k_predictor_pde(
    const double* __restrict__ luh, double* __restrict__ lduh, double* __restrict__ qf, double* __restrict__ ff,
    int* __restrict__ iters, const double* __restrict__ dtp, double dx0, double dx1, double dx2, double gamma,
    unsigned long long* __restrict__ sweep_count, PredictorDebug dbg, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  constexpr int MaxIterations = 2 * (N + 1);
  extern __shared__ double sm[];
  double* sQ = sm;                  // iterate slice [v][node]
  double* sF = sm + V * B3;         // flux slice [d][v][node]
  double* red = sm + 4 * V * B3;    // residual sum scratch
  double* sG = sm + 4 * V * B3 + 16; // gradients of the slice [d][v][node] (NCP)
  double* sQh = sm;                 // after the sweeps: time averages
  double* sFh = sm + V * B3;
  const DGTables& T = c_tab[N];
  const PdeParams P{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}};
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  const bool act = s < B3;
  const int z = act ? s / B2 : 0, y = act ? (s / B) % B : 0, x = act ? s % B : 0;
  const double dt = *dtp;
  const double w3 = (T.w[z] * T.w[y]) * T.w[x];
  const double iweight = 1.0 / w3;

  double u0[V], q[B][V];
#pragma unroll
  for (int v = 0; v < V; ++v) u0[v] = act ? luh[cell * V * B3 + v * B3 + s] : 1.0;
  // trivial initial guess (:68-78)
#pragma unroll
  for (int t = 0; t < B; ++t)
#pragma unroll
    for (int v = 0; v < V; ++v) q[t][v] = u0[v];

  double fh[3][V], sh[V];
  int it = 0;
  for (int iter = 0; iter < MaxIterations; ++iter) {
    ++it;
    double qn[B][V];
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) qn[t][v] = 0.0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      sh[v] = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) fh[d][v] = 0.0;
    }
#pragma unroll
    for (int t = 0; t < B; ++t) {
      __syncthreads();  // the previous slice is consumed
      if (act) {
        if (PDE::NCP) {
#pragma unroll
          for (int v = 0; v < V; ++v) sQ[v * B3 + s] = q[t][v];
        }
        double F[3][V];
        if (PDE::FLUX) {
          PDE::flux(q[t], F, P);
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) {
              sF[(d * V + v) * B3 + s] = F[d][v];
              fh[d][v] += F[d][v] * T.w[t];  // time average of the sweep (:326-397)
            }
        }
        if (DEBUG) {
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < V; ++v) dbg.lFi[((t * B3 + s) * 4 + d) * V + v] = PDE::FLUX ? F[d][v] : 0.0;
        }
      }
      __syncthreads();
      if (PDE::FLUX) {
        // pencil owners contract one (direction, variable, pencil) of the
        // flux slice with Kxi in place: sum_n F_n Kxi[n][l] at node l
        for (int k = s; k < 3 * V * B2; k += O::NT) {
          const int d = k / (V * B2), v = (k / B2) % V, ab = k % B2, a = ab / B, b = ab % B;
          const int base = d == 0 ? a * B2 + b * B : (d == 1 ? a * B2 + b : a * B + b);
          const int step = d == 0 ? 1 : (d == 1 ? B : B2);
          double* pf = sF + (d * V + v) * B3 + base;
          double pv[B];
#pragma unroll
          for (int n = 0; n < B; ++n) pv[n] = pf[n * step];
#pragma unroll
          for (int l = 0; l < B; ++l) {
            double c = 0.0;
#pragma unroll
            for (int n = 0; n < B; ++n) c += pv[n] * T.Kxi[n][l];
            pf[l * step] = c;
          }
        }
      }
      if (PDE::NCP) {
        // pencil owners: the B nodal derivatives of one (direction, variable,
        // pencil) from one read of its values
        for (int k = s; k < 3 * V * B2; k += O::NT) {
          const int d = k / (V * B2), v = (k / B2) % V, ab = k % B2, a = ab / B, b = ab % B;
          const int base = d == 0 ? a * B2 + b * B : (d == 1 ? a * B2 + b : a * B + b);
          const int step = d == 0 ? 1 : (d == 1 ? B : B2);
          const double idx = d == 0 ? 1.0 / dx0 : (d == 1 ? 1.0 / dx1 : 1.0 / dx2);
          double pq[B];
#pragma unroll
          for (int n = 0; n < B; ++n) pq[n] = sQ[v * B3 + base + n * step];
#pragma unroll
          for (int l = 0; l < B; ++l) {
            double g = 0.0;
#pragma unroll
            for (int n = 0; n < B; ++n) g += idx * pq[n] * T.dudx[l][n];
            sG[(d * V + v) * B3 + base + l * step] = g;
          }
        }
      }
      if (PDE::FLUX || PDE::NCP) __syncthreads();
      if (act) {
        // rhs0 = w^3 F0 luh (:118-133)
        double r[V];
        const double c0 = w3 * T.F0[t];
#pragma unroll
        for (int v = 0; v < V; ++v) r[v] = c0 * u0[v];
        if (PDE::FLUX) {  // rhs -= (w dt/dx) lFi Kxi per direction (:143-222)
          const double usx = ((T.w[t] * T.w[z]) * T.w[y]) * dt / dx0;
          const double usy = ((T.w[t] * T.w[z]) * T.w[x]) * dt / dx1;
          const double usz = ((T.w[t] * T.w[y]) * T.w[x]) * dt / dx2;
#pragma unroll
          for (int v = 0; v < V; ++v) {
            r[v] -= usx * sF[(0 * V + v) * B3 + s];
            r[v] -= usy * sF[(1 * V + v) * B3 + s];
            r[v] -= usz * sF[(2 * V + v) * B3 + s];
          }
        }
        if (PDE::SRC || PDE::NCP) {
          // gradients of the iterate (:136-222), fused source (:227-252)
          double G[3][V];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            G[0][v] = G[1][v] = G[2][v] = 0.0;
            if (PDE::NCP) {
#pragma unroll
              for (int d = 0; d < 3; ++d) G[d][v] = sG[(d * V + v) * B3 + s];
            }
          }
          double S[V];
          fused_source<PDE>(q[t], G, S, P);
          const double us = (((T.w[t] * T.w[z]) * T.w[y]) * T.w[x]) * dt;
#pragma unroll
          for (int v = 0; v < V; ++v) {
            r[v] += us * S[v];
            sh[v] += S[v] * T.w[t];
          }
          if (DEBUG) {
#pragma unroll
            for (int v = 0; v < V; ++v) dbg.lFi[((t * B3 + s) * 4 + 3) * V + v] = S[v];
          }
        } else if (DEBUG) {
#pragma unroll
          for (int v = 0; v < V; ++v) dbg.lFi[((t * B3 + s) * 4 + 3) * V + v] = 0.0;
        }
        // this time node's share of the new iterate: += (w^-3 rhs) iK1 (:270-296)
#pragma unroll
        for (int l = 0; l < B; ++l)
#pragma unroll
          for (int v = 0; v < V; ++v) qn[l][v] += iweight * r[v] * T.iK1[l][t];
      }
    }
    // exit condition (:299-306)
    double res = 0.0;
    if (act) {
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const double d = q[t][v] - qn[t][v];
          res += d * d;
          q[t][v] = qn[t][v];
        }
    }
    res = block_sum(res, red);
    if (res < 1e-14) break;
  }

  // time averages (:326-397) into LDS, then volume integral and extrapolation
  __syncthreads();
  if (act) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double a = 0.0;
#pragma unroll
      for (int t = 0; t < B; ++t) a += q[t][v] * T.w[t];
      sQh[v * B3 + s] = a;
      if (DEBUG) {
        dbg.lQhi[s * V + v] = a;
#pragma unroll
        for (int t = 0; t < B; ++t) dbg.lQi[(s * B + t) * V + v] = q[t][v];
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v = 0; v < V; ++v) sFh[(d * V + v) * B3 + s] = PDE::FLUX ? fh[d][v] : 0.0;
    if (DEBUG) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        dbg.lFhi[0 * B3 * V + ((z * B + y) * B + x) * V + v] = PDE::FLUX ? fh[0][v] : 0.0;
        dbg.lFhi[1 * B3 * V + ((z * B + x) * B + y) * V + v] = PDE::FLUX ? fh[1][v] : 0.0;
        dbg.lFhi[2 * B3 * V + ((y * B + x) * B + z) * V + v] = PDE::FLUX ? fh[2][v] : 0.0;
        dbg.lFhi[3 * B3 * V + s * V + v] = (PDE::SRC || PDE::NCP) ? sh[v] : 0.0;
      }
    }
  }
  __syncthreads();
  if (act) {
    double out[V];
    if (PDE::FLUX) {
      volume_integral_node<N>(sFh, s, z, y, x, dx0, dx1, dx2, out);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] = 0.0;
    }
    if (PDE::SRC || PDE::NCP) {  // += w^3 lShi (volumeIntegralNonlinear.cpph:106-124)
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] += w3 * sh[v];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) lduh[cell * V * B3 + v * B3 + s] = out[v];
  }
  // without a flux sFh is zero and so is lFhbnd, as the reference writes it (:418)
  extrapolate_faces<N>(sQh, sFh, s, qf + cell * 6 * V * B2, ff + cell * 6 * V * B2);
  if (s == 0) {
    if (iters) iters[cell] = it;
    if (sweep_count) atomicAdd(sweep_count, (unsigned long long)it);
  }
}

// One face, one wavefront: Rusanov flux from the face-averaged wave speeds
// plus, for a non-conservative product, the fluctuation ncp(Qavg, gradQ)
// with gradQ zero except the normal row QR - QL: FL = F + ncp/2,
// FR = F - ncp/2 (riemannSolverNonlinear.cpph:83-116).
template <int N, class PDE>
__device__ inline void riemann_face_pde(const double* __restrict__ QL, const double* __restrict__ QR,
                                        double* __restrict__ FL, double* __restrict__ FR, int d,
                                        const PdeParams& P, int lane) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  constexpr int NPL = (B2 + 63) / 64;
  const DGTables& T = c_tab[N];
  double aL[V], aR[V];
#pragma unroll
  for (int v = 0; v < V; ++v) aL[v] = aR[v] = 0.0;
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
      const double w = T.w[node / B] * T.w[node % B];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        aL[v] += w * QL[v * B2 + node];
        aR[v] += w * QR[v * B2 + node];
      }
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v) {
    aL[v] = wave_sum(aL[v]);
    aR[v] = wave_sum(aR[v]);
  }
  const double sL = PDE::max_abs_eigenvalue(aL, d, P);
  const double sR = PDE::max_abs_eigenvalue(aR, d, P);
  const double smax = sL < sR ? sR : sL;
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
      double ncp[V];
      if (PDE::NCP) {
        double G[3][V], Qa[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int i = v * B2 + node;
          const double jump = QR[i] - QL[i];
          G[0][v] = d == 0 ? jump : 0.0;
          G[1][v] = d == 1 ? jump : 0.0;
          G[2][v] = d == 2 ? jump : 0.0;
          Qa[v] = 0.5 * (QR[i] + QL[i]);
        }
        PDE::ncp(Qa, G, ncp, P);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int i = v * B2 + node;
        const double F = 0.5 * (FR[i] + FL[i]) - 0.5 * smax * (QR[i] - QL[i]);
        if (PDE::NCP) {
          FR[i] = F - 0.5 * ncp[v];
          FL[i] = F + 0.5 * ncp[v];
        } else {
          FL[i] = F;
          FR[i] = F;
        }
      }
    }
  }
}

// Every face of a grid, one wavefront per face (left cell face 2d+1, right
// cell face 2d; E/solvers/ADERDGSolver.cpp:2744-2825).
template <int N, class PDE>
__global__ __launch_bounds__(256) void k_riemann_pde(const double* __restrict__ qf, double* __restrict__ ff,
                                                     const int* __restrict__ nbr, int ncells, double gamma,
                                                     long long wave0, long long nwaves, Ens ens) {
  constexpr int B2 = Ord<N>::B2;
  const long long local = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (local >= nwaves) return;
  const long long wave = wave0 + local;
  const int d = (int)(wave / ncells);
  const int cl = (int)(wave % ncells);
  const int cr = nbr[(size_t)d * ncells + cl];
  if (cr < 0) return;  // a domain face of a non-periodic axis
  const PdeParams P{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}};
  riemann_face_pde<N, PDE>(qf + ((size_t)cl * 6 + 2 * d + 1) * V * B2, qf + ((size_t)cr * 6 + 2 * d) * V * B2,
                           ff + ((size_t)cl * 6 + 2 * d + 1) * V * B2, ff + ((size_t)cr * 6 + 2 * d) * V * B2, d, P,
                           lane);
}

// A list of independent faces: face i reads QL = q[2i], QR = q[2i+1] and
// updates FL = f[2i], FR = f[2i+1], normal direction dirs[i].
template <int N, class PDE>
__global__ __launch_bounds__(256) void k_riemann_list_pde(const double* __restrict__ q, double* __restrict__ f,
                                                          const int* __restrict__ dirs, long long nfaces,
                                                          double gamma, Ens ens) {
  constexpr int B2 = Ord<N>::B2;
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (i >= nfaces) return;
  const PdeParams P{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}};
  riemann_face_pde<N, PDE>(q + (size_t)(2 * i) * V * B2, q + (size_t)(2 * i + 1) * V * B2,
                           f + (size_t)(2 * i) * V * B2, f + (size_t)(2 * i + 1) * V * B2, dirs[i], P,
                           threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// The linear kernel family (the isLinear branch of the generated glue),
// orders 1..9.
constexpr int LINEAR_MAX_ORDER = 9;

// ---------------------------------------------------------------------------
// Point sources (kernels/aderdg/generic/c/3d/pointSource.cpph:28-145).
// Multiplies and adds stay separately rounded here, as in the reference's
// host build: the source position in reference coordinates decides whether a
// cell holds the source.

// Lagrange basis polynomial of GL node i at xi from its definition,
// prod_{m != i} (xi - xi_m) / (xi_i - xi_m) (the reference evaluates
// generated literals of the same polynomials, kernels/DGBasisFunctions.cpp).
template <int N>
__device__ inline double lagrange(int i, double xi) {
#pragma clang fp contract(off)
  const DGTables& T = c_tab[N];
  double num = 1.0, den = 1.0;
#pragma unroll
  for (int m = 0; m <= N; ++m) {
    if (m == i) continue;
    num *= xi - T.xi[m];
    den *= T.xi[i] - T.xi[m];
  }
  return num / den;
}

// The user pointSource(x, t, dt, forceVector, x0) (E/solvers/
// UserSolverInterface.h:91-95): the Ricker wavelet force of PointSrc.
__device__ inline void user_point_source(const double*, double t, double, const PointSrc& ps, double f[5],
                                         double x0[3]) {
#pragma clang fp contract(off)
  const double pa = M_PI * ps.f0 * (t - ps.t0);
  const double a = pa * pa;
  const double r = (1.0 - 2.0 * a) * exp(-a);
  for (int v = 0; v < 5; ++v) f[v] = ps.amp[v] * r;
  for (int d = 0; d < 3; ++d) x0[d] = ps.x0[d];
}

// PSi[n][z][y][x][v] of one node (pointSource.cpph:58-125): level n = 0 at t,
// n >= 1 at the GL time node n-1 of [t, t+dt]; the force at the node,
// projected onto the node's basis function at the source position, zero
// when the source lies outside the cell.
template <int N>
__device__ inline void point_source_node(int n, int z, int y, int x, double t, double dt, const double c[3],
                                         const double h[3], const PointSrc& ps, double out[5]) {
#pragma clang fp contract(off)
  const DGTables& T = c_tab[N];
  const double tn = n == 0 ? t : dt * T.xi[n - 1] + t;
  const double X[3] = {c[0] + h[0] * (T.xi[x] - 0.5), c[1] + h[1] * (T.xi[y] - 0.5), c[2] + h[2] * (T.xi[z] - 0.5)};
  double f[5], x0[3], r[3];
  user_point_source(X, tn, dt, ps, f, x0);
  for (int d = 0; d < 3; ++d) r[d] = (x0[d] - (c[d] - h[d] / 2.0)) / h[d];
  if (r[0] >= 0.0 && r[1] >= 0.0 && r[2] >= 0.0 && r[0] <= 1.0 && r[1] <= 1.0 && r[2] <= 1.0) {
    const double b = lagrange<N>(z, r[2]) * lagrange<N>(y, r[1]) * lagrange<N>(x, r[0]);
    const double den = T.w[z] * h[2] * T.w[y] * h[1] * T.w[x] * h[0];
    for (int v = 0; v < 5; ++v) out[v] = b * f[v] / den;
  } else {
    for (int v = 0; v < 5; ++v) out[v] = 0.0;
  }
}

// pointSource of one cell: thread per node, PSi [B+1][z][y][x][v].
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_point_source(double t, double dt, double c0, double c1, double c2,
                                                             double h0, double h1, double h2, double* __restrict__ psi,
                                                             Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  const int s = threadIdx.x;
  if (s >= B3) return;
  const int z = s / B2, y = (s / B) % B, x = s % B;
  const double c[3] = {c0, c1, c2}, h[3] = {h0, h1, h2};
  for (int n = 0; n <= B; ++n) {
    double o[5];
    point_source_node<N>(n, z, y, x, t, dt, c, h, ens.ps, o);
    for (int v = 0; v < V; ++v) psi[(n * B3 + s) * V + v] = o[v];
  }
}

// Cauchy-Kowalevski predictor + linear volume integral
// (spaceTimePredictorLinear.cpph:33-443, volumeIntegralLinear.cpph:21-55).
// Thread s owns node s and walks the time derivatives lQi[0..B]: the level
// passes through LDS, pencil owners contract each pencil with dudx once per
// direction (LDS bandwidth binds this kernel: one read per value and
// direction instead of B), node owners read their gradients back, the product tensor
// T = B_d dQ/dx_d gives the next level -sum_d T_d, and the Taylor time
// averages dt^i/(i+1)! accumulate in registers.
template <int N, class PDE, bool DEBUG, bool PS = false>  // PS: the point-source arm is compiled in
__global__ __launch_bounds__(Ord<N>::NT) void k_predictor_linear(
    const double* __restrict__ luh, double* __restrict__ lduh, double* __restrict__ qf, double* __restrict__ ff,
    int* __restrict__ iters, const double* __restrict__ dtp, double dx0, double dx1, double dx2, double gamma,
    unsigned long long* __restrict__ sweep_count, PredictorDebug dbg, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  extern __shared__ double sm[];
  double* sQ = sm;           // one derivative level [v][node]
  double* sG = sm + V * B3;  // its gradients [d][v][node]
  double* sQh = sm;          // afterwards: time averages [v][node]
  double* sFh = sm + V * B3; // and the averaged products [d][v][node]
  const DGTables& T = c_tab[N];
  const PdeParams P{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}};
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  const bool act = s < B3;
  const int z = act ? s / B2 : 0, y = act ? (s / B) % B : 0, x = act ? s % B : 0;
  const double dt = *dtp;
  double q[V], qh[V], fh[3][V];
#pragma unroll
  for (int v = 0; v < V; ++v) q[v] = qh[v] = act ? luh[cell * V * B3 + v * B3 + s] : 1.0;
  // point source (:78-102): the force at t enters the first time derivative,
  // its time integral sum_n w_n PSi[n+1] the zeroth
  double ps0[V], psint[V];
  if (PS && ens.ps.mode && act) {
    double c[3], h[3] = {dx0, dx1, dx2};
    if (ens.ps.mode == 2) {
      const long long ix = (long long)(cell % ens.ps.nx), iy = (long long)((cell / ens.ps.nx) % ens.ps.ny),
                      iz = (long long)(cell / ((size_t)ens.ps.nx * ens.ps.ny));
      c[0] = __dadd_rn(ens.ps.off[0], __dmul_rn((double)ix + 0.5, dx0));
      c[1] = __dadd_rn(ens.ps.off[1], __dmul_rn((double)iy + 0.5, dx1));
      c[2] = __dadd_rn(ens.ps.off[2], __dmul_rn((double)iz + 0.5, dx2));
    }
    for (int n = 0; n <= B; ++n) {
      double o[V];
      if (ens.ps.mode == 1) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = ens.ps.psi[((cell * (B + 1) + n) * B3 + s) * V + v];
      } else {
        point_source_node<N>(n, z, y, x, dtp[2], dt, c, h, ens.ps, o);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if (n == 0) {
          ps0[v] = o[v];
          psint[v] = 0.0;
        } else {
          psint[v] = __dadd_rn(psint[v], __dmul_rn(T.w[n - 1], o[v]));
        }
      }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) q[v] = qh[v] = __dadd_rn(q[v], __dmul_rn(dt, psint[v]));
  }
  double fac = 0.5 * dt;  // dtavFac of the level (:236-248)
  for (int i = 0; i < B; ++i) {
    __syncthreads();
    if (act) {
#pragma unroll
      for (int v = 0; v < V; ++v) sQ[v * B3 + s] = q[v];
    }
    __syncthreads();
    // pencil owners: one (direction, variable, pencil) item reads its B
    // values once and writes the B nodal derivatives sum_n (q_n / dx_d) dudx[l][n]
    for (int k = s; k < 3 * V * B2; k += O::NT) {
      const int d = k / (V * B2), v = (k / B2) % V, ab = k % B2, a = ab / B, b = ab % B;
      const int base = d == 0 ? a * B2 + b * B : (d == 1 ? a * B2 + b : a * B + b);
      const int step = d == 0 ? 1 : (d == 1 ? B : B2);
      const double idx = d == 0 ? 1.0 / dx0 : (d == 1 ? 1.0 / dx1 : 1.0 / dx2);
      double pq[B];
#pragma unroll
      for (int n = 0; n < B; ++n) pq[n] = sQ[v * B3 + base + n * step];
#pragma unroll
      for (int l = 0; l < B; ++l) {
        double g = 0.0;
#pragma unroll
        for (int n = 0; n < B; ++n) g += idx * pq[n] * T.dudx[l][n];
        sG[(d * V + v) * B3 + base + l * step] = g;
      }
    }
    __syncthreads();
    if (act) {
      double G[3][V], Tn[3][V];
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) G[d][v] = sG[(d * V + v) * B3 + s];
      PDE::ncp_linear(q, G, Tn, P);
      if (DEBUG) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          dbg.lQi[(i * B3 + s) * V + v] = q[v];
#pragma unroll
          for (int d = 0; d < 3; ++d) dbg.lFi[((i * B3 + s) * 3 + d) * V + v] = Tn[d][v];
        }
      }
      // time averages: level 0 initialises, level i >= 1 adds dtavFac_i times it
#pragma unroll
      for (int v = 0; v < V; ++v) {
        if (i > 0) qh[v] += fac * q[v];
#pragma unroll
        for (int d = 0; d < 3; ++d) fh[d][v] = i == 0 ? Tn[d][v] : fh[d][v] + fac * Tn[d][v];
      }
      if (i > 0) fac *= dt / (i + 2);
      // next level: lQi[i+1] = 0 - T_x - T_y - T_z (:196-207), the first
      // level starting from the force at t when there is a point source
#pragma unroll
      for (int v = 0; v < V; ++v) q[v] = (((PS && i == 0 && ens.ps.mode ? ps0[v] : 0.0) - Tn[0][v]) - Tn[1][v]) - Tn[2][v];
    }
  }
  if (DEBUG && act) {
#pragma unroll
    for (int v = 0; v < V; ++v) dbg.lQi[(B * B3 + s) * V + v] = q[v];
  }
  // a third of the force's time integral in every direction's block (:287-302)
  if (PS && ens.ps.mode && act) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v = 0; v < V; ++v) fh[d][v] += psint[v] / 3.0;
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      sQh[v * B3 + s] = qh[v];
#pragma unroll
      for (int d = 0; d < 3; ++d) sFh[(d * V + v) * B3 + s] = fh[d][v];
    }
    if (DEBUG) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        dbg.lQhi[s * V + v] = qh[v];
#pragma unroll
        for (int d = 0; d < 3; ++d) dbg.lFhi[(d * B3 + s) * V + v] = fh[d][v];
        dbg.lFhi[(3 * B3 + s) * V + v] = 0.0;
      }
    }
    // volume integral: lduh = -w^3 sum_d lFh[d]
    const double w3 = (T.w[z] * T.w[y]) * T.w[x];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double o = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) o -= w3 * fh[d][v];
      lduh[cell * V * B3 + v * B3 + s] = o;
    }
  }
  __syncthreads();
  extrapolate_faces<N>(sQh, sFh, s, qf + cell * 6 * V * B2, ff + cell * 6 * V * B2);
  if (s == 0 && iters) iters[cell] = 0;  // no Picard iterations
  (void)sweep_count;
}

// Linear Riemann solver on one face, one wavefront: face averages, smax from
// the eigenvalues, Bn = coefficientMatrix(mean of the averages); per node
// FL = (Bn - smax I)^T (QR - QL)/2, FR = (Bn + smax I)^T (QR - QL)/2
// (riemannSolverLinear.cpph:24-140).
template <int N, class PDE>
__device__ inline void riemann_face_linear(const double* __restrict__ QL, const double* __restrict__ QR,
                                           double* __restrict__ FL, double* __restrict__ FR, int d,
                                           const PdeParams& P, int lane) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  constexpr int NPL = (B2 + 63) / 64;
  const DGTables& T = c_tab[N];
  double aL[V], aR[V];
#pragma unroll
  for (int v = 0; v < V; ++v) aL[v] = aR[v] = 0.0;
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
      const double w = T.w[node / B] * T.w[node % B];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        aL[v] += w * QL[v * B2 + node];
        aR[v] += w * QR[v * B2 + node];
      }
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v) {
    aL[v] = wave_sum(aL[v]);
    aR[v] = wave_sum(aR[v]);
  }
  const double sL = PDE::max_abs_eigenvalue(aL, d, P);
  const double sR = PDE::max_abs_eigenvalue(aR, d, P);
  const double smax = sL < sR ? sR : sL;
  double Qa[V], Bm[V][V], Bp[V][V];
#pragma unroll
  for (int v = 0; v < V; ++v) Qa[v] = 0.5 * (aL[v] + aR[v]);
  PDE::coefficient_matrix(Qa, d, Bm, P);
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int j = 0; j < V; ++j) Bp[i][j] = Bm[i][j];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    Bm[i][i] -= smax;
    Bp[i][i] += smax;
  }
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
      double Qd[V];
#pragma unroll
      for (int v = 0; v < V; ++v) Qd[v] = 0.5 * (QR[v * B2 + node] - QL[v * B2 + node]);
#pragma unroll
      for (int kk = 0; kk < V; ++kk) {
        double l = 0.0, r = 0.0;
#pragma unroll
        for (int ll = 0; ll < V; ++ll) {
          l += Bm[ll][kk] * Qd[ll];
          r += Bp[ll][kk] * Qd[ll];
        }
        FL[kk * B2 + node] = l;
        FR[kk * B2 + node] = r;
      }
    }
  }
}

template <int N, class PDE>
__global__ __launch_bounds__(256) void k_riemann_linear(const double* __restrict__ qf, double* __restrict__ ff,
                                                        const int* __restrict__ nbr, int ncells, double gamma,
                                                        long long wave0, long long nwaves, Ens ens) {
  constexpr int B2 = Ord<N>::B2;
  const long long local = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (local >= nwaves) return;
  const long long wave = wave0 + local;
  const int d = (int)(wave / ncells);
  const int cl = (int)(wave % ncells);
  const int cr = nbr[(size_t)d * ncells + cl];
  const PdeParams P{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}};
  riemann_face_linear<N, PDE>(qf + ((size_t)cl * 6 + 2 * d + 1) * V * B2, qf + ((size_t)cr * 6 + 2 * d) * V * B2,
                              ff + ((size_t)cl * 6 + 2 * d + 1) * V * B2, ff + ((size_t)cr * 6 + 2 * d) * V * B2, d,
                              P, lane);
}

// volumeIntegralLinear of the per-cell entry: lFhi [d][node][V] (reference
// layout, d < 3) -> lduh [v][node].
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_volume_integral_linear(const double* __restrict__ lFhi,
                                                                       double* __restrict__ lduh) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  const int s = threadIdx.x;
  if (s >= B3) return;
  const DGTables& T = c_tab[N];
  const int z = s / B2, y = (s / B) % B, x = s % B;
  const double w3 = (T.w[z] * T.w[y]) * T.w[x];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    double o = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) o -= w3 * lFhi[(d * B3 + s) * V + v];
    lduh[v * B3 + s] = o;
  }
}

}  // namespace exa
