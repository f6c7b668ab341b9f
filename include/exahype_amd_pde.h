/*
 * exahype_amd_pde.h -- compile a project's own PDE into the ADER-DG kernels.
 *
 * The reference resolves the user solver's pointwise functions statically:
 * the generated Abstract<Solver> passes the user class as the SolverType
 * template argument of the generic kernels (TT/GenericADERDGSolverHeader.template:71-111,
 * E/solvers/UserSolverInterface.h:42-176).  Here a project writes the same
 * functions as static __device__ members of a struct, includes this header
 * in a HIP translation unit built for gfx950 (hipcc --offload-arch=gfx950,
 * with -I of this directory and of exahype_amd/csrc, the kernel templates),
 * and calls exa::register_pde<MyPDE>("name", &id) once after exa_init; the
 * returned id goes into exa_aderdg_config.pde.  The library then runs the
 * generic kernel family (FP64 Picard predictor with the flux / source /
 * non-conservative-product arms the PDE enables, Rusanov Riemann solver,
 * surface integral, update and CFL time step) on that PDE in every per-cell
 * entry point and on the device grid, including its initial conditions
 * (exa_grid_set_initial_conditions with a null exa_ic) and domain
 * boundaries (exa_grid_set_boundary / exa_aderdg_boundary_conditions with
 * EXA_BC_USER).
 *
 * The user PDE concept (reference names and argument meanings):
 *
 *   struct MyPDE {
 *     static constexpr int NumberOfVariables = 5;   // nVar (see below)
 *     static constexpr bool useFlux = true, useSource = false, useNCP = false;
 *     // F[d][v]: flux of Q in direction d        (UserSolverInterface.h:175)
 *     __device__ static void flux(const double* Q, double** F);
 *     // all eigenvalues of the flux Jacobian along d (GenericADERDGSolverHeader.template:90)
 *     __device__ static void eigenvalues(const double* Q, int d, double* lambda);
 *     // S(Q)                                       (UserSolverInterface.h:112)
 *     __device__ static void algebraicSource(const double* Q, double* S);
 *     // B(Q) gradQ with gradQ[d * nVar + v]         (UserSolverInterface.h:147)
 *     __device__ static void nonConservativeProduct(const double* Q, const double* gradQ, double* BgradQ);
 *     // initial state at x (ADERDGSolver.h:1183; w the quadrature weight)
 *     __device__ static void adjustPointSolution(const double* x, double w, double t, double dt, double* Q);
 *     // outside state / normal flux of a domain face node
 *     // (GenericADERDGSolverHeader.template:111)
 *     __device__ static void boundaryValues(const double* x, double t, double dt, int faceIndex,
 *                                           int normalNonZero, const double* fluxIn, const double* stateIn,
 *                                           double* fluxOut, double* stateOut);
 *   };
 *
 * algebraicSource and nonConservativeProduct are only called when their
 * flag is set (declare them anyway; empty bodies do).  The number of
 * variables is a template parameter of the concept, but the kernels of this
 * build hold NumberOfVariables = 5 per node (exa::V): another count is
 * rejected at compile time.  The PDE runs in FP64 at orders 1..7.
 */
#ifndef EXAHYPE_AMD_PDE_H
#define EXAHYPE_AMD_PDE_H

#include "exahype_amd.h"
#include "aderdg_pde_kernels.h"

namespace exa {

// The kernels' PDE interface (pde_generic.h) from the user's reference-style
// functions: the wave speed is max_v |lambda_v| accumulated as
// stableTimeStepSize.cpph:43-48 and the Rusanov solver do (std::max(|l|, m)
// from 0, a NaN propagating).
template <class U>
struct UserPde {
  static_assert(U::NumberOfVariables == V, "the kernels of this build hold 5 variables per node");
  static constexpr bool FLUX = U::useFlux, SRC = U::useSource, NCP = U::useNCP;
  __device__ static inline void flux(const double* Q, double F[3][V], const PdeParams&) {
    double* rows[3] = {F[0], F[1], F[2]};
    U::flux(Q, rows);
  }
  __device__ static inline double max_abs_eigenvalue(const double* Q, int d, const PdeParams&) {
    double lam[V];
    U::eigenvalues(Q, d, lam);
    double m = 0.0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double a = fabs(lam[v]);
      m = (a < m) ? m : a;
    }
    return m;
  }
  __device__ static inline void max_abs_eigenvalues3(const double* Q, const PdeParams& p, double m[3]) {
    for (int d = 0; d < 3; ++d) m[d] = max_abs_eigenvalue(Q, d, p);
  }
  __device__ static inline void source(const double* Q, double S[V], const PdeParams&) {
    if (SRC)
      U::algebraicSource(Q, S);
    else
      for (int v = 0; v < V; ++v) S[v] = 0.0;
  }
  __device__ static inline void ncp(const double* Q, const double G[3][V], double out[V], const PdeParams&) {
    if (NCP)
      U::nonConservativeProduct(Q, &G[0][0], out);
    else
      for (int v = 0; v < V; ++v) out[v] = 0.0;
  }
};

// adjustPointSolution at the GL nodes of every cell of a regular grid
// (solutionAdjustment.cpph:17-52; cell (ix, iy, iz) centred at
// off + (i + 1/2) h, x fastest), at t = 0, dt = 0.
template <int N, class U>
__global__ __launch_bounds__(Ord<N>::NT) void k_user_initial_conditions(double* __restrict__ luh, int nx, int ny,
                                                                        double ox, double oy, double oz,
                                                                        double dx0, double dx1, double dx2) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  const DGTables& T = c_tab[N];
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  if (s >= B3) return;
  const int z = s / B2, y = (s / B) % B, x = s % B;
  const int ix = (int)(cell % nx), iy = (int)((cell / nx) % ny), iz = (int)(cell / ((size_t)nx * ny));
  const double c0 = ox + (ix + 0.5) * dx0, c1 = oy + (iy + 0.5) * dx1, c2 = oz + (iz + 0.5) * dx2;
  double X[3];
  X[2] = c2 + dx2 * (T.xi[z] - 0.5);
  X[1] = c1 + dx1 * (T.xi[y] - 0.5);
  X[0] = c0 + dx0 * (T.xi[x] - 0.5);
  const double w = (T.w[z] * T.w[y]) * T.w[x];
  double Q[V];
  U::adjustPointSolution(X, w, 0.0, 0.0, Q);
#pragma unroll
  for (int v = 0; v < V; ++v) luh[cell * V * B3 + v * B3 + s] = Q[v];
}

// Face node (a, b) of face f of the cell at centre c, size h: x[normal] on
// the face, x at the GL nodes along the slower (a) and faster (b)
// tangential axes (boundaryConditions.cpph:55-81).
template <int N>
__device__ inline void user_face_point(int f, int a, int b, const double c[3], const double h[3], double x[3]) {
  const DGTables& T = c_tab[N];
  const int d = f >> 1, d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
  x[d] = c[d] + (-0.5 + (f & 1)) * h[d];
  x[d2] = c[d2] + h[d2] * (T.xi[a] - 0.5);
  x[d1] = c[d1] + h[d1] * (T.xi[b] - 0.5);
}

// The domain faces of a grid (bfaces[i] = cell * 6 + face): boundaryValues
// at every face node, then the Riemann problem with the outside state on
// the outer side (applyBoundaryConditions, E/solvers/ADERDGSolver.cpp:2854-2931).
// oq, of: [i][v][B^2] scratch for the outside state and flux.
template <int N, class U>
__global__ __launch_bounds__(256) void k_user_boundary_faces(double* __restrict__ qf, double* __restrict__ ff,
                                                             const int* __restrict__ bfaces, int nb,
                                                             double* __restrict__ oq, double* __restrict__ of,
                                                             const double* __restrict__ dt_cur,
                                                             const double* __restrict__ time, double off0,
                                                             double off1, double off2, double dx0, double dx1,
                                                             double dx2, int nx, int ny, double gamma, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  constexpr int NPL = (B2 + 63) / 64;
  const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= nb) return;
  const int code = bfaces[i];
  const long long cell = code / 6;
  const int f = code % 6, d = f >> 1;
  const int ix = (int)(cell % nx), iy = (int)((cell / nx) % ny), iz = (int)(cell / ((long long)nx * ny));
  const double h[3] = {dx0, dx1, dx2};
  const double c[3] = {off0 + (ix + 0.5) * dx0, off1 + (iy + 0.5) * dx1, off2 + (iz + 0.5) * dx2};
  double* QIn = qf + ((size_t)cell * 6 + f) * V * B2;
  double* FIn = ff + ((size_t)cell * 6 + f) * V * B2;
  double* QO = oq + (size_t)i * V * B2;
  double* FO = of + (size_t)i * V * B2;
  const double t = *time, dt = *dt_cur;
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
      double x[3], qi[V], fi[V], qo[V], fo[V];
      user_face_point<N>(f, node / B, node % B, c, h, x);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        qi[v] = QIn[v * B2 + node];
        fi[v] = FIn[v * B2 + node];
      }
      U::boundaryValues(x, t, dt, f, d, fi, qi, fo, qo);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        QO[v * B2 + node] = qo[v];
        FO[v * B2 + node] = fo[v];
      }
    }
  }
  const PdeParams P{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}};
  // each lane reads back only the nodes it wrote
  if ((f & 1) == 0)
    riemann_face_pde<N, UserPde<U>>(QO, QIn, FO, FIn, d, P, lane);
  else
    riemann_face_pde<N, UserPde<U>>(QIn, QO, FIn, FO, d, P, lane);
}

// boundaryConditions of one face in the reference layout (node-major,
// nd = 5 values per node): the per-face entry point.
template <int N, class U>
__global__ __launch_bounds__(64) void k_user_boundary_ref(double* __restrict__ FOut, double* __restrict__ QOut,
                                                          const double* __restrict__ FIn,
                                                          const double* __restrict__ QIn, double c0, double c1,
                                                          double c2, double s0, double s1, double s2, double t,
                                                          double dt, int faceIndex) {
  constexpr int B = Ord<N>::B, B2 = Ord<N>::B2;
  const double c[3] = {c0, c1, c2}, h[3] = {s0, s1, s2};
  for (int node = threadIdx.x; node < B2; node += blockDim.x) {
    double x[3];
    user_face_point<N>(faceIndex, node / B, node % B, c, h, x);
    U::boundaryValues(x, t, dt, faceIndex, faceIndex >> 1, FIn + (size_t)node * V, QIn + (size_t)node * V,
                      FOut + (size_t)node * V, QOut + (size_t)node * V);
  }
}

namespace user_detail {

// Per translation unit: the DG tables of this unit's copy of c_tab and the
// LDS limits of its predictors.
template <class U>
int init_unit(const void* tables) {
  if (hipMemcpyToSymbol(HIP_SYMBOL(c_tab), tables, sizeof(DGTables) * (EXA_MAX_ORDER + 1)) != hipSuccess) return -1;
  int r = 0;
  auto set = [&](const void* k, int bytes) {
    if (r == 0 && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) r = -1;
  };
  set(reinterpret_cast<const void*>(&k_predictor_pde<1, UserPde<U>, false>), PdeOrd<1>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<1, UserPde<U>, true>), PdeOrd<1>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<2, UserPde<U>, false>), PdeOrd<2>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<2, UserPde<U>, true>), PdeOrd<2>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<3, UserPde<U>, false>), PdeOrd<3>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<3, UserPde<U>, true>), PdeOrd<3>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<4, UserPde<U>, false>), PdeOrd<4>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<4, UserPde<U>, true>), PdeOrd<4>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<5, UserPde<U>, false>), PdeOrd<5>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<5, UserPde<U>, true>), PdeOrd<5>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<6, UserPde<U>, false>), PdeOrd<6>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<6, UserPde<U>, true>), PdeOrd<6>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<7, UserPde<U>, false>), PdeOrd<7>::LDS_DOUBLES * 8);
  set(reinterpret_cast<const void*>(&k_predictor_pde<7, UserPde<U>, true>), PdeOrd<7>::LDS_DOUBLES * 8);
  return r;
}

template <class U, int N>
void fill_order(exa_pde_kernels& k) {
  using P = UserPde<U>;
  constexpr int FULL = UPD_SURFACE | UPD_UPDATE | UPD_DT;
  k.predictor[N][0] = reinterpret_cast<const void*>(&k_predictor_pde<N, P, false>);
  k.predictor[N][1] = reinterpret_cast<const void*>(&k_predictor_pde<N, P, true>);
  k.riemann[N] = reinterpret_cast<const void*>(&k_riemann_pde<N, P>);
  k.riemann_list[N] = reinterpret_cast<const void*>(&k_riemann_list_pde<N, P>);
  k.update_dt[N] = reinterpret_cast<const void*>(&k_update<N, UPD_DT, P>);
  k.update_full[N] = reinterpret_cast<const void*>(&k_update<N, FULL, P>);
  k.initial_conditions[N] = reinterpret_cast<const void*>(&k_user_initial_conditions<N, U>);
  k.boundary_faces[N] = reinterpret_cast<const void*>(&k_user_boundary_faces<N, U>);
  k.boundary_ref[N] = reinterpret_cast<const void*>(&k_user_boundary_ref<N, U>);
  k.predictor_lds_bytes[N] = PdeOrd<N>::LDS_DOUBLES * (int)sizeof(double);
}

}  // namespace user_detail

// Instantiates the kernel family for U (orders 1..7) in the calling
// translation unit and registers it under name; *id receives the PDE id for
// exa_aderdg_config.pde.  Returns an EXA_* status.
template <class U>
int register_pde(const char* name, int* id) {
  exa_pde_kernels k{};
  k.version = EXA_PDE_KERNELS_VERSION;
  k.max_order = 7;
  k.init = &user_detail::init_unit<U>;
  user_detail::fill_order<U, 1>(k);
  user_detail::fill_order<U, 2>(k);
  user_detail::fill_order<U, 3>(k);
  user_detail::fill_order<U, 4>(k);
  user_detail::fill_order<U, 5>(k);
  user_detail::fill_order<U, 6>(k);
  user_detail::fill_order<U, 7>(k);
  return exa_pde_register(name, &k, id);
}

}  // namespace exa

#endif  // EXAHYPE_AMD_PDE_H
