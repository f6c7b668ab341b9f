// Finite-volume kernels for gfx950: Godunov (2D/3D), MUSCL-Hancock (3D),
// Rusanov flux, FV CFL time step, ghost-layer filling of periodic patch grids.
//
// Device layout of a patch: structure of arrays over the patch including
// its ghost layers, [patch][v][cell], cell = (k*S + j)*S + i (3D) or j*S + i
// (2D), S = patch + 2*ghost.  Updates are written in gather form: one thread
// per interior cell evaluates the Rusanov fluxes of its 2*DIM faces and
// applies them in the reference's order (x-, x+, y-, y+, z-, z+), so each
// cell's accumulation matches the reference's face loops
// (kernels/finitevolumes/godunov/c/{2d,3d}/godunov.cpph,
// kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph).
#pragma once
#include "common.h"
#include "pde_euler.h"

namespace exa {
namespace fv {

constexpr int V = 5;

struct PatchGeom {
  int P, G, S, dim;
  __host__ __device__ int ncells_box() const { return dim == 3 ? S * S * S : S * S; }
  __host__ __device__ int stride(int d) const { return d == 0 ? 1 : (d == 1 ? S : S * S); }
};

// rusanov<useNCP=false,useFlux=true>, kernels/finitevolumes/riemannsolvers/c/rusanov.cpph:24-156
template <class PDE>
__device__ inline double rusanov(const double* qL, const double* qR, int d, double gamma, double* f) {
  const double sL = PDE::max_abs_eigenvalue(qL, d, gamma);
  const double sR = PDE::max_abs_eigenvalue(qR, d, gamma);
  double s_max = -1.0;
  s_max = sL < s_max ? s_max : sL;
  s_max = sR < s_max ? s_max : sR;
  double FL[V], FR[V];
  PDE::flux_dir(qL, d, FL, gamma);
  PDE::flux_dir(qR, d, FR, gamma);
#pragma unroll
  for (int i = 0; i < V; ++i) {
    f[i] = 0.5 * s_max * (qL[i] - qR[i]);
    f[i] += 0.5 * (FL[i] + FR[i]);
  }
  return s_max;
}

__device__ inline void load_q(const double* base, int nbox, int cell, double* q) {
#pragma unroll
  for (int v = 0; v < V; ++v) q[v] = base[v * nbox + cell];
}

__device__ inline void interior_coords(const PatchGeom& g, int c, int& i, int& j, int& k) {
  i = c % g.P + g.G;
  j = (c / g.P) % g.P + g.G;
  k = g.dim == 3 ? c / (g.P * g.P) + g.G : 0;
}

__device__ inline int box_index(const PatchGeom& g, int i, int j, int k) {
  return g.dim == 3 ? (k * g.S + j) * g.S + i : j * g.S + i;
}

// The reference's face loops also run over the two patch-boundary faces of
// every line, so the first ghost layer of luh_new receives one face update:
// -= at the lower ghost (face ghost|first), += at the upper ghost (last|ghost).
// Enumerates those cells: item -> (direction d, side, transverse position).
__device__ inline void ghost_layer_cell(const PatchGeom& g, int item, int& d, int& side, int& idx, int& inner) {
  const int trans = g.dim == 3 ? g.P * g.P : g.P;
  d = item / (2 * trans);
  side = (item / trans) % 2;
  const int t = item % trans;
  const int a = g.dim == 3 ? t / g.P + g.G : 0, b = t % g.P + g.G;
  int c[3];
  const int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
  c[o1] = b;
  c[o2] = a;
  c[d] = side ? g.G + g.P : g.G - 1;
  idx = box_index(g, c[0], c[1], c[2]);
  c[d] = side ? g.G + g.P - 1 : g.G;
  inner = box_index(g, c[0], c[1], c[2]);
}

// Godunov update of every patch: luh_new(interior) = luh + flux differences.
// Returns (per patch, and as a global minimum) the admissible dt of the face
// wave speeds: min CFL/DIM * h / s_max (godunov.cpph:33-143).
template <int DIM, class PDE>
__global__ __launch_bounds__(256) void k_godunov(const double* __restrict__ luh, double* __restrict__ luh_new,
                                                 double* __restrict__ dt_patch, unsigned long long* dt_min,
                                                 PatchGeom g, double dx0, double dx1, double dx2,
                                                 const double* __restrict__ dtp, double gamma, double cfl) {
  __shared__ double red[8];
  const double dt = *dtp;
  const int nbox = g.ncells_box();
  const double* u = luh + (size_t)blockIdx.x * V * nbox;
  double* un = luh_new + (size_t)blockIdx.x * V * nbox;
  const double h[3] = {dx0 / g.P, dx1 / g.P, dx2 / g.P};
  const int ninterior = DIM == 3 ? g.P * g.P * g.P : g.P * g.P;
  double dtm = 1.0e300;
  for (int c = threadIdx.x; c < ninterior; c += blockDim.x) {
    int i, j, k;
    interior_coords(g, c, i, j, k);
    const int idx = box_index(g, i, j, k);
    double q[V], acc[V], qn[V], f[V];
    load_q(u, nbox, idx, q);
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = q[v];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double fac = dt / h[d];
      load_q(u, nbox, idx - g.stride(d), qn);
      double s = rusanov<PDE>(qn, q, d, gamma, f);
      double cand = exa_dt_sane(cfl / DIM * h[d] / s);
      dtm = cand < dtm ? cand : dtm;
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] += fac * f[v];
      load_q(u, nbox, idx + g.stride(d), qn);
      s = rusanov<PDE>(q, qn, d, gamma, f);
      cand = exa_dt_sane(cfl / DIM * h[d] / s);
      dtm = cand < dtm ? cand : dtm;
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] -= fac * f[v];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) un[v * nbox + idx] = acc[v];
  }
  const int nghost = 2 * DIM * (DIM == 3 ? g.P * g.P : g.P);
  for (int it = threadIdx.x; it < nghost; it += blockDim.x) {
    int d, side, idx, inner;
    ghost_layer_cell(g, it, d, side, idx, inner);
    double ql[V], qr[V], f[V];
    load_q(u, nbox, side ? inner : idx, ql);
    load_q(u, nbox, side ? idx : inner, qr);
    rusanov<PDE>(ql, qr, d, gamma, f);
    const double fac = dt / h[d];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double o = un[v * nbox + idx];
      un[v * nbox + idx] = side ? o + fac * f[v] : o - fac * f[v];
    }
  }
  dtm = block_min(dtm, red);
  if (threadIdx.x == 0) {
    if (dt_patch) dt_patch[blockIdx.x] = dtm;
    if (dt_min) atomic_min_pos_double(dt_min, dtm);
  }
}

__device__ inline double minmod(double a, double b) {
  if (a * b < 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

// MUSCL-Hancock (kernels/finitevolumes/musclhancock/c/3d/musclhancock.cpph:
// 27-309), one workgroup per patch, one launch:
//   phase 1  every interior subcell evaluates its half-step evolution term
//            Qt = sum_d (F_d(w_L,d) - F_d(w_R,d)) / h_d from its minmod slopes
//            and boundary-extrapolated states (:62-213) into LDS;
//   phase 2  every interior subcell rebuilds the evolved states of the faces it
//            touches -- its own and its neighbours' w = q -+ h/2 slope, plus
//            dt/2 Qt where the owner is interior -- and applies the Rusanov
//            fluxes on (w_R[i], w_L[i+1]) (:215-281); the first ghost layer
//            receives the patch-boundary face updates as in the reference,
//            from the interior subcell next to it, which has just formed
//            that face's flux for its own update.
// The states are recomputed from the patch values instead of being stored
// (six state arrays per patch would be 6x the patch's HBM traffic); they are
// the same operations on the same inputs, hence the same values.
__device__ inline void muscl_face_state(const double* u, int nbox, const PatchGeom& g, const double* sQt, int idx,
                                        int d, int side, double h, double dt, double* w) {
  double q[V], qm[V], qp[V];
  load_q(u, nbox, idx, q);
  load_q(u, nbox, idx - g.stride(d), qm);
  load_q(u, nbox, idx + g.stride(d), qp);
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const double slope = minmod(qp[v] - q[v], q[v] - qm[v]) / h;
    w[v] = side ? q[v] + 0.5 * h * slope : q[v] - 0.5 * h * slope;
  }
  const int i = idx % g.S, j = (idx / g.S) % g.S, k = idx / (g.S * g.S);
  const int b = g.G, e = g.G + g.P;
  if (i >= b && i < e && j >= b && j < e && k >= b && k < e) {
    const int ci = ((k - b) * g.P + (j - b)) * g.P + (i - b), ni = g.P * g.P * g.P;
#pragma unroll
    for (int v = 0; v < V; ++v) w[v] += 0.5 * dt * sQt[v * ni + ci];
  }
}

// PC: subcells per direction (compile time: the index arithmetic of the
// recomputed face states is then constant; the launch's geometry must match).
// One patch per workgroup: u (with its ghost layers) -> un (interior and the
// first ghost layer's face updates); sQt: V P^3 doubles of LDS.  Returns the
// thread's part of the admissible time step.
template <class PDE, int PC>
__device__ inline double muscl_patch(const double* __restrict__ u, double* __restrict__ un, double* sQt,
                                     double dx0, double dx1, double dx2, double dt, double gamma, double cfl) {
  const PatchGeom g{PC, 2, PC + 4, 3};
  const int nbox = g.ncells_box();
  const double h[3] = {dx0 / g.P, dx1 / g.P, dx2 / g.P};
  const int ninterior = g.P * g.P * g.P;
  // phase 1: half-step evolution term of every interior subcell
  for (int c = threadIdx.x; c < ninterior; c += blockDim.x) {
    int i, j, k;
    interior_coords(g, c, i, j, k);
    const int idx = box_index(g, i, j, k);
    double q[V];
    load_q(u, nbox, idx, q);
    double Qt[V], F[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Qt[v] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double qm[V], qp[V], wL[V], wR[V];
      load_q(u, nbox, idx - g.stride(d), qm);
      load_q(u, nbox, idx + g.stride(d), qp);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const double slope = minmod(qp[v] - q[v], q[v] - qm[v]) / h[d];
        wL[v] = q[v] - 0.5 * h[d] * slope;
        wR[v] = q[v] + 0.5 * h[d] * slope;
      }
      PDE::flux_dir(wL, d, F, gamma);
#pragma unroll
      for (int v = 0; v < V; ++v) Qt[v] += F[v] / h[d];
      PDE::flux_dir(wR, d, F, gamma);
#pragma unroll
      for (int v = 0; v < V; ++v) Qt[v] -= F[v] / h[d];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) sQt[v * ninterior + c] = Qt[v];
  }
  __syncthreads();
  // phase 2: Rusanov fluxes of the evolved face states
  double dtm = 1.0e300;
  for (int c = threadIdx.x; c < ninterior; c += blockDim.x) {
    int i, j, k;
    interior_coords(g, c, i, j, k);
    const int idx = box_index(g, i, j, k);
    const int ijk[3] = {i, j, k};
    double acc[V], a[V], b[V], f[V];
    load_q(u, nbox, idx, acc);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double fac = dt / h[d];
      muscl_face_state(u, nbox, g, sQt, idx - g.stride(d), d, 1, h[d], dt, a);
      muscl_face_state(u, nbox, g, sQt, idx, d, 0, h[d], dt, b);
      double s = rusanov<PDE>(a, b, d, gamma, f);
      double cand = exa_dt_sane(cfl / 3 * h[d] / s);
      dtm = cand < dtm ? cand : dtm;
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] += fac * f[v];
      if (ijk[d] == g.G) {  // the lower ghost's face with this subcell
        const int gi = idx - g.stride(d);
#pragma unroll
        for (int v = 0; v < V; ++v) un[v * nbox + gi] = un[v * nbox + gi] - fac * f[v];
      }
      muscl_face_state(u, nbox, g, sQt, idx, d, 1, h[d], dt, a);
      muscl_face_state(u, nbox, g, sQt, idx + g.stride(d), d, 0, h[d], dt, b);
      s = rusanov<PDE>(a, b, d, gamma, f);
      cand = exa_dt_sane(cfl / 3 * h[d] / s);
      dtm = cand < dtm ? cand : dtm;
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] -= fac * f[v];
      if (ijk[d] == g.G + g.P - 1) {  // the upper ghost's
        const int gi = idx + g.stride(d);
#pragma unroll
        for (int v = 0; v < V; ++v) un[v * nbox + gi] = un[v * nbox + gi] + fac * f[v];
      }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) un[v * nbox + idx] = acc[v];
  }
  return dtm;
}

// Waves per SIMD the MUSCL-Hancock kernels are compiled for: up to 11^3
// subcells three workgroups' half-step terms fit the LDS (53 KB each at
// P = 11), and the kernel is then held to 168 VGPRs to let them all run
// (4.33 against 4.55 ms per launch on 27^3 patches of 11^3); larger patches
// keep the compiler's choice.
__host__ __device__ constexpr int muscl_min_waves(int P) { return P <= 11 ? 3 : 1; }

template <class PDE, int PC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(muscl_min_waves(PC)))) void k_muscl(const double* __restrict__ luh, double* __restrict__ luh_new,
                                               double* __restrict__ dt_patch, unsigned long long* dt_min,
                                               PatchGeom, double dx0, double dx1, double dx2,
                                               const double* __restrict__ dtp, double gamma, double cfl) {
  extern __shared__ double sQt[];  // [v][interior subcell]
  __shared__ double red[8];
  constexpr int nbox = (PC + 4) * (PC + 4) * (PC + 4);
  double dtm = muscl_patch<PDE, PC>(luh + (size_t)blockIdx.x * V * nbox, luh_new + (size_t)blockIdx.x * V * nbox,
                                    sQt, dx0, dx1, dx2, *dtp, gamma, cfl);
  dtm = block_min(dtm, red);
  if (threadIdx.x == 0) {
    if (dt_patch) dt_patch[blockIdx.x] = dtm;
    if (dt_min) atomic_min_pos_double(dt_min, dtm);
  }
}

// FV CFL time step of every patch (commons.cpph:26-67 2D, :410-463 3D).
template <int DIM, class PDE>
__global__ __launch_bounds__(256) void k_fv_stable_dt(const double* __restrict__ luh, double* __restrict__ dt_patch,
                                                      unsigned long long* dt_min, PatchGeom g, double dx0,
                                                      double dx1, double dx2, double gamma, double cfl) {
  __shared__ double red[8];
  const int nbox = g.ncells_box();
  const double* u = luh + (size_t)blockIdx.x * V * nbox;
  const double dxs[3] = {dx0, dx1, dx2};
  const int ninterior = DIM == 3 ? g.P * g.P * g.P : g.P * g.P;
  double dtm = 1.0e300;
  for (int c = threadIdx.x; c < ninterior; c += blockDim.x) {
    int i, j, k;
    interior_coords(g, c, i, j, k);
    double q[V];
    load_q(u, nbox, box_index(g, i, j, k), q);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double s = PDE::max_abs_eigenvalue(q, d, gamma);
      const double cand = exa_dt_sane(cfl / DIM * (dxs[d] / g.P) / s);
      dtm = cand < dtm ? cand : dtm;
    }
  }
  dtm = block_min(dtm, red);
  if (threadIdx.x == 0) {
    if (dt_patch) dt_patch[blockIdx.x] = dtm;
    if (dt_min) atomic_min_pos_double(dt_min, dtm);
  }
}

// ---------------------------------------------------------------------------
// Boundary hooks of a patch (commons.cpph 2D :99-403, 3D :503-982).  Face
// blocks hold (layer k, slower transverse a, faster transverse b): 3D
// idx_bnd_x(P,P,G), idx_bnd_y(P,G,P), idx_bnd_z(G,P,P); 2D idx_bnd_x(P,G),
// idx_bnd_y(G,P).
__host__ __device__ inline int fv_bnd(int dim, int d, int P, int G, int k, int a, int b) {
  if (dim == 2) return d == 0 ? b * G + k : k * P + b;
  if (d == 0) return (a * P + b) * G + k;
  if (d == 1) return (a * G + k) * P + b;
  return (k * P + a) * P + b;
}

// the transverse axes of direction d: o1 faster, o2 slower (2 = none in 2D)
__host__ __device__ inline void fv_axes(int dim, int d, int& o1, int& o2) {
  o1 = dim == 2 ? 1 - d : (d == 0 ? 1 : 0);
  o2 = dim == 2 ? 2 : (d == 2 ? 1 : 2);
}

// Layer of the patch that boundaryLayerExtraction puts at layer k of the
// face block (side 1: patchEnd-G+k; side 0: patchBegin+k, but the 3D y and
// z faces read patchBegin for every k, commons.cpph:561, :591).
__host__ __device__ inline int fv_extract_layer(int dim, int d, int side, int P, int G, int k) {
  return side ? P + k : (dim == 3 && d > 0 ? G : G + k);
}

// boundaryValues of the Euler FV solvers at an outside subcell centre:
// outflow copies, a reflecting wall negates the normal momentum, exact data
// give the state of bc.exact at (x, t).
__device__ inline void fv_boundary_values(const exa_bc& bc, const double x[3], double t, int d, double gamma,
                                          const double (&qin)[V], double (&qout)[V]) {
  if (bc.kind == EXA_BC_EXACT) {
    exa_ic_point(bc.exact, gamma, x, t, qout);
    return;
  }
#pragma unroll
  for (int v = 0; v < V; ++v) qout[v] = qin[v];
  if (bc.kind == EXA_BC_REFLECTING) qout[1 + d] = -qin[1 + d];
}

// The outside subcell centre of boundaryConditions for layer k (0 the
// outermost) at transverse (a, b) of face (d, side) of a cell.
__device__ inline void fv_outside_centre(int dim, int d, int side, int P, int G, int k, int a, int b,
                                         const double c[3], const double h[3], double x[3]) {
  int o1, o2;
  fv_axes(dim, d, o1, o2);
  const int K = G - k - 1;  // mirror
  x[0] = x[1] = x[2] = 0.0;
  x[d] = side ? c[d] + 0.5 * h[d] * (1.0 + k * 1.0 / (2.0 * P)) : c[d] - 0.5 * h[d] * (1.0 + K * 1.0 / (2.0 * P));
  const double qa = (2.0 * a + 1.0) / (2.0 * P), qb = (2.0 * b + 1.0) / (2.0 * P);
  if (dim == 3) x[o2] = c[o2] + h[o2] * (qa - 0.5);
  // the 3D x+ face scales its y coordinate by size[2] (commons.cpph:780)
  x[o1] = c[o1] + h[(dim == 3 && d == 0 && side) ? 2 : o1] * (qb - 0.5);
}

// Domain boundaries of a grid: the boundary values, the cell geometry of
// every patch and the device-held time stamp of the step.
struct FvBoundary {
  exa_bc bc;
  double off[3], h[3];
  int nx, ny;
  const double* time;
  double gamma;
};

// Face ghost layers: from the neighbours' interiors (ghostLayerFilling,
// commons.cpph:337-402, :876-982), or, where a closed axis leaves no
// neighbour (nbr < 0), as mergeWithBoundaryData fills them
// (E/solvers/FiniteVolumesSolver.cpp:689-728): the mirrored layer of the
// boundary-layer extraction through boundaryValues.  nbr[(2*d + side)*npatch
// + p]: neighbour of patch p on side (0: -, 1: +) of direction d.
template <int DIM>
__global__ __launch_bounds__(256) void k_fv_ghosts(double* __restrict__ luh, const int* __restrict__ nbr,
                                                   PatchGeom g, int npatch, FvBoundary fb) {
  const int p = blockIdx.x;
  const int nbox = g.ncells_box();
  const int trans = DIM == 3 ? g.P * g.P : g.P;
  const int per_side = g.G * trans;
  const int total = 2 * DIM * per_side;
  double* me = luh + (size_t)p * V * nbox;
  for (int it = threadIdx.x; it < total; it += blockDim.x) {
    const int ds = it / per_side, r = it % per_side;
    const int d = ds / 2, side = ds % 2;
    const int kk = r / trans, t = r % trans;
    const int ta = DIM == 3 ? t / g.P : 0, tb = t % g.P;  // transverse interior indices
    int o1, o2;
    fv_axes(DIM, d, o1, o2);
    int ci[3], ni[3];
    ci[o1] = ni[o1] = tb + g.G;
    ci[o2] = ni[o2] = DIM == 3 ? ta + g.G : 0;
    ci[d] = side ? g.G + g.P + kk : kk;
    const int q = nbr[(2 * d + side) * npatch + p];
    if (q >= 0) {
      ni[d] = side ? g.G + kk : g.P + kk;
      const double* nb = luh + (size_t)q * V * nbox;
      const int cidx = box_index(g, ci[0], ci[1], ci[2]);
      const int nidx = box_index(g, ni[0], ni[1], ni[2]);
#pragma unroll
      for (int v = 0; v < V; ++v) me[v * nbox + cidx] = nb[v * nbox + nidx];
    } else {
      const int K = g.G - kk - 1;
      ni[d] = fv_extract_layer(DIM, d, side, g.P, g.G, K);
      const int px = p % fb.nx, py = (p / fb.nx) % fb.ny, pz = p / (fb.nx * fb.ny);
      const double c[3] = {fb.off[0] + (px + 0.5) * fb.h[0], fb.off[1] + (py + 0.5) * fb.h[1],
                           DIM == 3 ? fb.off[2] + (pz + 0.5) * fb.h[2] : 0.0};
      double x[3], qin[V], qout[V];
      fv_outside_centre(DIM, d, side, g.P, g.G, kk, ta, tb, c, fb.h, x);
      load_q(me, nbox, box_index(g, ni[0], ni[1], ni[2]), qin);
      fv_boundary_values(fb.bc, x, *fb.time, d, fb.gamma, qin, qout);
      const int cidx = box_index(g, ci[0], ci[1], ci[2]);
#pragma unroll
      for (int v = 0; v < V; ++v) me[v * nbox + cidx] = qout[v];
    }
  }
}

// The per-patch hooks on one patch in the reference layout (node-major,
// nd values per subcell), one thread per (layer, transverse) entry:
// op 0 boundaryLayerExtraction (dst = luhbnd, src = luh), 1
// ghostLayerFillingAtBoundary (dst = luh, src = luhbnd), 2 ghostLayerFilling
// (dst = luh, src = the neighbour's luh), 3 boundaryConditions (dst =
// stateOut, src = stateIn; face d, side given by faceIndex).
struct FvHookArgs {
  int op, dim, d, side, P, G, nd;
  exa_bc bc;
  double c[3], h[3], t, gamma;
};

static __global__ __launch_bounds__(256) void k_fv_patch_hook(double* __restrict__ dst, const double* __restrict__ src,
                                                       FvHookArgs A) {
  const int dim = A.dim, P = A.P, G = A.G, S = P + 2 * G, d = A.d, side = A.side, nd = A.nd;
  const int na = dim == 3 ? P : 1;
  const int total = na * P * G;
  int o1, o2;
  fv_axes(dim, d, o1, o2);
  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < total; it += gridDim.x * blockDim.x) {
    const int k = it % G, b = (it / G) % P, a = it / (G * P);
    int x[3] = {0, 0, 0};
    x[o1] = G + b;
    if (dim == 3) x[o2] = G + a;
    auto cell = [&](int layer) {
      int y[3] = {x[0], x[1], x[2]};
      y[d] = layer;
      return dim == 3 ? ((size_t)y[2] * S + y[1]) * S + y[0] : (size_t)y[1] * S + y[0];
    };
    const size_t bk = (size_t)fv_bnd(dim, d, P, G, k, a, b);
    if (A.op == 0) {
      const size_t s = cell(fv_extract_layer(dim, d, side, P, G, k));
      for (int v = 0; v < nd; ++v) dst[bk * nd + v] = src[s * nd + v];
    } else if (A.op == 1) {
      const size_t o = cell(side ? G + P + k : k);
      for (int v = 0; v < nd; ++v) dst[o * nd + v] = src[bk * nd + v];
    } else if (A.op == 2) {
      const size_t o = cell(side ? G + P + k : k), s = cell(side ? G + k : P + k);
      for (int v = 0; v < nd; ++v) dst[o * nd + v] = src[s * nd + v];
    } else {
      const size_t bK = (size_t)fv_bnd(dim, d, P, G, G - k - 1, a, b);
      double xo[3], qin[V], qout[V];
      fv_outside_centre(dim, d, side, P, G, k, a, b, A.c, A.h, xo);
#pragma unroll
      for (int v = 0; v < V; ++v) qin[v] = src[bK * nd + v];
      fv_boundary_values(A.bc, xo, A.t, d, A.gamma, qin, qout);
#pragma unroll
      for (int v = 0; v < V; ++v) dst[bk * nd + v] = qout[v];
      for (int v = V; v < nd; ++v) dst[bk * nd + v] = src[bK * nd + v];
    }
  }
}

// Initial conditions at the subcell centres (commons.cpph:69-97).
template <int DIM>
__global__ __launch_bounds__(256) void k_fv_initial_conditions(double* __restrict__ luh, exa_ic ic, double gamma,
                                                               PatchGeom g, int nx, int ny, double ox, double oy,
                                                               double oz, double dx0, double dx1, double dx2) {
  const int p = blockIdx.x;
  const int nbox = g.ncells_box();
  const int px = p % nx, py = (p / nx) % ny, pz = p / (nx * ny);
  const double cx = ox + (px + 0.5) * dx0, cy = oy + (py + 0.5) * dx1, cz = oz + (pz + 0.5) * dx2;
  const int ninterior = DIM == 3 ? g.P * g.P * g.P : g.P * g.P;
  for (int c = threadIdx.x; c < ninterior; c += blockDim.x) {
    int i, j, k;
    interior_coords(g, c, i, j, k);
    const double qr = (2.0 * (i - g.G) + 1.0) / (2.0 * g.P);
    const double qs = (2.0 * (j - g.G) + 1.0) / (2.0 * g.P);
    const double qt = (2.0 * (k - g.G) + 1.0) / (2.0 * g.P);
    double X[3] = {cx + dx0 * (qr - 0.5), cy + dx1 * (qs - 0.5), DIM == 3 ? cz + dx2 * (qt - 0.5) : 0.0};
    double Q[V];
    exa_ic_point(ic, gamma, X, 0.0, Q);
    const int idx = box_index(g, i, j, k);
#pragma unroll
    for (int v = 0; v < V; ++v) luh[(size_t)p * V * nbox + v * nbox + idx] = Q[v];
  }
}

}  // namespace fv
}  // namespace exa
