// Kernels of the limiting ADER-DG step on a static adaptive mesh (one launch
// per stage over all leaves of every level, or over a device-built list of
// limiter patches).  The reference does this work cell by cell in its mesh
// traversals: E/solvers/ADERDGSolver.cpp (Riemann problems, face data of
// virtual cells), E/solvers/LimitingADERDGSolver.cpp:721-1490 (status switch,
// DMP, status spreading, rollback and recomputation).
//
// Leaves are numbered by (level, z, y, x); a leaf's face data are blocks
// [leaf][face][v][a*B+b] of the qf/ff arrays, and the face data of virtual
// cells (coarse-level cells next to a coarser leaf, virtual fine cells on a
// coarse leaf's boundary) are further blocks appended after the leaves'.
// Limiter patches live in pools of `capacity` patches [v][S^3] (S = 2N+1+4,
// two ghost layers), handed out from a free stack.
#pragma once
#include "aderdg_kernels.h"
#include "fv_kernels.h"
#include "limiter_kernels.h"

namespace exa {
namespace amrg {

// LimiterStatus, E/ADERDGCellDescription.def:423-453
enum : int { ST_OK = 0, ST_NOT4 = 1, ST_NOT3 = 2, ST_NOT2 = 3, ST_NOT1 = 4, ST_TROUBLED = 5 };

// counters kept on the device (summed over the steps)
enum : int {
  CNT_IRREGULAR_STEPS = 0,  // steps with a newly troubled cell (local recomputation)
  CNT_REFINE_REQUESTS,      // troubled cells below the finest level (the reference refines)
  CNT_COARSE_FINE,          // finest cells in the limiter domain next to a coarser cell
  CNT_MIRRORED,             // FV ghost faces filled from the patch itself
  CNT_CANNOT_COMMUNICATE,   // face pairs the reference aborts on
  CNT_PATCH_OVERFLOW,       // limiter patches that found no free slot
  CNT_TROUBLED,             // troubled finest cells, summed over steps
  CNT_FV_UPDATES,           // FV updates of finest cells (both passes)
  NCNT = 8
};

constexpr int NOBS = 5;   // DMP observables: the five conserved variables
constexpr int GHOST = 2;  // MUSCL-Hancock ghost width

__device__ inline bool is_fv(int s) { return s >= ST_NOT2; }
// mergeNeighboursBasedOnLimiterStatus, LimitingADERDGSolver.cpp:1695-1763
__device__ inline bool dg_merge(int a, int b) {
  const bool da = a == ST_OK || a == ST_NOT3 || a == ST_NOT4, db = b == ST_OK || b == ST_NOT3 || b == ST_NOT4;
  if (da && db) return true;
  return ((a == ST_NOT1 || a == ST_NOT2) && (b == ST_NOT3 || b == ST_NOT4)) ||
         ((b == ST_NOT1 || b == ST_NOT2) && (a == ST_NOT3 || a == ST_NOT4));
}
__device__ inline bool fv_merge(int a, int b) {
  if (is_fv(a) && is_fv(b)) return true;
  return ((a == ST_NOT1 || a == ST_NOT2) && (b == ST_NOT3 || b == ST_NOT4)) ||
         ((b == ST_NOT1 || b == ST_NOT2) && (a == ST_NOT3 || a == ST_NOT4));
}
// neighboursCanCommunicate, LimitingADERDGSolver.cpp:1682-1686
__device__ inline bool can_communicate(int a, int b) {
  return (a > ST_OK || b <= ST_NOT3) && (b > ST_OK || a <= ST_NOT3);
}
// std::min / std::max as the reference calls them
__device__ inline double smin(double a, double b) { return b < a ? b : a; }
__device__ inline double smax(double a, double b) { return a < b ? b : a; }

// Face relation of a leaf face (face_kind): the neighbour across it is a
// leaf of the same level, a coarser leaf, or a refined cell whose nine
// children on the face are leaves.
enum : int { FK_SAME = 0, FK_COARSER = 1, FK_FINER = 2 };

// A Rusanov problem: left block l (its face 2d+1), right block r (face 2d);
// la/lb: the two leaves when both are leaves of the finest level (the limiter
// status decides), else -1.
struct RFace {
  long long l, r;
  int d, la, lb, pad;
};

// ---------------------------------------------------------------------------
// small helpers

__global__ void k_fill(double* __restrict__ p, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_int(int* __restrict__ p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void k_copy_int(int* __restrict__ dst, const int* __restrict__ src, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Initial conditions at the GL nodes of every leaf (solutionAdjustment,
// kernels/aderdg/generic/c/3d/solutionAdjustment.cpph:17-52).
// geo[leaf] = (level, ix, iy, iz); h[level] = cell width.
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_amr_ic(double* __restrict__ luh, const int4* __restrict__ geo,
                                                       const double* __restrict__ h, exa_ic ic, double gamma) {
  constexpr int B = N + 1, B2 = B * B, B3 = B2 * B;
  const DGTables& T = c_tab[N];
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  if (s >= B3) return;
  const int4 g = geo[cell];
  const double dx = h[g.x];
  const int z = s / B2, y = (s / B) % B, x = s % B;
  const double c0 = (g.y + 0.5) * dx, c1 = (g.z + 0.5) * dx, c2 = (g.w + 0.5) * dx;
  double X[3];
  X[2] = c2 + dx * (T.xi[z] - 0.5);
  X[1] = c1 + dx * (T.xi[y] - 0.5);
  X[0] = c0 + dx * (T.xi[x] - 0.5);
  double Q[V];
  exa_ic_point(ic, gamma, X, 0.0, Q);
#pragma unroll
  for (int v = 0; v < V; ++v) luh[cell * V * B3 + v * B3 + s] = Q[v];
}

// FV initial conditions at the subcell centres of the listed patches
// (solutionAdjustment, kernels/finitevolumes/commons/c/commons.cpph:466-500):
// the limiter's own initial data where a patch exists from the start.
template <int P>
__global__ __launch_bounds__(256) void k_amr_fv_ic(double* __restrict__ pool, const int* __restrict__ cells,
                                                   const int* __restrict__ slots, const int* __restrict__ count,
                                                   const int4* __restrict__ geo, const double* __restrict__ h,
                                                   exa_ic ic, double gamma) {
  if ((int)blockIdx.x >= *count) return;
  constexpr int G = GHOST, S = P + 2 * G, S3 = S * S * S;
  const int4 g = geo[cells[blockIdx.x]];
  const double dx = h[g.x];
  const double c0 = (g.y + 0.5) * dx, c1 = (g.z + 0.5) * dx, c2 = (g.w + 0.5) * dx;
  double* o = pool + (size_t)slots[blockIdx.x] * V * S3;
  for (int e = threadIdx.x; e < P * P * P; e += blockDim.x) {
    const int i = e % P, j = (e / P) % P, k = e / (P * P);
    const double qi = (2.0 * i + 1.0) / (2.0 * P), qj = (2.0 * j + 1.0) / (2.0 * P), qk = (2.0 * k + 1.0) / (2.0 * P);
    double X[3] = {c0 + dx * (qi - 0.5), c1 + dx * (qj - 0.5), c2 + dx * (qk - 0.5)};
    double Q[V];
    exa_ic_point(ic, gamma, X, 0.0, Q);
    const int idx = ((k + G) * S + j + G) * S + i + G;
#pragma unroll
    for (int v = 0; v < V; ++v) o[(size_t)v * S3 + idx] = Q[v];
  }
}

// ---------------------------------------------------------------------------
// Riemann problems of the face list (solveRiemannProblemAtInterface,
// ADERDGSolver.cpp:2744-2825), skipping the pairs the limiter status
// excludes.  One wavefront per face.
template <int N>
__global__ __launch_bounds__(256) void k_riemann_faces(const double* __restrict__ qf, double* __restrict__ ff,
                                                       const RFace* __restrict__ faces, long long nfaces,
                                                       const int* __restrict__ status, double gamma,
                                                       unsigned long long* __restrict__ cnt) {
  constexpr int B2 = Ord<N>::B2;
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= nfaces) return;
  const RFace f = faces[i];
  if (f.la >= 0) {
    const int a = status[f.la], b = status[f.lb];
    if (lane == 0 && !can_communicate(a, b)) atomicAdd(&cnt[CNT_CANNOT_COMMUNICATE], 1ull);
    if (!dg_merge(a, b)) return;
  }
  riemann_face<N>(qf + (size_t)f.l * V * B2, qf + (size_t)f.r * V * B2, ff + (size_t)f.l * V * B2,
                  ff + (size_t)f.r * V * B2, f.d, gamma, lane);
}

// ---------------------------------------------------------------------------
// Metadata merge of every leaf face (Jacobi: reads the *_in arrays, writes
// the *_out arrays).  Limiter status (ADERDGSolver::mergeWithLimiterStatus,
// ADERDGSolver.cpp:2486-2504) on finest-level leaves: a face next to a
// same-level leaf takes max(0, max(own, other) - 1), a face next to a coarser
// cell merges with Ok.  DMP min/max (mergeSolutionMinMaxOnFace,
// LimitingADERDGSolver.cpp:1816-1850, with the observables copied to virtual
// fine cells and reduced into refined coarse-level cells,
// ADERDGSolver.cpp:2320-2345, :2442-2470) on every leaf when minmax != 0.
__global__ void k_merge_meta(const int* __restrict__ fw_in, int* __restrict__ fw_out,
                             const double* __restrict__ fmin_in, const double* __restrict__ fmax_in,
                             double* __restrict__ fmin_out, double* __restrict__ fmax_out,
                             const int* __restrict__ face_kind, const int* __restrict__ face_ref,
                             const int* __restrict__ anc_kids, const int* __restrict__ level, int finest,
                             long long first, long long nc, int minmax, const int* __restrict__ only_if) {
  if (only_if && !*only_if) return;
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  const bool fin = level[q] == finest;
  for (int f = 0; f < 6; ++f) {
    const int kind = face_kind[q * 6 + f], ref = face_ref[q * 6 + f], fo = f ^ 1;
    if (fin) {
      const int own = min(fw_in[q * 6 + f], (int)ST_TROUBLED);
      int m = own;
      if (kind == FK_SAME) m = max(own, min(fw_in[(long long)ref * 6 + fo], (int)ST_TROUBLED));
      fw_out[q * 6 + f] = max(0, m - 1);
    } else if (fw_out != fw_in) {
      fw_out[q * 6 + f] = fw_in[q * 6 + f];
    }
    if (!minmax) continue;
    const double* mi = fmin_in + (q * 6 + f) * NOBS;
    const double* ma = fmax_in + (q * 6 + f) * NOBS;
    double lo[NOBS], hi[NOBS];
    if (kind == FK_FINER) {
      for (int v = 0; v < NOBS; ++v) {
        lo[v] = 1.7976931348623157e308;
        hi[v] = -1.7976931348623157e308;
      }
      for (int k = 0; k < 9; ++k) {
        const long long c = anc_kids[ref * 9 + k];
        for (int v = 0; v < NOBS; ++v) {
          lo[v] = smin(fmin_in[(c * 6 + fo) * NOBS + v], lo[v]);
          hi[v] = smax(fmax_in[(c * 6 + fo) * NOBS + v], hi[v]);
        }
      }
    } else {
      for (int v = 0; v < NOBS; ++v) {
        lo[v] = fmin_in[((long long)ref * 6 + fo) * NOBS + v];
        hi[v] = fmax_in[((long long)ref * 6 + fo) * NOBS + v];
      }
    }
    for (int v = 0; v < NOBS; ++v) {
      fmin_out[(q * 6 + f) * NOBS + v] = smin(mi[v], lo[v]);
      fmax_out[(q * 6 + f) * NOBS + v] = smax(ma[v], hi[v]);
    }
  }
}

// ---------------------------------------------------------------------------
// Lists of finest-level leaves, built on the device (atomic append; every
// listed item is processed independently, so the order does not matter).
enum : int {
  LIST_FV_OLD = 0,     // status at the start of the step in the FV range
  LIST_FV_NEW = 1,     // status after the update in the FV range
  LIST_NEW_SLOT = 2,   // patch allocated in this step
  LIST_DG_PATCH = 3,   // holds a patch, advanced by ADER-DG in this step
};
__global__ void k_build_list(int mode, const int* __restrict__ status_old, const int* __restrict__ status,
                             const int* __restrict__ slot_of, const int* __restrict__ new_slot,
                             const int* __restrict__ fv_done, long long first, long long nc,
                             int* __restrict__ cells, int* __restrict__ slots, int* __restrict__ count,
                             const int* __restrict__ only_if) {
  if (only_if && !*only_if) return;
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  const int slot = slot_of[q];
  if (slot < 0) return;
  bool take = false;
  switch (mode) {
    case LIST_FV_OLD: take = is_fv(status_old[q]); break;
    case LIST_FV_NEW: take = is_fv(status[q]); break;
    case LIST_NEW_SLOT: take = new_slot[q] != 0; break;
    default: take = fv_done[q] == 0; break;
  }
  if (!take) return;
  const int i = atomicAdd(count, 1);
  cells[i] = (int)q;
  slots[i] = slot;
}

// FV ghost layers of the listed patches from the interiors of their
// same-level neighbours' patches (ghostLayerFilling,
// kernels/finitevolumes/commons/c/commons.cpph:876-982) where the limiter
// statuses exchange them (fv_merge); otherwise the patch's own boundary
// layers mirrored.  One workgroup per patch.
template <int P>
__global__ __launch_bounds__(256) void k_ghost_fill(double* __restrict__ pool, const int* __restrict__ cells,
                                                    const int* __restrict__ slots, const int* __restrict__ count,
                                                    const int* __restrict__ slot_of, const int* __restrict__ status,
                                                    const int* __restrict__ nbr6,
                                                    unsigned long long* __restrict__ cnt) {
  if ((int)blockIdx.x >= *count) return;
  constexpr int G = GHOST, S = P + 2 * G, S3 = S * S * S;
  const int q = cells[blockIdx.x];
  double* me = pool + (size_t)slots[blockIdx.x] * V * S3;
  __shared__ const double* src[6];
  if (threadIdx.x < 6) {
    const int f = threadIdx.x;
    const int r = nbr6[(size_t)q * 6 + f];
    const int rs = r >= 0 ? slot_of[r] : -1;
    if (rs >= 0 && fv_merge(status[q], status[r])) {
      src[f] = pool + (size_t)rs * V * S3;
    } else {
      src[f] = nullptr;
      atomicAdd(&cnt[CNT_MIRRORED], 1ull);
    }
  }
  __syncthreads();
  constexpr int per_face = G * P * P;
  const int total = 6 * per_face * V;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int v = e / (6 * per_face);
    const int r0 = e % (6 * per_face);
    const int f = r0 / per_face, r1 = r0 % per_face;
    const int k = r1 / (P * P), t = r1 % (P * P);
    const int a = t / P + G, b = t % P + G;  // the two tangential indices (slow, fast)
    const int d = f >> 1, side = f & 1;
    int c[3];
    // tangential axes in index order (x fastest)
    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
    c[t0] = b;
    c[t1] = a;
    c[d] = side ? G + P + k : k;
    const int dst = (c[2] * S + c[1]) * S + c[0];
    const double* from = src[f];
    if (from) {
      c[d] = side ? G + k : P + k;  // the neighbour's interior layers next to the shared face
    } else {
      from = me;
      c[d] = side ? G + P - 1 - k : 2 * G - 1 - k;  // own boundary layers, mirrored
    }
    const int sidx = (c[2] * S + c[1]) * S + c[0];
    me[(size_t)v * S3 + dst] = from[(size_t)v * S3 + sidx];
  }
}

// MUSCL-Hancock update of the listed patches: old pool -> new pool.
template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fv::muscl_min_waves(P)))) void k_muscl_list(const double* __restrict__ pool_old, double* __restrict__ pool_new,
                                                    const int* __restrict__ slots, const int* __restrict__ count,
                                                    double dx, const InstParams* __restrict__ inst, int level,
                                                    double gamma) {
  if ((int)blockIdx.x >= *count) return;
  extern __shared__ double sQt[];
  constexpr int S3 = (P + 4) * (P + 4) * (P + 4);
  const size_t o = (size_t)slots[blockIdx.x] * V * S3;
  // the kernel's own CFL (kernels/finitevolumes/musclhancock/c/musclhancock.h:29)
  fv::muscl_patch<EulerPDE, P>(pool_old + o, pool_new + o, sQt, dx, dx, dx, inst[level].dt_cur, gamma, 0.99);
}

// DG <-> FV projections of the listed cells (Limiter.cpp:23-112).
template <int P>
__global__ __launch_bounds__(256) void k_project_fv_list(const double* __restrict__ luh, double* __restrict__ pool,
                                                         const int* __restrict__ cells, const int* __restrict__ slots,
                                                         const int* __restrict__ count,
                                                         const lim::LimTables* __restrict__ tab) {
  if ((int)blockIdx.x >= *count) return;
  extern __shared__ double sm[];
  constexpr int B = (P + 1) / 2, S = P + 2 * GHOST;
  lim::project_on_fv_block_t<B>(luh + (size_t)cells[blockIdx.x] * V * B * B * B,
                                pool + (size_t)slots[blockIdx.x] * V * S * S * S, tab, GHOST, sm);
}
template <int P>
__global__ __launch_bounds__(256) void k_project_dg_list(const double* __restrict__ pool, double* __restrict__ luh,
                                                         const int* __restrict__ cells, const int* __restrict__ slots,
                                                         const int* __restrict__ count,
                                                         const lim::LimTables* __restrict__ tab,
                                                         int* __restrict__ fv_done,
                                                         unsigned long long* __restrict__ cnt) {
  if ((int)blockIdx.x >= *count) return;
  extern __shared__ double sm[];
  constexpr int B = (P + 1) / 2, S = P + 2 * GHOST;
  const int q = cells[blockIdx.x];
  lim::project_on_dg_block_t<B>(pool + (size_t)slots[blockIdx.x] * V * S * S * S, luh + (size_t)q * V * B * B * B,
                                tab, GHOST, sm);
  if (threadIdx.x == 0) {
    fv_done[q] = 1;
    atomicAdd(&cnt[CNT_FV_UPDATES], 1ull);
  }
}

// Rollback of the listed cells that ADER-DG advanced in this step to the
// previous solution (ADERDGSolver::rollbackSolution).
__global__ void k_rollback_list(double* __restrict__ luh, const double* __restrict__ luh_prev,
                                const int* __restrict__ cells, const int* __restrict__ count,
                                const int* __restrict__ fv_done, int B3) {
  if ((int)blockIdx.x >= *count) return;
  const int q = cells[blockIdx.x];
  if (fv_done[q]) return;
  const size_t o = (size_t)q * V * B3;
  for (int e = threadIdx.x; e < V * B3; e += blockDim.x) luh[o + e] = luh_prev[o + e];
}

// Min/max of the interior subcells of the listed patches
// (findCellLocalLimiterMinAndMax, Limiter.cpp:167-205).  out [cell][2][NOBS].
template <int P>
__global__ __launch_bounds__(256) void k_patch_min_max(const double* __restrict__ pool, const int* __restrict__ cells,
                                                       const int* __restrict__ slots, const int* __restrict__ count,
                                                       double* __restrict__ out) {
  if ((int)blockIdx.x >= *count) return;
  constexpr int G = GHOST, S = P + 2 * G, S3 = S * S * S, P3 = P * P * P;
  __shared__ double red[2][4];
  const double* u = pool + (size_t)slots[blockIdx.x] * V * S3;
  const int q = cells[blockIdx.x];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int v = 0; v < NOBS; ++v) {
    double mn = 1.7976931348623157e308, mx = -1.7976931348623157e308;
    for (int e = threadIdx.x; e < P3; e += blockDim.x) {
      const int x = e % P + G, y = (e / P) % P + G, z = e / (P * P) + G;
      const double a = u[(size_t)v * S3 + (z * S + y) * S + x];
      mn = smin(mn, a);
      mx = smax(mx, a);
    }
    mn = wave_min(mn);
    mx = -wave_min(-mx);
    if (lane == 0) {
      red[0][wid] = mn;
      red[1][wid] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
        mn = smin(mn, red[0][w]);
        mx = smax(mx, red[1][w]);
      }
      out[((size_t)q * 2 + 0) * NOBS + v] = mn;
      out[((size_t)q * 2 + 1) * NOBS + v] = mx;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// DMP, admissibility and the new limiter status of every leaf after the
// update: discreteMaximumPrincipleAndMinAndMaxSearch (Limiter.cpp:114-165)
// against the merged face blocks, isPhysicallyAdmissible of the Euler solver
// (positive density and energy minima), determineLimiterStatusAfterSolutionUpdate
// (LimitingADERDGSolver.cpp:933-991).  mm: cell min/max of the new solution,
// lmm: limiter-patch min/max of the cells the FV scheme advanced.
__global__ void k_status(const double* __restrict__ mm, const double* __restrict__ lmm, double* __restrict__ fmin,
                         double* __restrict__ fmax, const int* __restrict__ status_old, int* __restrict__ status,
                         int* __restrict__ fw, int* __restrict__ cure, const int* __restrict__ level, int finest,
                         long long nc, double relax, double scaling, int steps_till_cured, int* __restrict__ irregular,
                         unsigned long long* __restrict__ cnt) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  bool dmp = true;
  double lmin[NOBS];
  for (int v = 0; v < NOBS; ++v) {
    double bmin = fmin[(q * 6 + 0) * NOBS + v], bmax = fmax[(q * 6 + 0) * NOBS + v];
    for (int i = 1; i < 6; ++i) {
      bmin = smin(bmin, fmin[(q * 6 + i) * NOBS + v]);
      bmax = smax(bmax, fmax[(q * 6 + i) * NOBS + v]);
    }
    double sd = (bmax - bmin) * scaling;
    sd = smax(sd, relax);
    const double lo = mm[(q * 2 + 0) * NOBS + v], hi = mm[(q * 2 + 1) * NOBS + v];
    if (lo < bmin - sd || hi > bmax + sd) dmp = false;
    lmin[v] = lo;
    fmin[(q * 6 + 0) * NOBS + v] = lo;
    fmax[(q * 6 + 0) * NOBS + v] = hi;
    fmin[(q * 6 + 1) * NOBS + v] = bmin;
    fmax[(q * 6 + 1) * NOBS + v] = bmax;
  }
  const bool troubled = !(dmp && lmin[0] > 0.0 && lmin[4] > 0.0);
  if (level[q] != finest) {
    if (troubled) atomicAdd(&cnt[CNT_REFINE_REQUESTS], 1ull);
    return;
  }
  const int s_old = status_old[q];
  if (is_fv(s_old)) {
    for (int i = 0; i < 6; ++i)
      for (int v = 0; v < NOBS; ++v) {
        fmin[(q * 6 + i) * NOBS + v] = lmm[(q * 2 + 0) * NOBS + v];
        fmax[(q * 6 + i) * NOBS + v] = lmm[(q * 2 + 1) * NOBS + v];
      }
  }
  int s = s_old;
  int* w = fw + q * 6;
  if (troubled) {
    atomicAdd(&cnt[CNT_TROUBLED], 1ull);
    cure[q] = steps_till_cured + 1;
    if (s_old != ST_TROUBLED && s_old != ST_NOT1) atomicOr(irregular, 1);
    s = ST_TROUBLED;
    for (int i = 0; i < 6; ++i) w[i] = s;
  } else if (s_old == ST_TROUBLED) {
    s = ST_TROUBLED;
    for (int i = 0; i < 6; ++i) w[i] = s;
    cure[q] -= 1;
    if (cure[q] == 0) {
      s = ST_NOT1;
      for (int i = 0; i < 6; ++i) w[i] = s;
      cure[q] = steps_till_cured;
    }
  }
  s = w[0];
  for (int i = 1; i < 6; ++i) s = max(s, w[i]);
  status[q] = s;
  for (int i = 0; i < 6; ++i) w[i] = s;
}

// One status-spreading update after a merge of the facewise statuses
// (updateLimiterStatusDuringLimiterStatusSpreading, :369-379): merged values
// in fw_m; a troubled cell stays troubled.
__global__ void k_spread(const int* __restrict__ fw_m, int* __restrict__ fw, int* __restrict__ status,
                         long long first, long long nc, const int* __restrict__ only_if) {
  if (!*only_if) return;
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  int s;
  if (status[q] >= ST_TROUBLED) {
    s = status[q];
  } else {
    s = fw_m[q * 6];
    for (int i = 1; i < 6; ++i) s = max(s, fw_m[q * 6 + i]);
  }
  status[q] = s;
  for (int i = 0; i < 6; ++i) fw[q * 6 + i] = s;
}

// After the spreading: troubled cells restart their cure count
// (reinitialiseSolvers, LimitingADERDGSolver.cpp:1358-1361).
__global__ void k_cure_reset(const int* __restrict__ status, int* __restrict__ cure, long long first, long long nc,
                             int steps_till_cured, const int* __restrict__ only_if) {
  if (!*only_if) return;
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  if (status[q] == ST_TROUBLED) cure[q] = 1 + steps_till_cured;
}

// Limiter patches: released by cells back to Ok, taken by cells of the
// limiter domain without one (a free stack of slot indices).
__global__ void k_slots_free(const int* __restrict__ status, int* __restrict__ slot_of, int* __restrict__ stack,
                             int* __restrict__ top, long long first, long long nc) {
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  if (slot_of[q] >= 0 && status[q] == ST_OK) {
    stack[atomicAdd(top, 1)] = slot_of[q];
    slot_of[q] = -1;
  }
}
__global__ void k_slots_alloc(const int* __restrict__ status, int* __restrict__ slot_of, int* __restrict__ new_slot,
                              const int* __restrict__ stack, int* __restrict__ top, long long first, long long nc,
                              unsigned long long* __restrict__ cnt) {
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  new_slot[q] = 0;
  if (slot_of[q] < 0 && status[q] >= ST_NOT4) {
    const int i = atomicSub(top, 1) - 1;
    if (i < 0) {
      atomicAdd(top, 1);
      atomicAdd(&cnt[CNT_PATCH_OVERFLOW], 1ull);
      return;
    }
    slot_of[q] = stack[i];
    new_slot[q] = 1;
  }
}

// The face blocks of every leaf from its own min/max after a recomputation
// (determineMinAndMax, LimitingADERDGSolver.cpp:1050-1116): the limiter
// patch's for finest cells in the FV range, the DG solution's otherwise.
__global__ void k_minmax_reset(const double* __restrict__ mm, const double* __restrict__ lmm,
                               double* __restrict__ fmin, double* __restrict__ fmax, const int* __restrict__ status,
                               const int* __restrict__ level, int finest, long long nc,
                               const int* __restrict__ only_if) {
  if (!*only_if) return;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  const double* src = (level[q] == finest && is_fv(status[q])) ? lmm : mm;
  for (int i = 0; i < 6; ++i)
    for (int v = 0; v < NOBS; ++v) {
      fmin[(q * 6 + i) * NOBS + v] = src[(q * 2 + 0) * NOBS + v];
      fmax[(q * 6 + i) * NOBS + v] = src[(q * 2 + 1) * NOBS + v];
    }
}

// ADER-DG stable time step of every leaf (stableTimeStepSize.cpph:21-61)
// into its level's running minimum.  A node whose wave speed is not a number
// (a negative pressure or a non-finite value) has no admissible step: in a
// cell the FV scheme advances (finest level, status NeighbourOfTroubled2 and
// above) its DG solution is only the projection of the subcells and the
// node is skipped, as the reference's std::min(dt, NaN) skips it; in every
// other cell it becomes 0, which the sync call reports.
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_amr_dt(const double* __restrict__ luh, const int* __restrict__ status,
                                                       const int* __restrict__ level, int finest,
                                                       InstParams* __restrict__ inst, double cfl) {
  using O = Ord<N>;
  constexpr int B3 = O::B3;
  __shared__ double red[16];
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  const int l = level[cell];
  InstParams& P = inst[l];
  const bool skip_nan = l == finest && is_fv(status[cell]);
  double dtn = 1.0e300;
  if (s < B3) {
    double u[V];
#pragma unroll
    for (int v = 0; v < V; ++v) u[v] = luh[cell * V * B3 + v * B3 + s];
    double m[3];
    EulerPDE::max_abs_eigenvalues3(u, P.gamma, m);
    double den = 0.0;
    den += m[0] / P.dx[0];
    den += m[1] / P.dx[1];
    den += m[2] / P.dx[2];
    const double cand = cfl * exa_pnpm(N) / den;
    dtn = (skip_nan && !(cand == cand)) ? 1.0e300 : exa_dt_sane(cand);
  }
  dtn = block_min(dtn, red);
  if (s == 0) atomic_min_pos_double(&P.dt_next, dtn);
}

// End of a step: counters, the global time step of all levels (global time
// stepping: every level advances with the minimum over all leaves).
__global__ void k_step_end(InstParams* __restrict__ inst, int nlev, const int* __restrict__ irregular,
                           const int* __restrict__ status, const int* __restrict__ nbr6, long long first, long long nc,
                           unsigned long long* __restrict__ cnt) {
  const long long q = first + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nc && status[q] > ST_OK) {
    bool cf = false;
    for (int f = 0; f < 6; ++f) cf |= nbr6[q * 6 + f] < 0;
    if (cf) atomicAdd(&cnt[CNT_COARSE_FINE], 1ull);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (*irregular) cnt[CNT_IRREGULAR_STEPS] += 1;
    unsigned long long m = inst[0].dt_next;
    for (int l = 1; l < nlev; ++l) m = inst[l].dt_next < m ? inst[l].dt_next : m;
    for (int l = 0; l < nlev; ++l) {
      inst[l].time += inst[l].dt_cur;
      inst[l].dt_cur = __longlong_as_double((long long)m);
      inst[l].dt_next = (unsigned long long)__double_as_longlong(1.0e300);
    }
  }
}

// First time step: the minimum over all leaves into every level's dt_cur.
__global__ void k_dt_take_all(InstParams* __restrict__ inst, int nlev) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  unsigned long long m = inst[0].dt_next;
  for (int l = 1; l < nlev; ++l) m = inst[l].dt_next < m ? inst[l].dt_next : m;
  for (int l = 0; l < nlev; ++l) {
    inst[l].dt_cur = __longlong_as_double((long long)m);
    inst[l].dt_next = (unsigned long long)__double_as_longlong(1.0e300);
  }
}
__global__ void k_dt_reset_all(InstParams* __restrict__ inst, int nlev) {
  if (blockIdx.x == 0 && threadIdx.x < nlev)
    inst[threadIdx.x].dt_next = (unsigned long long)__double_as_longlong(1.0e300);
}

}  // namespace amrg
}  // namespace exa
