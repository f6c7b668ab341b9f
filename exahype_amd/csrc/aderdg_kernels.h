// ADER-DG kernels for gfx950 (3D, nonlinear, conservative flux only).
//
// Device layout of a cell (structure of arrays inside the cell block, cells
// contiguous in HBM):
//   luh, lduh : [cell][v][node]            node = (z*B + y)*B + x
//   qf, ff    : [cell][face][v][a*B + b]   face order x-,x+,y-,y+,z-,z+,
//                                          (a,b) = (z,y) | (z,x) | (y,x)
// One workgroup owns one cell; thread s owns spatial node s.  All summation
// orders and expression groupings follow the reference generic kernels so
// the FP64 results agree with them to rounding.
#pragma once
#include "common.h"
#include "pde_euler.h"
#include "pde_generic.h"
#include "mfma_util.h"
#include <type_traits>

namespace exa {

constexpr int V = 5;

template <int N>
struct Ord {
  static constexpr int B = N + 1, B2 = B * B, B3 = B2 * B, B4 = B2 * B2;
  static constexpr int NT = round_up64(B3);          // threads per cell
  // LDS stride between the pencils of a direction layout: odd, so that lanes
  // addressing consecutive pencils (phase P) or the same node of consecutive
  // pencils (phases F and A of the y and z layouts) fall into distinct banks
  static constexpr int PS = (B % 2 == 0) ? B + 1 : B;
  static constexpr int BLK = B2 * PS;                // one (direction, variable) block of pencils
  static constexpr int LDS_DOUBLES = 6 * V * BLK + 3 * B3 + 16;  // flux slices + pencil constants + reduction scratch
  // staged predictor: one flux slice + pencil constants during the sweeps,
  // time averages sQh/sFh afterwards; the reduction scratch of the residual
  // sum aliases the flux slice (80 KB at N = 7: two workgroups per CU)
  static constexpr int LDS_STAGED = 3 * V * BLK + 3 * B3 > 4 * V * B3 ? 3 * V * BLK + 3 * B3 : 4 * V * B3;
};

// ---------------------------------------------------------------------------
// Volume integral and face extrapolation from the time-averaged predictor held
// in LDS: sQh[v][node], sFh[d][v][node] (d = x,y,z) -- shared by the fused
// predictor and the stand-alone volume-integral entry point.
// volumeIntegralNonlinear.cpph:26-125, aderExtrapolatorNonlinear :400-519.
template <int N>
__device__ inline void volume_integral_node(const double* sFh, int s, int z, int y, int x,
                                            double idx0, double idx1, double idx2, double out[V]) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  const DGTables& T = c_tab[N];
  const double usx = (T.w[z] * T.w[y]) / idx0;
  const double usy = (T.w[z] * T.w[x]) / idx1;
  const double usz = (T.w[y] * T.w[x]) / idx2;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < B; ++m) acc += (T.Kxi[x][m] * sFh[(0 * V + v) * B3 + z * B2 + y * B + m]) * usx;
#pragma unroll
    for (int m = 0; m < B; ++m) acc += (T.Kxi[y][m] * sFh[(1 * V + v) * B3 + z * B2 + m * B + x]) * usy;
#pragma unroll
    for (int m = 0; m < B; ++m) acc += (T.Kxi[z][m] * sFh[(2 * V + v) * B3 + m * B2 + y * B + x]) * usz;
    out[v] = acc;
  }
}

template <int N>
__device__ inline void extrapolate_faces(const double* sQh, const double* sFh, int s, double* qf_cell,
                                         double* ff_cell) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  if (s >= 3 * B2) return;
  const DGTables& T = c_tab[N];
  const int d = s / B2, ab = s % B2, a = ab / B, b = ab % B;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    double qL = 0.0, qR = 0.0, fL = 0.0, fR = 0.0;
#pragma unroll
    for (int l = 0; l < B; ++l) {
      const int idx = d == 0 ? a * B2 + b * B + l : (d == 1 ? a * B2 + l * B + b : l * B2 + a * B + b);
      const double qv = sQh[v * B3 + idx];
      const double fv = sFh[(d * V + v) * B3 + idx];
      fL += fv * T.FL[l];
      fR += fv * T.FR[l];
      qL += qv * T.FL[l];
      qR += qv * T.FR[l];
    }
    qf_cell[((2 * d) * V + v) * B2 + ab] = qL;
    qf_cell[((2 * d + 1) * V + v) * B2 + ab] = qR;
    ff_cell[((2 * d) * V + v) * B2 + ab] = fL;
    ff_cell[((2 * d + 1) * V + v) * B2 + ab] = fR;
  }
}

// Debug outputs of the per-cell entry point, reference layouts with nVar=5.
struct PredictorDebug {
  double* lQi;   // [z][y][x][t][V]
  double* lFi;   // [t][z][y][x][4][V]
  double* lQhi;  // [z][y][x][V]
  double* lFhi;  // 4 blocks of B^3*V
};

// Ensemble of independent grids advanced by the same launches: every cell
// belongs to an instance with its own mesh width, gamma, time and dt.
struct InstParams {
  double dx[3];
  double gamma;
  double dt_cur;
  unsigned long long dt_next;  // bit pattern of the running positive minimum
  double time;
  double off[3];
  long long base;  // first cell of the instance in the group's arrays
  int nx, ny, nz, pad;
};
struct Ens {
  const int* cell_inst;  // instance of every cell; nullptr: a single grid
  InstParams* inst;
  const exa_ic* ics;  // per-instance initial conditions
  double grav;        // gravity g along -z of the Euler-with-gravity PDE; 0: no source
  double adv[3];      // advection velocity of the linear-advection PDEs
  PointSrc ps;        // point source of the linear predictor (mode 0: none)
  // 1: the Euler predictors write lduh = luh + (dt / w^3) * (volume
  // integral) -- the update's volume part, for k_riemann_update, which then
  // needs neither luh nor the bare volume integral; 0: lduh = volume integral
  int fold_update;
};

// Algebraic source of Euler under gravity g along -z (the useSource arm of
// spaceTimePredictorNonlinear.cpph:227-252): S = (0, 0, 0, -rho g, -m_z g).
// Component v of S at state q (only v = 3, 4 are nonzero).
template <typename T>
__device__ inline T gravity_source(const T* q, int v, T g) {
  return v == 3 ? -q[0] * g : (v == 4 ? -q[3] * g : T(0));
}

// ---------------------------------------------------------------------------
// Picard sweeps of the space-time predictor for one cell held by one
// workgroup (aderPicardLoopNonlinear, spaceTimePredictorNonlinear.cpph:43-317).
//
// Thread s owns spatial node s and keeps its whole time history q[t][v] in
// registers.  Per time slice t, three phases separated by barriers:
//   F  node owners evaluate the flux and store F_d in direction-contiguous
//      layouts (x: [z][y][x], y: [z][x][y], z: [y][x][z]), so every pencil is
//      one contiguous run of B values;
//   P  pencil owners (3*V*B^2 items: direction, variable, pencil) read their
//      pencil once, scale it by the pencil constant c_d = ((w_t w_a) w_b)
//      dt/dx_d (a table in LDS), and produce all B outputs of the contraction
//      with Kxi, whose entries are uniform across the workgroup (scalar
//      operands), stored in the same direction layout;
//   A  node owners combine rhs = w^3 F0[t] u - Px - Py - Pz and accumulate the
//      thread-local time contraction qn[t'] += (rhs / w^3) iK1[t'][t].
// T is the arithmetic type of the sweep: double for the plain method and the
// correction sweeps, float for the low-precision sweeps of the mixed
// predictor.  Returns the sweeps executed; stops once
// sum((q_new-q_old)^2) < tol2 over the cell (block-uniform decision).
template <int N, typename T, bool ACC_FH, bool DEBUG, bool PAIR = false>
__device__ inline int picard_sweeps(T (&q)[Ord<N>::B][V], const double (&u)[V], T* sF, T* sP, const T* sc,
                                    double* red, bool act, int s, int z, int y, int x, double gamma, int max_sweeps,
                                    double tol2, double (&fh)[3][V], PredictorDebug dbg, double grav = 0.0,
                                    double dt = 0.0, double* sh = nullptr) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3, PS = O::PS, BLK = O::BLK;
  constexpr int NITEMS = 3 * V * B2;
  constexpr bool DP = std::is_same<T, double>::value;
  const DGTables& Td = c_tab[N];
  const DGTablesF& Tf = c_tabf[N];
  // tz is an opaque zero (see the asm in the slice loop): the Kxi and iK1
  // reads through it are re-issued every slice as scalar loads instead of
  // being hoisted out of the sweep into registers the kernel cannot spare
  int tz = 0;
  auto W = [&](int i) -> T { return DP ? (T)Td.w[i] : (T)Tf.w[i]; };
  auto KXI = [&](int a, int b) -> T { return DP ? (T)c_tab[N + tz].Kxi[a][b] : (T)c_tabf[N + tz].Kxi[a][b]; };
  auto IK1 = [&](int a, int b) -> T { return DP ? (T)c_tab[N + tz].iK1[a][b] : (T)c_tabf[N + tz].iK1[a][b]; };
  // (F0 too: w^3 F0[t] u would otherwise be hoisted for every t, 30 registers)
  auto F0 = [&](int i) -> T { return DP ? (T)c_tab[N + tz].F0[i] : (T)c_tabf[N + tz].F0[i]; };

  T uu[V];
#pragma unroll
  for (int v = 0; v < V; ++v) uu[v] = (T)u[v];
  const T w3 = W(z) * W(y) * W(x);
  const T iw3 = T(1.0) / w3;
  const T tg = (T)gamma;
  // offsets of this node in the three direction layouts (flux and partials)
  const int ox = (z * B + y) * PS + x, oy = (z * B + x) * PS + y, oz = (y * B + x) * PS + z;

  // per-thread item tables: LDS offset and pencil-constant index of every
  // item a thread owns, the same for every slice and sweep (paired form up
  // to N=5 only; at N=7 the extra registers would push the sweeps into
  // scratch)
  constexpr bool TAB = B <= 6;
  constexpr bool PTAB = PAIR && TAB;
  constexpr int PR = TAB ? ((PAIR ? 2 : 1) * NITEMS + O::NT - 1) / O::NT : 1;
  int po[PR], pc[PR];
#pragma unroll
  for (int r = 0; r < (TAB ? PR : 0); ++r) {
    const int i = s + r * O::NT;
    const int kdv = i / B2, ab = i - kdv * B2;   // kdv = slice-in-pair * 3V + d*V + v
    const int k = kdv / (3 * V), d = (kdv - k * 3 * V) / V;
    po[r] = kdv * BLK + ab * PS;
    pc[r] = (d * B + k) * B2 + ab;
  }
  int it = 0;
  while (it < max_sweeps) {
    ++it;
    T qn[B][V];
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) qn[t][v] = T(0);
    if constexpr (!PAIR) {
#pragma unroll
      for (int t = 0; t < B; ++t) {
        asm volatile("" : "+s"(tz));
        // phase F
        if (act) {
          T F[3][V];
          EulerPDE::flux<T>(q[t], F, tg);
#pragma unroll
          for (int v = 0; v < V; ++v) {
            sF[(0 * V + v) * BLK + ox] = F[0][v];
            sF[(1 * V + v) * BLK + oy] = F[1][v];
            sF[(2 * V + v) * BLK + oz] = F[2][v];
          }
        }
        __syncthreads();
        // phase P: item = (d*V + v, pencil ab); pencils and partials are
        // contiguous runs of B in the direction layouts, so the item needs no
        // direction-dependent addressing
        auto item = [&](const T c, int off) {
          const T* pf = sF + off;
          T f[B];
#pragma unroll
          for (int n = 0; n < B; ++n) f[n] = c * pf[n];
          T* pp = sP + off;
#pragma unroll
          for (int l = 0; l < B; ++l) {
            T acc = T(0);
#pragma unroll
            for (int n = 0; n < B; ++n) acc += f[n] * KXI(n, l);
            pp[l] = acc;
          }
        };
        if constexpr (TAB) {
#pragma unroll
          for (int r = 0; r < PR; ++r)
            if (s + r * O::NT < NITEMS) item(sc[pc[r] + t * B2], po[r]);
        } else {
          for (int i = s; i < NITEMS; i += O::NT) {
            const int dv = i / B2, ab = i - dv * B2, d = dv / V;
            item(sc[(d * B + t) * B2 + ab], dv * BLK + ab * PS);
          }
        }
        __syncthreads();
        // phase A
        if (act) {
          const T c0 = w3 * F0(t);
          const T cs = grav != 0.0 ? (((W(t) * W(z)) * W(y)) * W(x)) * (T)dt : T(0);  // source weight
#pragma unroll
          for (int v = 0; v < V; ++v) {
            T r = ((c0 * uu[v] - sP[(0 * V + v) * BLK + ox]) - sP[(1 * V + v) * BLK + oy]) - sP[(2 * V + v) * BLK + oz];
            if (v >= 3 && grav != 0.0) r += cs * gravity_source<T>(q[t], v, (T)grav);
            const T ri = iw3 * r;
#pragma unroll
            for (int t2 = 0; t2 < B; ++t2) qn[t2][v] += ri * IK1(t2, t);
          }
        }
      }
    } else {
      // two slices per phase, contraction in place (an item reads its whole
      // pencil before writing it): slice t0 in sF, slice t0+1 in sP (which
      // follows sF directly); the last slice of an odd B goes alone.  Three
      // barriers per pair.  Item i of a pair is pencil i/B2 (slice-major,
      // then direction and variable) at i%B2; its LDS offset and pencil
      // constant index are the same for every pair and sweep (po, pc).
#pragma unroll
      for (int t0 = 0; t0 < B; t0 += 2) {
        asm volatile("" : "+s"(tz));
        const int ns = t0 + 1 < B ? 2 : 1;
        if (act) {
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            if (k < ns) {
              T* buf = k == 0 ? sF : sP;
              T F[3][V];
              EulerPDE::flux<T>(q[t0 + k], F, tg);
#pragma unroll
              for (int v = 0; v < V; ++v) {
                buf[(0 * V + v) * BLK + ox] = F[0][v];
                buf[(1 * V + v) * BLK + oy] = F[1][v];
                buf[(2 * V + v) * BLK + oz] = F[2][v];
              }
            }
          }
        }
        __syncthreads();
        if constexpr (PTAB) {
#pragma unroll
          for (int r = 0; r < PR; ++r) {
            if (s + r * O::NT < ns * NITEMS) {
              const T c = sc[pc[r] + t0 * B2];
              T* pf = sF + po[r];
              T f[B];
#pragma unroll
              for (int n = 0; n < B; ++n) f[n] = c * pf[n];
#pragma unroll
              for (int l = 0; l < B; ++l) {
                T acc = T(0);
#pragma unroll
                for (int n = 0; n < B; ++n) acc += f[n] * KXI(n, l);
                pf[l] = acc;
              }
            }
          }
        } else {
          for (int i = s; i < ns * NITEMS; i += O::NT) {
            const int k = i >= NITEMS ? 1 : 0;
            const int ii = i - k * NITEMS;
            const int dv = ii / B2, ab = ii - dv * B2, d = dv / V;
            const T c = sc[(d * B + t0 + k) * B2 + ab];
            T* pf = (k == 0 ? sF : sP) + dv * BLK + ab * PS;
            T f[B];
#pragma unroll
            for (int n = 0; n < B; ++n) f[n] = c * pf[n];
#pragma unroll
            for (int l = 0; l < B; ++l) {
              T acc = T(0);
#pragma unroll
              for (int n = 0; n < B; ++n) acc += f[n] * KXI(n, l);
              pf[l] = acc;
            }
          }
        }
        __syncthreads();
        if (act) {
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            if (k < ns) {
              const T* buf = k == 0 ? sF : sP;
              const T c0 = w3 * F0(t0 + k);
              const T cs = grav != 0.0 ? (((W(t0 + k) * W(z)) * W(y)) * W(x)) * (T)dt : T(0);
#pragma unroll
              for (int v = 0; v < V; ++v) {
                T r =
                    ((c0 * uu[v] - buf[(0 * V + v) * BLK + ox]) - buf[(1 * V + v) * BLK + oy]) - buf[(2 * V + v) * BLK + oz];
                if (v >= 3 && grav != 0.0) r += cs * gravity_source<T>(q[t0 + k], v, (T)grav);
                const T ri = iw3 * r;
#pragma unroll
                for (int t2 = 0; t2 < B; ++t2) qn[t2][v] += ri * IK1(t2, t0 + k);
              }
            }
          }
        }
        __syncthreads();
      }
    }
    double res = 0.0;
    if (act) {
      T rt = T(0);
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const T dd = q[t][v] - qn[t][v];
          rt += dd * dd;
        }
      res = (double)rt;
    }
    res = block_sum(res, red);
    const bool done = res < tol2 || it == max_sweeps;
    if (done && ACC_FH && act) {
      // time-averaged fluxes of this last sweep (aderPredictorNonlinear
      // :326-397): the fluxes of its input iterate q, summed over t in order
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < V; ++v) fh[d][v] = 0.0;
      if (sh) sh[0] = sh[1] = 0.0;
#pragma unroll
      for (int t = 0; t < B; ++t) {
        T F[3][V];
        EulerPDE::flux<T>(q[t], F, tg);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int v = 0; v < V; ++v) {
            fh[d][v] += (double)F[d][v] * Td.w[t];
            if (DEBUG) dbg.lFi[((t * B3 + s) * 4 + d) * V + v] = (double)F[d][v];
          }
        // the time-averaged source of the same iterate (lShi, components 3, 4)
        if (sh && grav != 0.0) {
          sh[0] += (double)gravity_source<T>(q[t], 3, (T)grav) * Td.w[t];
          sh[1] += (double)gravity_source<T>(q[t], 4, (T)grav) * Td.w[t];
        }
        if (DEBUG)
#pragma unroll
          for (int v = 0; v < V; ++v)
            dbg.lFi[((t * B3 + s) * 4 + 3) * V + v] = grav != 0.0 ? (double)gravity_source<T>(q[t], v, (T)grav) : 0.0;
      }
    }
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) q[t][v] = qn[t][v];
    if (done) break;
  }
  return it;
}

// ---------------------------------------------------------------------------
// The same Picard sweeps with one barrier per time slice ("node-owner"
// contraction).  Per slice t:
//   F  node owners evaluate the flux, scale F_d by the pencil constant
//      c_d = ((w_t w_a) w_b) dt/dx_d of their own pencil (a table in LDS) and
//      store c_d F_d in node order [z][y][x] into buffer t&1;
//   C  after the barrier every node owner contracts its three pencils itself:
//      P_d = sum_n (c_d F_d)[pencil, n] Kxi[n][l_d] -- lanes of one pencil read
//      the same LDS words (broadcast), x-neighbours read consecutive words --
//      and accumulates the time contraction as before.
// The products and sums are the ones the pencil-owner form evaluates, in the
// same order.  Buffers alternate between slices, so slice t+1's stores never
// meet slice t's loads and a single barrier per slice suffices.
// sF: 2 buffers of 3*V*B^3; sc: c_d table [d][t][a*B+b].
template <int N, typename T, bool ACC_FH, bool DEBUG>
__device__ inline int picard_sweeps_node(T (&q)[Ord<N>::B][V], const double (&u)[V], T* sF, const T* sc,
                                         double* red, bool act, int s, int z, int y, int x, double gamma,
                                         int max_sweeps, double tol2, double (&fh)[3][V], PredictorDebug dbg,
                                         double grav = 0.0, double dt = 0.0, double* sh = nullptr) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  constexpr bool DP = std::is_same<T, double>::value;
  const DGTables& Td = c_tab[N];
  const DGTablesF& Tf = c_tabf[N];
  // tz is an opaque zero (see the asm in the slice loop): the Kxi and iK1
  // reads through it are re-issued every slice as scalar loads instead of
  // being hoisted out of the sweep into registers the kernel cannot spare
  int tz = 0;
  auto W = [&](int i) -> T { return DP ? (T)Td.w[i] : (T)Tf.w[i]; };
  auto KXI = [&](int a, int b) -> T { return DP ? (T)c_tab[N + tz].Kxi[a][b] : (T)c_tabf[N + tz].Kxi[a][b]; };
  auto IK1 = [&](int a, int b) -> T { return DP ? (T)c_tab[N + tz].iK1[a][b] : (T)c_tabf[N + tz].iK1[a][b]; };
  // (F0 too: w^3 F0[t] u would otherwise be hoisted for every t, 30 registers)
  auto F0 = [&](int i) -> T { return DP ? (T)c_tab[N + tz].F0[i] : (T)c_tabf[N + tz].F0[i]; };

  T uu[V];
#pragma unroll
  for (int v = 0; v < V; ++v) uu[v] = (T)u[v];
  const T w3 = W(z) * W(y) * W(x);
  const T iw3 = T(1.0) / w3;
  const T tg = (T)gamma;
  const int px = (z * B + y) * B, py = z * B2 + x, pz = y * B + x;  // pencil bases
  const int cxo = z * B + y, cyo = z * B + x, czo = y * B + x;      // pencil (a,b) of this node

  int it = 0;
  while (it < max_sweeps) {
    ++it;
    T qn[B][V];
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) qn[t][v] = T(0);
#pragma unroll
    for (int t = 0; t < B; ++t) {
      asm volatile("" : "+s"(tz));
      T* buf = sF + (t & 1) * 3 * V * B3;
      if (act) {
        T F[3][V];
        EulerPDE::flux<T>(q[t], F, tg);
        const T cx = sc[(0 * B + t) * B2 + cxo], cy = sc[(1 * B + t) * B2 + cyo], cz = sc[(2 * B + t) * B2 + czo];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          buf[(0 * V + v) * B3 + s] = cx * F[0][v];
          buf[(1 * V + v) * B3 + s] = cy * F[1][v];
          buf[(2 * V + v) * B3 + s] = cz * F[2][v];
        }
      }
      __syncthreads();
      if (act) {
        const T c0 = w3 * F0(t);
        const T cs = grav != 0.0 ? (((W(t) * W(z)) * W(y)) * W(x)) * (T)dt : T(0);
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const T* bx = buf + (0 * V + v) * B3 + px;
          const T* by = buf + (1 * V + v) * B3 + py;
          const T* bz = buf + (2 * V + v) * B3 + pz;
          T ax = T(0), ay = T(0), az = T(0);
#pragma unroll
          for (int n = 0; n < B; ++n) ax += bx[n] * KXI(n, x);
#pragma unroll
          for (int n = 0; n < B; ++n) ay += by[n * B] * KXI(n, y);
#pragma unroll
          for (int n = 0; n < B; ++n) az += bz[n * B2] * KXI(n, z);
          T r = ((c0 * uu[v] - ax) - ay) - az;
          if (v >= 3 && grav != 0.0) r += cs * gravity_source<T>(q[t], v, (T)grav);
          const T ri = iw3 * r;
#pragma unroll
          for (int t2 = 0; t2 < B; ++t2) qn[t2][v] += ri * IK1(t2, t);
        }
      }
    }
    double res = 0.0;
    if (act) {
      T rt = T(0);
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const T dd = q[t][v] - qn[t][v];
          rt += dd * dd;
        }
      res = (double)rt;
    }
    res = block_sum(res, red);
    const bool done = res < tol2 || it == max_sweeps;
    if (done && ACC_FH && act) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < V; ++v) fh[d][v] = 0.0;
      if (sh) sh[0] = sh[1] = 0.0;
#pragma unroll
      for (int t = 0; t < B; ++t) {
        T F[3][V];
        EulerPDE::flux<T>(q[t], F, tg);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int v = 0; v < V; ++v) {
            fh[d][v] += (double)F[d][v] * Td.w[t];
            if (DEBUG) dbg.lFi[((t * B3 + s) * 4 + d) * V + v] = (double)F[d][v];
          }
        // the time-averaged source of the same iterate (lShi, components 3, 4)
        if (sh && grav != 0.0) {
          sh[0] += (double)gravity_source<T>(q[t], 3, (T)grav) * Td.w[t];
          sh[1] += (double)gravity_source<T>(q[t], 4, (T)grav) * Td.w[t];
        }
        if (DEBUG)
#pragma unroll
          for (int v = 0; v < V; ++v)
            dbg.lFi[((t * B3 + s) * 4 + 3) * V + v] = grav != 0.0 ? (double)gravity_source<T>(q[t], v, (T)grav) : 0.0;
      }
    }
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) q[t][v] = qn[t][v];
    if (done) break;
  }
  return it;
}

// ---------------------------------------------------------------------------
// Low-precision Picard sweeps with the spatial contraction on the matrix
// cores (mixed-precision predictor only).  Per time slice t:
//   F  node owners evaluate the FP32 flux, scale F_d by their pencil's
//      constant c_d and store it in the direction layouts (pencils contiguous);
//   P  the 3*V*B^2 pencils, taken two at a time, are the rows of
//      D = A * Kbd with A[m] = [pencil 2m | pencil 2m+1 | 0] (K = 16) and
//      Kbd = blockdiag(Kxi, Kxi) (16 x 16): one 16x16 MFMA per 32 pencils,
//      BF16 operands (v_mfma_f32_16x16x16_bf16) or FP32 operands
//      (ceil(2B/4) x v_mfma_f32_16x16x4_f32), FP32 accumulation; row m,
//      column n of D is output node n%B of pencil 2m + n/B, so D's valid
//      columns are again contiguous pencils;
//   A  node owners combine rhs and accumulate the time contraction (VALU, FP32).
// Operand layouts of the 16x16 MFMAs (64 lanes, l = lane): A[l%16][4(l/16)+j],
// B[4(l/16)+j][l%16], D[4(l/16)+i][l%16] (x16 variant: j = 0..3; x4 variant:
// A[l%16][l/16], B[l/16][l%16]).

// Kbd entry (k, n) for B <= 8.
template <int B>
__device__ inline float kbd(const float (&K)[10][10], int k, int n) {
  if (k < B && n < B) return K[k][n];
  if (k >= B && k < 2 * B && n >= B && n < 2 * B) return K[k - B][n - B];
  return 0.0f;
}

template <int N, bool BF16>
struct MfmaKxi {
  exa_s4 b16;    // BF16: Kbd[4(l/16)+j][l%16]
  float b32[4];  // FP32: Kbd[4s + l/16][l%16], s = 0..3
  __device__ inline void init(int lane) {
    constexpr int B = N + 1;
    const DGTablesF& Tf = c_tabf[N];
    const int n = lane & 15, kq = lane >> 4;
    if (BF16) {
#pragma unroll
      for (int j = 0; j < 4; ++j) b16[j] = f32_to_bf16_rne(kbd<B>(Tf.Kxi, 4 * kq + j, n));
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) b32[s] = kbd<B>(Tf.Kxi, 4 * s + kq, n);
    }
  }
};

// All pencils of one slice: sF -> sP (both [dv][ab][n], FP32).
template <int N, bool BF16>
__device__ inline void contract_mfma(const float* sF, float* sP, const MfmaKxi<N, BF16>& kf) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  constexpr int NITEMS = 3 * V * B2;
  constexpr int NTILES = (NITEMS + 31) / 32;
  constexpr int NWAVES = O::NT / 64;
  constexpr int KS = (2 * B + 3) / 4;  // FP32 K-steps of 4
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  for (int tile = wave; tile < NTILES; tile += NWAVES) {
    const int row0 = tile * 32 + 2 * m;  // first pencil of this A row
    // entry k of the row is node k%B of pencil row0 + k/B; pencils past the
    // last (NITEMS may be odd) read as zero
    const int lim = (NITEMS - row0) * B;  // entries of this row inside sF
    const float* a = sF + row0 * B;
    exa_f4 acc = {0.f, 0.f, 0.f, 0.f};
    if (BF16) {
      exa_s4 av;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 4 * kq + j;
        av[j] = f32_to_bf16_rne((k < 2 * B && k < lim) ? a[k] : 0.0f);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, kf.b16, acc, 0, 0, 0);
    } else {
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + kq;
        const float av = (k < 2 * B && k < lim) ? a[k] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, kf.b32[s], acc, 0, 0, 0);
      }
    }
    const int n = m;  // output column of this lane: node n%B of pencil 2(row) + n/B
    if (n < 2 * B) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = tile * 32 + 2 * (4 * kq + i);
        if (r * B + n < NITEMS * B) sP[r * B + n] = acc[i];
      }
    }
  }
}

template <int N, bool BF16>
__device__ inline int picard_sweeps_mfma(float (&q)[Ord<N>::B][V], const double (&u)[V], float* sF, float* sP,
                                         const float* sc, double* red, bool act, int s, int z, int y, int x,
                                         double gamma, int max_sweeps, double tol2) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  static_assert(B <= 8, "block-diagonal pairing needs 2B <= 16");
  const DGTablesF& Tf = c_tabf[N];
  MfmaKxi<N, BF16> kf;
  kf.init(threadIdx.x & 63);

  float uu[V];
#pragma unroll
  for (int v = 0; v < V; ++v) uu[v] = (float)u[v];
  const float w3 = Tf.w[z] * Tf.w[y] * Tf.w[x];
  const float iw3 = 1.0f / w3;
  const float tg = (float)gamma;
  const int ox = (z * B + y) * B + x, oy = (z * B + x) * B + y, oz = (y * B + x) * B + z;
  const int abx = z * B + y, aby = z * B + x, abz = y * B + x;

  int tz = 0;  // opaque zero, as in picard_sweeps
  int it = 0;
  while (it < max_sweeps) {
    ++it;
    float qn[B][V];
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) qn[t][v] = 0.0f;
#pragma unroll
    for (int t = 0; t < B; ++t) {
      asm volatile("" : "+s"(tz));
      if (act) {
        float F[3][V];
        EulerPDE::flux<float>(q[t], F, tg);
        const float cx = sc[(0 * B + t) * B2 + abx], cy = sc[(1 * B + t) * B2 + aby], cz = sc[(2 * B + t) * B2 + abz];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          sF[(0 * V + v) * B3 + ox] = cx * F[0][v];
          sF[(1 * V + v) * B3 + oy] = cy * F[1][v];
          sF[(2 * V + v) * B3 + oz] = cz * F[2][v];
        }
      }
      __syncthreads();
      contract_mfma<N, BF16>(sF, sP, kf);
      __syncthreads();
      if (act) {
        const float c0 = w3 * c_tabf[N + tz].F0[t];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const float r = ((c0 * uu[v] - sP[(0 * V + v) * B3 + ox]) - sP[(1 * V + v) * B3 + oy]) - sP[(2 * V + v) * B3 + oz];
          const float ri = iw3 * r;
#pragma unroll
          for (int t2 = 0; t2 < B; ++t2) qn[t2][v] += ri * c_tabf[N + tz].iK1[t2][t];
        }
      }
    }
    double res = 0.0;
    if (act) {
      float rt = 0.0f;
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const float dd = q[t][v] - qn[t][v];
          rt += dd * dd;
        }
      res = (double)rt;
    }
    res = block_sum(res, red);
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) q[t][v] = qn[t][v];
    if (res < tol2) break;
  }
  return it;
}

// Pencil constants c_d(t, a, b) = ((w_t w_a) w_b) dt / dx_d of the sweeps,
// [d][t][a*B+b], in the sweep's arithmetic type.
template <int N, typename T>
__device__ inline void pencil_constants(T* sc, double dt, double dx0, double dx1, double dx2) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  constexpr bool DP = std::is_same<T, double>::value;
  const DGTables& Td = c_tab[N];
  const DGTablesF& Tf = c_tabf[N];
  const T tdx[3] = {(T)dx0, (T)dx1, (T)dx2};
  const T tdt = (T)dt;
  for (int i = threadIdx.x; i < 3 * B * B2; i += blockDim.x) {
    const int d = i / (B * B2), t = (i / B2) % B, a = (i % B2) / B, b = i % B;
    const T wt = DP ? (T)Td.w[t] : (T)Tf.w[t];
    const T wa = DP ? (T)Td.w[a] : (T)Tf.w[a];
    const T wb = DP ? (T)Td.w[b] : (T)Tf.w[b];
    sc[i] = ((wt * wa) * wb) * tdt / tdx[d];
  }
}

// Arithmetic of the predictor's Picard sweeps (kernel template argument).
enum : int { PRED_FP64 = 0, PRED_FP32 = 1, PRED_FP32_MFMA = 2, PRED_BF16_MFMA = 3 };

// ---------------------------------------------------------------------------
// Space-time predictor + time averages + face extrapolation + volume
// integral.  spaceTimePredictorNonlinear.cpph:43-577 and
// volumeIntegralNonlinear.cpph:26-125 fused into one pass over the cell.
// MIXED=false: the plain FP64 method (Picard to sum(dq^2) < 1e-14 or 2(N+1)
// sweeps, as the reference).  MIXED=true: FP32 sweeps until the residual
// reaches the FP32 rounding floor of the cell, then `fp64_sweeps` FP64
// correction sweeps (stopping early on the FP64 criterion); the averages,
// extrapolation and volume integral are FP64 from the last FP64 sweep.
// Contraction form of the sweeps (kernel template argument).
enum : int { FORM_PENCIL = 0, FORM_NODE = 1, FORM_PENCIL_PAIR = 2 };

// SRC: the PDE has the gravity source (Euler with gravity, FP64 only); the
// source code is compiled out of the plain Euler kernels.
// USKIP (FP32 sweeps only): a cell whose nodes differ from one state by less
// than FP32 resolves (4 eps |u|max) gives the FP32 sweeps nothing to do and
// takes the FP64 sweeps alone, exactly as the FP64 predictor (a uniform cell
// starts at the fixed point and stops after one sweep).  A separate variant
// because the check's registers cost the plain mixed kernel ~1 %; the
// adaptive mesh, mostly uniform away from its features, uses it.
template <int N, bool DEBUG, int PREC, int WPE, int FORM, bool SRC = false, bool USKIP = false>
__global__ __launch_bounds__(Ord<N>::NT) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_predictor(
    const double* __restrict__ luh, double* __restrict__ lduh, double* __restrict__ qf,
    double* __restrict__ ff, int* __restrict__ iters, const double* __restrict__ dtp, double dx0,
    double dx1, double dx2, double gamma, int fp64_sweeps, unsigned long long* __restrict__ sweep_count,
    PredictorDebug dbg, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  extern __shared__ double sm[];
  constexpr int BLK = O::BLK;
  double* scd = sm + 6 * V * BLK;                         // pencil constants, FP64 sweeps
  float* scf = reinterpret_cast<float*>(sm + 3 * V * BLK);  // pencil constants, FP32 sweeps
  double* red = sm + 6 * V * BLK + 3 * B3;                 // reduction scratch
  const DGTables& T = c_tab[N];

  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  const bool act = s < B3;
  const int z = act ? s / B2 : 0, y = act ? (s / B) % B : 0, x = act ? s % B : 0;
  constexpr bool MIXED = PREC != PRED_FP64;
  double dt;
  if (ens.cell_inst) {
    const InstParams& P = ens.inst[ens.cell_inst[cell]];
    dt = P.dt_cur;
    dx0 = P.dx[0];
    dx1 = P.dx[1];
    dx2 = P.dx[2];
    gamma = P.gamma;
  } else {
    dt = *dtp;
  }
  pencil_constants<N, double>(scd, dt, dx0, dx1, dx2);
  if (MIXED) pencil_constants<N, float>(scf, dt, dx0, dx1, dx2);
  __syncthreads();

  double u[V];
#pragma unroll
  for (int v = 0; v < V; ++v) u[v] = act ? luh[cell * V * B3 + v * B3 + s] : 1.0;

  double q[B][V];
  double fh[3][V];
  double sh[2] = {0.0, 0.0};  // time-averaged source, components 3 and 4 (Euler with gravity)
  const double grav = SRC ? ens.grav : 0.0;
  constexpr int MaxIterations = 2 * (N + 1);
  int it = 0;
  if constexpr (MIXED) {
    static_assert(!USKIP || PREC == PRED_FP32, "the uniform-cell variant is built for FP32 sweeps");
    // rounding floor of sum((dq)^2) over the cell in the low precision:
    // count * (4 eps * |u|max)^2, eps the unit roundoff of the contraction
    // operands (FP32 2^-24, BF16 2^-8)
    double um = 0.0, dev = 0.0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      um = fmax(um, fabs(u[v]));
      if constexpr (USKIP)
        if (act) dev = fmax(dev, fabs(u[v] - luh[cell * V * B3 + v * B3]));
    }
    um = -block_min(-um, red);
    const double eps = PREC == PRED_BF16_MFMA ? 3.90625e-3 : 5.96e-8;
    bool uniform = false;
    if constexpr (USKIP) uniform = -block_min(-dev, red) <= 4.0 * eps * um;
    int max64 = fp64_sweeps;
    if (uniform) {
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) q[t][v] = u[v];
      max64 = MaxIterations;
    } else {
      float qs[B][V];
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) qs[t][v] = (float)u[v];
      const double floor_lp = (double)(V * O::B4) * (4.0 * eps * um) * (4.0 * eps * um);
      const double tol_lp = fmax(1e-7 * 1e-7, floor_lp);
      float* sFf = reinterpret_cast<float*>(sm);
      if constexpr (PREC == PRED_FP32_MFMA || PREC == PRED_BF16_MFMA)
        it = picard_sweeps_mfma<N, PREC == PRED_BF16_MFMA>(qs, u, sFf, sFf + 3 * V * B3, scf, red, act, s, z, y, x,
                                                           gamma, MaxIterations, tol_lp);
      else if constexpr (FORM == FORM_NODE)
        it = picard_sweeps_node<N, float, false, false>(qs, u, sFf, scf, red, act, s, z, y, x, gamma, MaxIterations,
                                                        tol_lp, fh, dbg);
      else
        it = picard_sweeps<N, float, false, false, FORM == FORM_PENCIL_PAIR>(qs, u, sFf, sFf + 3 * V * BLK, scf, red,
                                                                             act, s, z, y, x, gamma, MaxIterations,
                                                                             tol_lp, fh, dbg);
#pragma unroll
      for (int t = 0; t < B; ++t)
#pragma unroll
        for (int v = 0; v < V; ++v) q[t][v] = (double)qs[t][v];
    }
    if constexpr (FORM == FORM_NODE)
      it += picard_sweeps_node<N, double, true, DEBUG>(q, u, sm, scd, red, act, s, z, y, x, gamma, max64,
                                                       1e-7 * 1e-7, fh, dbg, grav, dt, sh);
    else
      it += picard_sweeps<N, double, true, DEBUG, FORM == FORM_PENCIL_PAIR>(q, u, sm, sm + 3 * V * BLK, scd, red, act,
                                                                            s, z, y, x, gamma, max64,
                                                                            1e-7 * 1e-7, fh, dbg, grav, dt, sh);
  } else {
#pragma unroll
    for (int t = 0; t < B; ++t)
#pragma unroll
      for (int v = 0; v < V; ++v) q[t][v] = u[v];
    if constexpr (FORM == FORM_NODE)
      it = picard_sweeps_node<N, double, true, DEBUG>(q, u, sm, scd, red, act, s, z, y, x, gamma, MaxIterations,
                                                      1e-7 * 1e-7, fh, dbg, grav, dt, sh);
    else
      it = picard_sweeps<N, double, true, DEBUG, FORM == FORM_PENCIL_PAIR>(q, u, sm, sm + 3 * V * BLK, scd, red, act, s,
                                                                           z, y, x, gamma, MaxIterations, 1e-7 * 1e-7,
                                                                           fh, dbg, grav, dt, sh);
  }

  // time averages (aderPredictorNonlinear :326-397)
  double* sQh = sm;
  double* sFh = sm + V * B3;
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
      for (int v = 0; v < V; ++v) sFh[(d * V + v) * B3 + s] = fh[d][v];
    if (DEBUG) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        dbg.lFhi[0 * B3 * V + ((z * B + y) * B + x) * V + v] = fh[0][v];
        dbg.lFhi[1 * B3 * V + ((z * B + x) * B + y) * V + v] = fh[1][v];
        dbg.lFhi[2 * B3 * V + ((y * B + x) * B + z) * V + v] = fh[2][v];
        dbg.lFhi[3 * B3 * V + s * V + v] = v >= 3 ? sh[v - 3] : 0.0;
      }
    }
  }
  __syncthreads();
  if (act) {
    double out[V];
    volume_integral_node<N>(sFh, s, z, y, x, dx0, dx1, dx2, out);
    if (SRC && grav != 0.0) {  // source term of the volume integral: += w^3 lShi
      const double weight = T.w[z] * T.w[y] * T.w[x];
      out[3] += weight * sh[0];
      out[4] += weight * sh[1];
    }
    if (ens.fold_update) {
      const double updateSize = dt / (T.w[z] * T.w[y] * T.w[x]);
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] = u[v] + out[v] * updateSize;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) lduh[cell * V * B3 + v * B3 + s] = out[v];
  }
  extrapolate_faces<N>(sQh, sFh, s, qf + cell * 6 * V * B2, ff + cell * 6 * V * B2);
  if (s == 0) {
    if (iters) iters[cell] = it;
    if (sweep_count) atomicAdd(sweep_count, (unsigned long long)it);
  }
}

// ---------------------------------------------------------------------------
// Space-time predictor for the orders whose space-time state does not fit the
// registers of a thread per node (N >= 8: the iterate and its successor alone
// would take 2 B V doubles = 200 VGPRs per thread at N = 9, while B^3 = 1000
// threads per workgroup leave 128).  FP64 only; the same products and sums in
// the same order as k_predictor (spaceTimePredictorNonlinear.cpph:43-577,
// volumeIntegralNonlinear.cpph:26-125).
//
// Persistent grid: workgroup g owns staging slot g, two space-time arrays
// [t][v][node] of B^4 V doubles in global memory (800 KB at N = 9, resident
// in the last-level cache for a grid of one workgroup per CU), and walks the
// cells g, g + gridDim.x, ...  Every slot word is only ever touched by the
// thread of its node, so the slot needs no barriers.  One Picard sweep:
//   for every slice t:  F  node owners load q[t] from the slot, evaluate the
//                          flux and store it in the direction layouts;
//                       P  pencil items contract in place with Kxi, scaled
//                          by the pencil constant (LDS table);
//                       A  node owners store rhs / w^3 into the second array;
//   then every node owner applies iK1 to its own rhs history, takes the
//   residual against q and overwrites the history with the new iterate; the
//   two arrays swap roles between sweeps.
//
// MIX: mixed precision (EXA_PREC_FP32).  The same sweep first runs on FP32
// copies -- iterate and rhs as FP32 arrays in the first half of the slot,
// flux slice and pencil constants as FP32 in the LDS, FP32 tables -- until
// the residual reaches the FP32 rounding floor of the cell (or 2(N+1)
// sweeps); the iterate is then widened into the slot's second array and at
// most fp64_sweeps FP64 sweeps follow (stopping early on the FP64
// criterion), from which the averages, extrapolation and volume integral
// come, as in k_predictor's mixed form.
template <int N, bool DEBUG, bool SRC = false, bool MIX = false>
__global__ __launch_bounds__(Ord<N>::NT) void k_predictor_staged(
    const double* __restrict__ luh, double* __restrict__ lduh, double* __restrict__ qf, double* __restrict__ ff,
    int* __restrict__ iters, const double* __restrict__ dtp, double dx0, double dx1, double dx2, double gamma,
    long long ncells, double* __restrict__ stage, unsigned long long* __restrict__ sweep_count,
    PredictorDebug dbg, Ens ens, int fp64_sweeps) {
  static_assert(!(MIX && SRC), "the gravity source runs in the FP64 predictor only");
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3, PS = O::PS, BLK = O::BLK;
  constexpr int NITEMS = 3 * V * B2;
  constexpr int SQ = B * V * B3;  // doubles of one staged space-time array
  constexpr int ROUNDS = (NITEMS + O::NT - 1) / O::NT;
  constexpr int MaxIterations = 2 * (N + 1);
  extern __shared__ double sm[];
  double* sF = sm;                     // flux slice, direction layouts, contracted in place
  double* sc = sm + 3 * V * BLK;       // pencil constants [d][t][a*B+b]
  double* red = sm;                    // residual sum scratch (between slice loops only)
  double* sQh = sm;                    // after the sweeps: time averages
  double* sFh = sm + V * B3;
  const DGTables& T = c_tab[N];
  const int s = threadIdx.x;
  const bool act = s < B3;
  const int z = act ? s / B2 : 0, y = act ? (s / B) % B : 0, x = act ? s % B : 0;
  const int ox = (z * B + y) * PS + x, oy = (z * B + x) * PS + y, oz = (y * B + x) * PS + z;
  const double w3 = T.w[z] * T.w[y] * T.w[x];
  const double iw3 = 1.0 / w3;
  double dt = ens.cell_inst ? 0.0 : *dtp;
  double* const slot = stage + (size_t)blockIdx.x * 2 * SQ;
  // opaque zero: keeps the Kxi / iK1 reads scalar loads inside the loops
  int tz = 0;
  const double grav = SRC ? ens.grav : 0.0;  // Euler with gravity: algebraic source (0: none)

  for (long long cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    if (ens.cell_inst) {  // per-cell mesh width, time step and gamma
      const InstParams& P = ens.inst[ens.cell_inst[cell]];
      dt = P.dt_cur;
      dx0 = P.dx[0];
      dx1 = P.dx[1];
      dx2 = P.dx[2];
      gamma = P.gamma;
    }
    __syncthreads();  // the previous cell is done with the LDS
    pencil_constants<N, double>(sc, dt, dx0, dx1, dx2);
    double u[V];
#pragma unroll
    for (int v = 0; v < V; ++v) u[v] = act ? luh[cell * V * B3 + v * B3 + s] : 1.0;
    double* q = slot;
    double* r = slot + SQ;
    int it = 0;
    if constexpr (MIX) {
      // ---- FP32 phase ----
      float* qL = reinterpret_cast<float*>(slot);
      float* rL = qL + SQ;
      float* sFf = reinterpret_cast<float*>(sm);
      float* scf = sFf + 3 * V * BLK;  // inside the flux slice's second half: the FP64 constants stay
      __syncthreads();                 // the FP64 constants are written (pencil_constants above)
      pencil_constants<N, float>(scf, dt, dx0, dx1, dx2);
      const float gf = (float)gamma;
      float uf[V];
      double um = 0.0;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        uf[v] = (float)u[v];
        um = fmax(um, act ? fabs(u[v]) : 0.0);
      }
      um = -block_min(-um, red);
      // rounding floor of sum(dq^2) over the cell in FP32: count * (4 eps |u|max)^2
      const double floor_lp = (double)(V * O::B4) * (4.0 * 5.96e-8 * um) * (4.0 * 5.96e-8 * um);
      const double tol_lp = fmax(1e-7 * 1e-7, floor_lp);
      const float w3f = (float)w3, iw3f = (float)iw3;
#pragma unroll 1
      for (int t = 0; t < B; ++t) {
        asm volatile("" : "+s"(tz));
        const int sl = s + tz;
        if (act)
#pragma unroll
          for (int v = 0; v < V; ++v) qL[(t * V + v) * B3 + sl] = uf[v];
      }
      __syncthreads();
      bool done_lp = false;
      while (!done_lp) {
        ++it;
#pragma unroll 1
        for (int t = 0; t < B; ++t) {
          asm volatile("" : "+s"(tz));
          const int sl = s + tz;
          if (act) {
            float Q[V], F[3][V];
#pragma unroll
            for (int v = 0; v < V; ++v) Q[v] = qL[(t * V + v) * B3 + sl];
            EulerPDE::flux<float>(Q, F, gf);
#pragma unroll
            for (int v = 0; v < V; ++v) {
              sFf[(0 * V + v) * BLK + ox] = F[0][v];
              sFf[(1 * V + v) * BLK + oy] = F[1][v];
              sFf[(2 * V + v) * BLK + oz] = F[2][v];
            }
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < ROUNDS; ++k) {
            asm volatile("" : "+s"(tz));
            const int i = s + k * O::NT;
            if (i < NITEMS) {
              const int dv = i / B2, ab = i - dv * B2, d = dv / V;
              const float c = scf[(d * B + t) * B2 + ab];
              float* pf = sFf + dv * BLK + ab * PS;
              float f[B];
#pragma unroll
              for (int n = 0; n < B; ++n) f[n] = c * pf[n];
#pragma unroll
              for (int l = 0; l < B; ++l) {
                float acc = 0.0f;
#pragma unroll
                for (int n = 0; n < B; ++n) acc += f[n] * c_tabf[N + tz].Kxi[n][l];
                pf[l] = acc;
              }
            }
          }
          __syncthreads();
          if (act) {
            const float c0 = w3f * c_tabf[N + tz].F0[t];
#pragma unroll
            for (int v = 0; v < V; ++v) {
              const float rr = ((c0 * uf[v] - sFf[(0 * V + v) * BLK + ox]) - sFf[(1 * V + v) * BLK + oy]) -
                               sFf[(2 * V + v) * BLK + oz];
              rL[(t * V + v) * B3 + sl] = iw3f * rr;
            }
          }
          __syncthreads();
        }
        double res = 0.0;
        const int sa = act ? s : 0;
#pragma unroll 1
        for (int v = 0; v < V; ++v) {
          asm volatile("" : "+s"(tz));
          const int sv = sa + tz;
          float ri[B];
#pragma unroll
          for (int t = 0; t < B; ++t) ri[t] = rL[(t * V + v) * B3 + sv];
#pragma unroll
          for (int t2 = 0; t2 < B; ++t2) {
            asm volatile("" : "+s"(tz));
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < B; ++t) acc += ri[t] * c_tabf[N + tz].iK1[t2][t];
            const double dd = (double)(qL[(t2 * V + v) * B3 + sv] - acc);
            if (act) {
              res += dd * dd;
              rL[(t2 * V + v) * B3 + sv] = acc;
            }
          }
        }
        res = block_sum(res, red);
        __syncthreads();
        done_lp = __builtin_amdgcn_readfirstlane((res < tol_lp || it == MaxIterations) ? 1 : 0) != 0;
        if (!done_lp) {
          float* tmp = qL;
          qL = rL;
          rL = tmp;
        }
      }
      // the FP32 iterate (rL, in the slot's first half) widened into the
      // second FP64 array, which becomes the input of the first FP64 sweep;
      // every word is read and written by its own node's thread
#pragma unroll 1
      for (int t = 0; t < B; ++t) {
        asm volatile("" : "+s"(tz));
        const int sl = s + tz;
        if (act)
#pragma unroll
          for (int v = 0; v < V; ++v) r[(t * V + v) * B3 + sl] = (double)rL[(t * V + v) * B3 + sl];
      }
      double* tmp = q;
      q = r;
      r = tmp;
      __syncthreads();  // the flux slice is free again
    } else {
#pragma unroll 1
      for (int t = 0; t < B; ++t) {
        asm volatile("" : "+s"(tz));
        const int sl = s + tz;
        if (act)
#pragma unroll
          for (int v = 0; v < V; ++v) q[(t * V + v) * B3 + sl] = u[v];
      }
      __syncthreads();
    }
    const int max64 = MIX ? it + fp64_sweeps : MaxIterations;
    bool done = false;
    while (!done) {
      ++it;
#pragma unroll 1
      for (int t = 0; t < B; ++t) {
        asm volatile("" : "+s"(tz));
        const int sl = s + tz;
        if (act) {
          double Q[V], F[3][V];
#pragma unroll
          for (int v = 0; v < V; ++v) Q[v] = q[(t * V + v) * B3 + sl];
          EulerPDE::flux<double>(Q, F, gamma);
#pragma unroll
          for (int v = 0; v < V; ++v) {
            sF[(0 * V + v) * BLK + ox] = F[0][v];
            sF[(1 * V + v) * BLK + oy] = F[1][v];
            sF[(2 * V + v) * BLK + oz] = F[2][v];
          }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ROUNDS; ++k) {
          asm volatile("" : "+s"(tz));
          const int i = s + k * O::NT;
          if (i < NITEMS) {
            const int dv = i / B2, ab = i - dv * B2, d = dv / V;
            const double c = sc[(d * B + t) * B2 + ab];
            double* pf = sF + dv * BLK + ab * PS;
            double f[B];
#pragma unroll
            for (int n = 0; n < B; ++n) f[n] = c * pf[n];
#pragma unroll
            for (int l = 0; l < B; ++l) {
              double acc = 0.0;
#pragma unroll
              for (int n = 0; n < B; ++n) acc += f[n] * c_tab[N + tz].Kxi[n][l];
              pf[l] = acc;
            }
          }
        }
        __syncthreads();
        if (act) {
          const double c0 = w3 * c_tab[N + tz].F0[t];
          double cs = 0.0, src[V] = {0.0, 0.0, 0.0, 0.0, 0.0};
          if (grav != 0.0) {  // rhs += ((((w_t w_z) w_y) w_x) dt) S(q[t])
            double Qs[V] = {q[(t * V + 0) * B3 + sl], 0.0, 0.0, q[(t * V + 3) * B3 + sl], 0.0};
            cs = (((T.w[t] * T.w[z]) * T.w[y]) * T.w[x]) * dt;
            src[3] = gravity_source<double>(Qs, 3, grav);
            src[4] = gravity_source<double>(Qs, 4, grav);
          }
#pragma unroll
          for (int v = 0; v < V; ++v) {
            double rr =
                ((c0 * u[v] - sF[(0 * V + v) * BLK + ox]) - sF[(1 * V + v) * BLK + oy]) - sF[(2 * V + v) * BLK + oz];
            if (v >= 3 && grav != 0.0) rr += cs * src[v];
            r[(t * V + v) * B3 + sl] = iw3 * rr;
          }
        }
        __syncthreads();
      }
      // (threads past the last node read node 0 and store nothing)
      double res = 0.0;
      const int sa = act ? s : 0;
#pragma unroll 1
      for (int v = 0; v < V; ++v) {
        asm volatile("" : "+s"(tz));
        const int sv = sa + tz;
        double ri[B];
#pragma unroll
        for (int t = 0; t < B; ++t) ri[t] = r[(t * V + v) * B3 + sv];
#pragma unroll
        for (int t2 = 0; t2 < B; ++t2) {
          asm volatile("" : "+s"(tz));
          double acc = 0.0;
#pragma unroll
          for (int t = 0; t < B; ++t) acc += ri[t] * c_tab[N + tz].iK1[t2][t];
          const double dd = q[(t2 * V + v) * B3 + sv] - acc;
          if (act) {
            res += dd * dd;
            r[(t2 * V + v) * B3 + sv] = acc;
          }
        }
      }
      res = block_sum(res, red);
      __syncthreads();  // every thread has read red before the flux slice reuses it
      // block-uniform (every thread holds the same sum); readfirstlane makes
      // that visible to the compiler, which keeps the loops scalar
      done = __builtin_amdgcn_readfirstlane((res < 1e-7 * 1e-7 || it == max64) ? 1 : 0) != 0;
      if (!done) {
        double* tmp = q;
        q = r;
        r = tmp;
      }
    }
    // q: the input iterate of the last sweep (its fluxes are the time-averaged
    // fluxes, aderPredictorNonlinear :326-397); r: the final iterate
    double fh[3][V];
    double sh[2] = {0.0, 0.0};  // time-averaged source, components 3 and 4
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v = 0; v < V; ++v) fh[d][v] = 0.0;
#pragma unroll 1
    for (int t = 0; t < B; ++t) {
      asm volatile("" : "+s"(tz));
      const int sl = s + tz;
      if (act) {
        double Q[V], F[3][V];
#pragma unroll
        for (int v = 0; v < V; ++v) Q[v] = q[(t * V + v) * B3 + sl];
        EulerPDE::flux<double>(Q, F, gamma);
        const double wt = T.w[t];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int v = 0; v < V; ++v) {
            fh[d][v] += F[d][v] * wt;
            if (DEBUG) dbg.lFi[((t * B3 + s) * 4 + d) * V + v] = F[d][v];
          }
        if (grav != 0.0) {
          sh[0] += gravity_source<double>(Q, 3, grav) * wt;
          sh[1] += gravity_source<double>(Q, 4, grav) * wt;
        }
        if (DEBUG)
#pragma unroll
          for (int v = 0; v < V; ++v)
            dbg.lFi[((t * B3 + s) * 4 + 3) * V + v] = grav != 0.0 ? gravity_source<double>(Q, v, grav) : 0.0;
      }
    }
    if (act) {
#pragma unroll 1
      for (int v = 0; v < V; ++v) {
        double a = 0.0;
#pragma unroll
        for (int t = 0; t < B; ++t) {
          const double qt = r[(t * V + v) * B3 + s];
          a += qt * T.w[t];
          if (DEBUG) dbg.lQi[(s * B + t) * V + v] = qt;
        }
        sQh[v * B3 + s] = a;
        if (DEBUG) dbg.lQhi[s * V + v] = a;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < V; ++v) sFh[(d * V + v) * B3 + s] = fh[d][v];
      if (DEBUG) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          dbg.lFhi[0 * B3 * V + ((z * B + y) * B + x) * V + v] = fh[0][v];
          dbg.lFhi[1 * B3 * V + ((z * B + x) * B + y) * V + v] = fh[1][v];
          dbg.lFhi[2 * B3 * V + ((y * B + x) * B + z) * V + v] = fh[2][v];
          dbg.lFhi[3 * B3 * V + s * V + v] = v >= 3 ? sh[v - 3] : 0.0;
        }
      }
    }
    __syncthreads();
    if (act) {
      double out[V];
      volume_integral_node<N>(sFh, s, z, y, x, dx0, dx1, dx2, out);
      if (grav != 0.0) {  // source term of the volume integral: += w^3 lShi
        const double weight = T.w[z] * T.w[y] * T.w[x];
        out[3] += weight * sh[0];
        out[4] += weight * sh[1];
      }
      if (ens.fold_update) {
        // luh read again through an opaque index, so that the state is not
        // held in registers through the sweeps
        int so = s;
        asm volatile("" : "+v"(so));
        const double updateSize = dt / w3;
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = luh[cell * V * B3 + v * B3 + so] + out[v] * updateSize;
      }
#pragma unroll
      for (int v = 0; v < V; ++v) lduh[cell * V * B3 + v * B3 + s] = out[v];
    }
    extrapolate_faces<N>(sQh, sFh, s, qf + cell * 6 * V * B2, ff + cell * 6 * V * B2);
    if (s == 0) {
      if (iters) iters[cell] = it;
      if (sweep_count) atomicAdd(sweep_count, (unsigned long long)it);
    }
  }
}

// Stand-alone volume integral from lFhi in the reference layout.
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_volume_integral_ref(const double* __restrict__ lFhi,
                                                                    double* __restrict__ lduh, double dx0,
                                                                    double dx1, double dx2, int src,
                                                                    int flux) {
  using O = Ord<N>;
  constexpr int B = O::B, B3 = O::B3;
  extern __shared__ double sm[];
  const int s = threadIdx.x;
  const bool act = s < B3;
  const int z = act ? s / O::B2 : 0, y = act ? (s / B) % B : 0, x = act ? s % B : 0;
  double* sFh = sm + V * B3;
  if (act) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      sFh[(0 * V + v) * B3 + s] = lFhi[0 * B3 * V + ((z * B + y) * B + x) * V + v];
      sFh[(1 * V + v) * B3 + s] = lFhi[1 * B3 * V + ((z * B + x) * B + y) * V + v];
      sFh[(2 * V + v) * B3 + s] = lFhi[2 * B3 * V + ((y * B + x) * B + z) * V + v];
    }
  }
  __syncthreads();
  if (act) {
    double out[V];
    if (flux) {  // without a flux only the source part is integrated (useFlux = false)
      volume_integral_node<N>(sFh, s, z, y, x, dx0, dx1, dx2, out);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] = 0.0;
    }
    if (src) {  // source term: += w^3 lShi (volumeIntegralNonlinear.cpph, useSourceOrNCP branch)
      const double weight = c_tab[N].w[z] * c_tab[N].w[y] * c_tab[N].w[x];
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] += weight * lFhi[3 * B3 * V + s * V + v];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) lduh[v * B3 + s] = out[v];
  }
}

// ---------------------------------------------------------------------------
// One face, one wavefront (lane = threadIdx.x % 64): the Rusanov wave speed
// of a face, the larger maximum eigenvalue of the two face-averaged states
// (riemannSolverNonlinear.cpph:52-83; per-lane partial sums, then a
// wavefront sum).
template <int N>
__device__ inline double riemann_face_speed(const double* __restrict__ QL, const double* __restrict__ QR, int d,
                                            double gamma, int lane) {
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
  const double sL = EulerPDE::max_abs_eigenvalue(aL, d, gamma);
  const double sR = EulerPDE::max_abs_eigenvalue(aR, d, gamma);
  return sL < sR ? sR : sL;
}

// One face, one wavefront: Rusanov flux from the face-averaged wave speeds;
// FL and FR enter holding the extrapolated normal fluxes and both leave
// holding the numerical flux.
template <int N>
__device__ inline void riemann_face(const double* __restrict__ QL, const double* __restrict__ QR,
                                    double* __restrict__ FL, double* __restrict__ FR, int d, double gamma, int lane) {
  constexpr int B2 = Ord<N>::B2;
  constexpr int NPL = (B2 + 63) / 64;
  const double smax = riemann_face_speed<N>(QL, QR, d, gamma, lane);
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int i = v * B2 + node;
        const double F = 0.5 * (FR[i] + FL[i]) - 0.5 * smax * (QR[i] - QL[i]);
        FL[i] = F;
        FR[i] = F;
      }
    }
  }
}

// Rusanov flux on every face of a grid: one wavefront per face.
// riemannSolverNonlinear.cpph:38-120; caller solveRiemannProblemAtInterface
// (E/solvers/ADERDGSolver.cpp:2744-2825): left cell face 2d+1, right cell face 2d.
// nbr[d*ncells + c] = neighbour of cell c in +d direction.
template <int N>
__global__ __launch_bounds__(256) void k_riemann(const double* __restrict__ qf, double* __restrict__ ff,
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
  if (cr < 0) return;  // a domain face of a non-periodic axis (k_boundary_faces)
  if (ens.cell_inst) gamma = ens.inst[ens.cell_inst[cl]].gamma;
  riemann_face<N>(qf + ((size_t)cl * 6 + 2 * d + 1) * V * B2, qf + ((size_t)cr * 6 + 2 * d) * V * B2,
                  ff + ((size_t)cl * 6 + 2 * d + 1) * V * B2, ff + ((size_t)cr * 6 + 2 * d) * V * B2, d, gamma, lane);
}

// Rusanov flux of a list of independent faces (the batched per-face entry):
// face i reads QL = q[2i], QR = q[2i+1] and updates FL = f[2i], FR = f[2i+1]
// ([v][a*B+b] blocks of V B^2 doubles), normal direction dirs[i].
template <int N>
__global__ __launch_bounds__(256) void k_riemann_list(const double* __restrict__ q, double* __restrict__ f,
                                                      const int* __restrict__ dirs, long long nfaces, double gamma) {
  constexpr int B2 = Ord<N>::B2;
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (i >= nfaces) return;
  riemann_face<N>(q + (size_t)(2 * i) * V * B2, q + (size_t)(2 * i + 1) * V * B2, f + (size_t)(2 * i) * V * B2,
                  f + (size_t)(2 * i + 1) * V * B2, dirs[i], gamma, threadIdx.x & 63);
}

// ---------------------------------------------------------------------------
// Boundary values of the Euler solvers at one face node (the user
// boundaryValues, kinds EXA_BC_*): outflow copies the inside state and flux;
// a reflecting wall negates the normal momentum and takes that state's
// normal flux; exact data averages the state of bc.exact and its normal flux
// over [t, t+dt] at the order's Gauss-Legendre time nodes.
template <int N>
__device__ inline void euler_boundary_values(const exa_bc& bc, const double x[3], double t, double dt, int d,
                                             double gamma, const double (&Qin)[V], const double (&Fin)[V],
                                             double (&Qout)[V], double (&Fout)[V]) {
  if (bc.kind == EXA_BC_REFLECTING) {
#pragma unroll
    for (int v = 0; v < V; ++v) Qout[v] = Qin[v];
    Qout[1 + d] = -Qin[1 + d];
    EulerPDE::flux_dir(Qout, d, Fout, gamma);
  } else if (bc.kind == EXA_BC_EXACT) {
    const DGTables& T = c_tab[N];
#pragma unroll
    for (int v = 0; v < V; ++v) Qout[v] = Fout[v] = 0.0;
    for (int i = 0; i <= N; ++i) {
      double Q[V], F[V];
      exa_ic_point(bc.exact, gamma, x, t + T.xi[i] * dt, Q);
      EulerPDE::flux_dir(Q, d, F, gamma);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        Qout[v] += T.w[i] * Q[v];
        Fout[v] += T.w[i] * F[v];
      }
    }
  } else {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      Qout[v] = Qin[v];
      Fout[v] = Fin[v];
    }
  }
}

// The two tangential axes of a face with normal d, faster one first (the
// (a, b) = (slower, faster) node order of the face blocks).
__device__ inline void face_axes(int d, int& d1, int& d2) {
  d1 = d == 0 ? 1 : 0;
  d2 = d == 2 ? 1 : 2;
}

// Domain faces of a grid with non-periodic axes, one wavefront per face:
// boundaryConditions at every face node (boundaryConditions.cpph:24-85),
// then the Riemann problem with the outside state on the outer side
// (applyBoundaryConditions, E/solvers/ADERDGSolver.cpp:2854-2931).
// bfaces[i] = cell * 6 + face; oq, of hold face i's outside state and flux
// ([i][v][B^2]); the time and dt of the step are read from the device.
template <int N>
__global__ __launch_bounds__(256) void k_boundary_faces(double* __restrict__ qf, double* __restrict__ ff,
                                                        const int* __restrict__ bfaces, int nb,
                                                        double* __restrict__ oq, double* __restrict__ of,
                                                        exa_bc bc, double gamma, const double* __restrict__ dt_cur,
                                                        const double* __restrict__ time, double off0, double off1,
                                                        double off2, double dx0, double dx1, double dx2, int nx,
                                                        int ny) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  constexpr int NPL = (B2 + 63) / 64;
  const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= nb) return;
  const DGTables& T = c_tab[N];
  const int code = bfaces[i];
  const long long cell = code / 6;
  const int f = code % 6, d = f >> 1, side = f & 1;
  const int ix = (int)(cell % nx), iy = (int)((cell / nx) % ny), iz = (int)(cell / ((long long)nx * ny));
  const double h[3] = {dx0, dx1, dx2};
  const double centre[3] = {off0 + (ix + 0.5) * dx0, off1 + (iy + 0.5) * dx1, off2 + (iz + 0.5) * dx2};
  int d1, d2;
  face_axes(d, d1, d2);
  double* QIn = qf + ((size_t)cell * 6 + f) * V * B2;
  double* FIn = ff + ((size_t)cell * 6 + f) * V * B2;
  double* QO = oq + (size_t)i * V * B2;
  double* FO = of + (size_t)i * V * B2;
  const double t = *time, dt = *dt_cur;
  double x[3];
  x[d] = centre[d] + (-0.5 + side) * h[d];
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    const int node = lane + 64 * k;
    if (node < B2) {
      x[d2] = centre[d2] + h[d2] * (T.xi[node / B] - 0.5);
      x[d1] = centre[d1] + h[d1] * (T.xi[node % B] - 0.5);
      double qi[V], fi[V], qo[V], fo[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        qi[v] = QIn[v * B2 + node];
        fi[v] = FIn[v * B2 + node];
      }
      euler_boundary_values<N>(bc, x, t, dt, d, gamma, qi, fi, qo, fo);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        QO[v * B2 + node] = qo[v];
        FO[v * B2 + node] = fo[v];
      }
    }
  }
  // each lane reads back only the nodes it wrote
  if (side == 0)
    riemann_face<N>(QO, QIn, FO, FIn, d, gamma, lane);
  else
    riemann_face<N>(QIn, QO, FIn, FO, d, gamma, lane);
}

// boundaryConditions of one face in the reference layout (the per-face
// entry): node-major QIn [B^2][nd], FIn [B^2][V] in, QOut, FOut out.
template <int N>
__global__ __launch_bounds__(64) void k_boundary_conditions_ref(double* __restrict__ FOut, double* __restrict__ QOut,
                                                                const double* __restrict__ FIn,
                                                                const double* __restrict__ QIn, exa_bc bc,
                                                                double gamma, double c0, double c1, double c2,
                                                                double s0, double s1, double s2, double t,
                                                                double dt, int faceIndex, int d, int nd) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2;
  const DGTables& T = c_tab[N];
  const double centre[3] = {c0, c1, c2}, size[3] = {s0, s1, s2};
  const int side = faceIndex - 2 * d;
  int d1, d2;
  face_axes(d, d1, d2);
  for (int node = threadIdx.x; node < B2; node += blockDim.x) {
    double x[3];
    x[d] = centre[d] + (-0.5 + side) * size[d];
    x[d2] = centre[d2] + size[d2] * (T.xi[node / B] - 0.5);
    x[d1] = centre[d1] + size[d1] * (T.xi[node % B] - 0.5);
    double qi[V], fi[V], qo[V], fo[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      qi[v] = QIn[(size_t)node * nd + v];
      fi[v] = FIn[(size_t)node * V + v];
    }
    euler_boundary_values<N>(bc, x, t, dt, d, gamma, qi, fi, qo, fo);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      QOut[(size_t)node * nd + v] = qo[v];
      FOut[(size_t)node * V + v] = fo[v];
    }
    for (int v = V; v < nd; ++v) QOut[(size_t)node * nd + v] = QIn[(size_t)node * nd + v];
  }
}

// ---------------------------------------------------------------------------
// Surface integral + solution update + per-node CFL dt with a block minimum
// and one atomic per cell.  surfaceIntegralNonlinear.cpph:21-86,
// solutionUpdate.cpph:23-47, stableTimeStepSize.cpph:21-61.
// UPD_LINEAR: the surface integral of the linear scheme, whose Riemann
// solver returns fluctuations: both faces enter with a plus sign
// (surfaceIntegralLinear.cpph:24-90).
enum : int { UPD_SURFACE = 1, UPD_UPDATE = 2, UPD_DT = 4, UPD_WRITE_LDUH = 8, UPD_LINEAR = 16 };

template <int N, int MODE, class PDE = PdeEuler>
__global__ __launch_bounds__(Ord<N>::NT) void k_update(double* __restrict__ luh, double* __restrict__ lduh,
                                                       const double* __restrict__ ff,
                                                       const double* __restrict__ dtp,
                                                       unsigned long long* __restrict__ dt_next, double dx0,
                                                       double dx1, double dx2, double gamma, double cfl, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  __shared__ double red[16];
  const DGTables& T = c_tab[N];
  const size_t cell = blockIdx.x;
  if (ens.cell_inst) {
    InstParams& P = ens.inst[ens.cell_inst[cell]];
    dtp = &P.dt_cur;
    dt_next = &P.dt_next;
    dx0 = P.dx[0];
    dx1 = P.dx[1];
    dx2 = P.dx[2];
    gamma = P.gamma;
  }
  const int s = threadIdx.x;
  const bool act = s < B3;
  const int z = act ? s / B2 : 0, y = act ? (s / B) % B : 0, x = act ? s % B : 0;
  double dtn = 1.0e300;
  if (act) {
    double du[V], u[V];
    // both state loads issued before any arithmetic
    if (MODE & (UPD_UPDATE | UPD_DT)) {
#pragma unroll
      for (int v = 0; v < V; ++v) u[v] = luh[cell * V * B3 + v * B3 + s];
    }
    if (MODE & (UPD_SURFACE | UPD_UPDATE)) {
#pragma unroll
      for (int v = 0; v < V; ++v) du[v] = lduh[cell * V * B3 + v * B3 + s];
    }
    if (MODE & UPD_SURFACE) {
      const double* F = ff + cell * 6 * V * B2;
      const double usx = (T.w[z] * T.w[y]) / dx0;
#pragma unroll
      for (int v = 0; v < V; ++v)
        du[v] -= (MODE & UPD_LINEAR)
                     ? (F[(1 * V + v) * B2 + z * B + y] * T.FR[x] + F[(0 * V + v) * B2 + z * B + y] * T.FL[x]) * usx
                     : (F[(1 * V + v) * B2 + z * B + y] * T.FR[x] - F[(0 * V + v) * B2 + z * B + y] * T.FL[x]) * usx;
      const double usy = (T.w[z] * T.w[x]) / dx1;
#pragma unroll
      for (int v = 0; v < V; ++v)
        du[v] -= (MODE & UPD_LINEAR)
                     ? (F[(3 * V + v) * B2 + z * B + x] * T.FR[y] + F[(2 * V + v) * B2 + z * B + x] * T.FL[y]) * usy
                     : (F[(3 * V + v) * B2 + z * B + x] * T.FR[y] - F[(2 * V + v) * B2 + z * B + x] * T.FL[y]) * usy;
      const double usz = (T.w[y] * T.w[x]) / dx2;
#pragma unroll
      for (int v = 0; v < V; ++v)
        du[v] -= (MODE & UPD_LINEAR)
                     ? (F[(5 * V + v) * B2 + y * B + x] * T.FR[z] + F[(4 * V + v) * B2 + y * B + x] * T.FL[z]) * usz
                     : (F[(5 * V + v) * B2 + y * B + x] * T.FR[z] - F[(4 * V + v) * B2 + y * B + x] * T.FL[z]) * usz;
    }
    if (MODE & UPD_WRITE_LDUH) {
#pragma unroll
      for (int v = 0; v < V; ++v) lduh[cell * V * B3 + v * B3 + s] = du[v];
    }
    if (MODE & UPD_UPDATE) {
      const double weight = T.w[z] * T.w[y] * T.w[x];
      const double updateSize = *dtp / weight;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        u[v] += du[v] * updateSize;
        luh[cell * V * B3 + v * B3 + s] = u[v];
      }
    }
    if (MODE & UPD_DT) {
      double m[3];
      PDE::max_abs_eigenvalues3(u, PdeParams{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}}, m);
      double den = 0.0;
      den += m[0] / dx0;
      den += m[1] / dx1;
      den += m[2] / dx2;
      dtn = exa_dt_sane(cfl * exa_pnpm(N) / den);
    }
  }
  if (MODE & UPD_DT) {
    dtn = block_min(dtn, red);
    if (s == 0) atomic_min_pos_double(dt_next, dtn);
  }
}

// Riemann problems of a cell's six faces, surface integral, solution update
// and CFL minimum in one launch (one workgroup per cell), after a predictor
// that folded the volume part of the update into lduh (Ens::fold_update:
// lduh = luh + dt / w^3 * volume integral; the new luh is lduh - dt / w^3 *
// surface integral, so luh is not read).  Face f = 2d + side
// belongs to half-wavefront f (32 lanes, nodes lane, lane + 32, ...): its
// Rusanov flux from this cell's extrapolated state and flux on f and the
// neighbour's on the opposite face, into LDS, the six faces concurrently.
// Every interior face is thus solved by both of its cells, with the same
// operations on the same inputs, so both see the same flux, and the face
// arrays are only read.  A domain face (no neighbour) takes the flux the
// boundary kernel left in ff.  The surface integral, update and time step
// are then k_update's (surfaceIntegralNonlinear.cpph:21-86,
// solutionUpdate.cpph:23-47, stableTimeStepSize.cpph:21-61).
// Neighbours: nx > 0 -- a box of nx * ny * nz cells in x-fastest order,
// wrapped along the axes whose bit is set in periodic, none past the ends of
// the others (computed, so that every load of the cell is issued at once);
// nx = 0 -- the table nbr[d * nc + c] (+d), nbr[(3 + d) * nc + c] (-d), -1
// for none.
__device__ inline double half_wave_sum(double v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int N>
struct RiemannUpdate {
  // at most 256 threads per cell: the faces take six half-wavefronts, and the
  // nodes of the update loop over the threads, so that high orders keep
  // several cells in flight per CU
  static constexpr int NT = Ord<N>::NT > 256 ? 256 : Ord<N>::NT;
  static constexpr int NIT = (Ord<N>::B3 + NT - 1) / NT;  // update nodes per thread
};

template <int N>
__global__ __launch_bounds__(RiemannUpdate<N>::NT) void k_riemann_update(double* __restrict__ luh,
                                                                         const double* __restrict__ lduh,
                                                                         const double* __restrict__ qf,
                                                                         const double* __restrict__ ff,
                                                                         const int* __restrict__ nbr,
                                                                         long long ncells, int nx, int ny, int nz,
                                                                         int periodic,
                                                                         const double* __restrict__ dtp,
                                                                         unsigned long long* __restrict__ dt_next,
                                                                         double dx0, double dx1, double dx2,
                                                                         double gamma, double cfl, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3, NT = RiemannUpdate<N>::NT, NIT = RiemannUpdate<N>::NIT;
  constexpr int NH = NT / 32, NP = (B2 + 31) / 32;
  __shared__ double sF[6 * V * B2];
  __shared__ double red[16];
  const DGTables& T = c_tab[N];
  const long long cell = blockIdx.x;
  if (ens.cell_inst) {
    InstParams& P = ens.inst[ens.cell_inst[cell]];
    dtp = &P.dt_cur;
    dt_next = &P.dt_next;
    dx0 = P.dx[0];
    dx1 = P.dx[1];
    dx2 = P.dx[2];
    gamma = P.gamma;
  }
  const int s = threadIdx.x, hw = s >> 5, hl = s & 31;
  double p0[V];  // one node per thread: loaded before the faces
  if (NIT == 1 && s < B3) {
#pragma unroll
    for (int v = 0; v < V; ++v) p0[v] = lduh[cell * V * B3 + v * B3 + s];
  }
  // the two half-waves of a wavefront take faces 2k and 2k + 1 (NH is even),
  // so both take the same number of faces and every lane of a half-wave is
  // active at its shuffles
  for (int f = hw; f < 6; f += NH) {
    const int d = f >> 1, side = f & 1;
    long long nb;
    if (nx > 0) {  // cell indices fit 32 bits: a grid of 2^31 cells would not fit the device
      const unsigned c = (unsigned)cell;
      int ix = (int)(c % (unsigned)nx), iy = (int)((c / (unsigned)nx) % (unsigned)ny);
      int iz = (int)(c / ((unsigned)nx * (unsigned)ny));
      const int nd = d == 0 ? nx : (d == 1 ? ny : nz);
      int j = (d == 0 ? ix : (d == 1 ? iy : iz)) + (side ? 1 : -1);
      bool none = false;
      if (j < 0 || j >= nd) {
        if ((periodic >> d) & 1)
          j = j < 0 ? nd - 1 : 0;
        else
          none = true;
      }
      if (d == 0)
        ix = j;
      else if (d == 1)
        iy = j;
      else
        iz = j;
      nb = none ? -1 : ((long long)iz * ny + iy) * nx + ix;
    } else {
      nb = nbr[(size_t)(side ? d : 3 + d) * ncells + cell];
    }
    const double* Qo = qf + ((size_t)cell * 6 + f) * V * B2;
    const double* Fo = ff + ((size_t)cell * 6 + f) * V * B2;
    double* out = sF + f * V * B2;
    if (nb < 0) {
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        const int node = hl + 32 * k;
        if (node < B2) {
#pragma unroll
          for (int v = 0; v < V; ++v) out[v * B2 + node] = Fo[v * B2 + node];
        }
      }
      continue;
    }
    const double* Qn = qf + ((size_t)nb * 6 + (f ^ 1)) * V * B2;
    const double* Fn = ff + ((size_t)nb * 6 + (f ^ 1)) * V * B2;
    const double* QL = side ? Qo : Qn;
    const double* QR = side ? Qn : Qo;
    const double* FL = side ? Fo : Fn;
    const double* FR = side ? Fn : Fo;
    double aL[V], aR[V];
#pragma unroll
    for (int v = 0; v < V; ++v) aL[v] = aR[v] = 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int node = hl + 32 * k;
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
      aL[v] = half_wave_sum(aL[v]);
      aR[v] = half_wave_sum(aR[v]);
    }
    const double sL = EulerPDE::max_abs_eigenvalue(aL, d, gamma);
    const double sR = EulerPDE::max_abs_eigenvalue(aR, d, gamma);
    const double smax = sL < sR ? sR : sL;
    // several nodes per lane: the states are read again (from cache) rather
    // than held through the sums
    if (NP > 1) asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int node = hl + 32 * k;
      if (node < B2) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int i = v * B2 + node;
          out[i] = 0.5 * (FR[i] + FL[i]) - 0.5 * smax * (QR[i] - QL[i]);
        }
      }
    }
  }
  __syncthreads();
  double dtn = 1.0e300;
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int n = s + it * NT;
    if (n < B3) {
      const int z = n / B2, y = (n / B) % B, x = n % B;
      double du[V], u[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        u[v] = NIT == 1 ? p0[v] : lduh[cell * V * B3 + v * B3 + n];
        du[v] = 0.0;
      }
      const double usx = (T.w[z] * T.w[y]) / dx0;
#pragma unroll
      for (int v = 0; v < V; ++v)
        du[v] -= (sF[(1 * V + v) * B2 + z * B + y] * T.FR[x] - sF[(0 * V + v) * B2 + z * B + y] * T.FL[x]) * usx;
      const double usy = (T.w[z] * T.w[x]) / dx1;
#pragma unroll
      for (int v = 0; v < V; ++v)
        du[v] -= (sF[(3 * V + v) * B2 + z * B + x] * T.FR[y] - sF[(2 * V + v) * B2 + z * B + x] * T.FL[y]) * usy;
      const double usz = (T.w[y] * T.w[x]) / dx2;
#pragma unroll
      for (int v = 0; v < V; ++v)
        du[v] -= (sF[(5 * V + v) * B2 + y * B + x] * T.FR[z] - sF[(4 * V + v) * B2 + y * B + x] * T.FL[z]) * usz;
      const double weight = T.w[z] * T.w[y] * T.w[x];
      const double updateSize = *dtp / weight;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        u[v] += du[v] * updateSize;
        luh[cell * V * B3 + v * B3 + n] = u[v];
      }
      double m[3];
      PdeEuler::max_abs_eigenvalues3(u, PdeParams{gamma, {ens.adv[0], ens.adv[1], ens.adv[2]}}, m);
      double den = 0.0;
      den += m[0] / dx0;
      den += m[1] / dx1;
      den += m[2] / dx2;
      const double cand = exa_dt_sane(cfl * exa_pnpm(N) / den);
      dtn = cand < dtn ? cand : dtn;
    }
  }
  dtn = block_min(dtn, red);
  if (s == 0) atomic_min_pos_double(dt_next, dtn);
}

// The solution at the (N+1)^3 equidistant vertices of every cell, as a
// Cartesian-vertices plotter samples it (E/plotters/VTK/
// ADERDG2CartesianVTK.cpp:288-336, same products and summation order):
// out [cell][vertex][V], vertex (z*B + y)*B + x at offset + (x, y, z) dx / N.
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_sample_cartesian(const double* __restrict__ luh,
                                                                 double* __restrict__ out) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  const DGTables& T = c_tab[N];
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  if (s >= B3) return;
  const int i2 = s / B2, i1 = (s / B) % B, i0 = s % B;
  const double* u = luh + cell * V * B3;
#pragma unroll 1
  for (int v = 0; v < V; ++v) {
    double acc = 0.0;
    for (int j2 = 0; j2 < B; ++j2)
      for (int j1 = 0; j1 < B; ++j1)
#pragma unroll
        for (int j0 = 0; j0 < B; ++j0)
          acc += T.eq[j1][i1] * T.eq[j0][i0] * T.eq[j2][i2] * u[v * B3 + (j2 * B + j1) * B + j0];
    out[(cell * B3 + s) * V + v] = acc;
  }
}

// Initial conditions at the GL nodes (solutionAdjustment.cpph:17-52).
template <int N>
__global__ __launch_bounds__(Ord<N>::NT) void k_initial_conditions(double* __restrict__ luh, exa_ic ic,
                                                                   double gamma, int nx, int ny, double ox,
                                                                   double oy, double oz, double dx0,
                                                                   double dx1, double dx2, Ens ens) {
  using O = Ord<N>;
  constexpr int B = O::B, B2 = O::B2, B3 = O::B3;
  const DGTables& T = c_tab[N];
  const size_t cell = blockIdx.x;
  const int s = threadIdx.x;
  if (s >= B3) return;
  size_t local = cell;
  if (ens.cell_inst) {
    const int i = ens.cell_inst[cell];
    const InstParams& P = ens.inst[i];
    ic = ens.ics[i];
    gamma = P.gamma;
    nx = P.nx;
    ny = P.ny;
    ox = P.off[0];
    oy = P.off[1];
    oz = P.off[2];
    dx0 = P.dx[0];
    dx1 = P.dx[1];
    dx2 = P.dx[2];
    local = cell - (size_t)P.base;
  }
  const int z = s / B2, y = (s / B) % B, x = s % B;
  const int ix = (int)(local % nx), iy = (int)((local / nx) % ny), iz = (int)(local / ((size_t)nx * ny));
  const double c0 = ox + (ix + 0.5) * dx0, c1 = oy + (iy + 0.5) * dx1, c2 = oz + (iz + 0.5) * dx2;
  double X[3];
  X[2] = c2 + dx2 * (T.xi[z] - 0.5);
  X[1] = c1 + dx1 * (T.xi[y] - 0.5);
  X[0] = c0 + dx0 * (T.xi[x] - 0.5);
  double Q[V];
  exa_ic_point(ic, gamma, X, 0.0, Q);
#pragma unroll
  for (int v = 0; v < V; ++v) luh[cell * V * B3 + v * B3 + s] = Q[v];
}

}  // namespace exa
