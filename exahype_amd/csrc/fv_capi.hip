// C-ABI of the finite-volume kernels: per-patch drop-ins for the
// FiniteVolumesSolver surface and a device-resident periodic patch grid.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "fv_kernels.h"
#include "host_util.h"

using exa::fv::PatchGeom;
static constexpr int FV_V = exa::fv::V;

struct exa_fv_solver {
  exa_fv_config cfg;
  int device = 0;
  int ghost = 1;
  hipStream_t stream = nullptr;
  std::mutex mu;
  DevBuf a, b, st;
  PatchGeom geom() const {
    PatchGeom g;
    g.P = cfg.patch;
    g.G = ghost;
    g.S = cfg.patch + 2 * ghost;
    g.dim = cfg.dim;
    return g;
  }
};

// MUSCL-Hancock kernel of a patch size (2..15 subcells per direction).
using MusclFn = void (*)(const double*, double*, double*, unsigned long long*, PatchGeom, double, double, double,
                         const double*, double, double);
static MusclFn muscl_kernel(int P) {
  switch (P) {
#define EXA_MUSCL_CASE(PP) \
  case PP:                 \
    return &exa::fv::k_muscl<EulerPDE, PP>;
    EXA_MUSCL_CASE(2) EXA_MUSCL_CASE(3) EXA_MUSCL_CASE(4) EXA_MUSCL_CASE(5) EXA_MUSCL_CASE(6) EXA_MUSCL_CASE(7)
    EXA_MUSCL_CASE(8) EXA_MUSCL_CASE(9) EXA_MUSCL_CASE(10) EXA_MUSCL_CASE(11) EXA_MUSCL_CASE(12)
    EXA_MUSCL_CASE(13) EXA_MUSCL_CASE(14) EXA_MUSCL_CASE(15)
#undef EXA_MUSCL_CASE
    default: return nullptr;
  }
}

// LDS of the MUSCL-Hancock kernel: the half-step term of every interior subcell.
static size_t muscl_lds_bytes(const PatchGeom& g) { return sizeof(double) * FV_V * (size_t)g.P * g.P * g.P; }

extern "C" int exa_fv_create(const exa_fv_config* cfg, exa_fv_solver** out) {
  EXA_ARG(cfg && out, "null pointer");
  EXA_ARG(cfg->dim == 2 || cfg->dim == 3, "dim must be 2 or 3");
  EXA_ARG(cfg->patch >= 2 && cfg->patch <= 32, "patch size 2..32");
  EXA_ARG(cfg->scheme == EXA_FV_GODUNOV || (cfg->scheme == EXA_FV_MUSCLHANCOCK && cfg->dim == 3),
          "scheme: Godunov (2D/3D) or MUSCL-Hancock (3D)");
  EXA_ARG(cfg->scheme != EXA_FV_MUSCLHANCOCK || cfg->patch <= 15,
          "MUSCL-Hancock patches up to 15^3 (the half-step terms of a patch are held in LDS)");
  EXA_ARG(cfg->nvar == 5 && cfg->npar == 0, "5 variables, no parameters");
  EXA_ARG((cfg->pde == EXA_PDE_EULER && cfg->dim == 3) || (cfg->pde == EXA_PDE_EULERFV2D && cfg->dim == 2),
          "pde: EXA_PDE_EULER in 3D, EXA_PDE_EULERFV2D in 2D");
  EXA_ARG(cfg->cfl > 0.0 && cfg->gamma > 1.0, "cfl/gamma");
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  exa_fv_solver* s = new exa_fv_solver();
  s->cfg = *cfg;
  s->device = dev;
  s->ghost = cfg->scheme == EXA_FV_MUSCLHANCOCK ? 2 : 1;
  hipError_t e = hipSuccess;
  if (cfg->scheme == EXA_FV_MUSCLHANCOCK)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(muscl_kernel(cfg->patch)),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)muscl_lds_bytes(s->geom()));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    exa_set_error("hipStreamCreate: %s", hipGetErrorString(e));
    return EXA_ERR_HIP;
  }
  *out = s;
  return EXA_OK;
}

extern "C" int exa_fv_destroy(exa_fv_solver* s) {
  if (!s) return EXA_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return EXA_OK;
}

static void patch_ref_to_soa(const double* ref, double* soa, int nbox) {
  for (int c = 0; c < nbox; ++c)
    for (int v = 0; v < FV_V; ++v) soa[v * nbox + c] = ref[c * FV_V + v];
}
static void patch_soa_to_ref(const double* soa, double* ref, int nbox) {
  for (int c = 0; c < nbox; ++c)
    for (int v = 0; v < FV_V; ++v) ref[c * FV_V + v] = soa[v * nbox + c];
}

// Launches the update of `np` patches: in -> out (out interior overwritten).
static int launch_fv_update(const exa_fv_solver* s, hipStream_t st, int np, const double* in, double* out,
                            double* dt_patch, unsigned long long* dt_min, const double dx[3], const double* dt) {
  const PatchGeom g = s->geom();
  const double gam = s->cfg.gamma;
  if (s->cfg.scheme == EXA_FV_MUSCLHANCOCK) {
    // the kernel's own CFL (kernels/finitevolumes/musclhancock/c/musclhancock.h:29)
    hipLaunchKernelGGL(muscl_kernel(g.P), dim3(np), dim3(256), muscl_lds_bytes(g), st, in, out, dt_patch, dt_min, g,
                       dx[0], dx[1], dx[2], dt, gam, 0.99);
  } else if (s->cfg.dim == 3) {
    hipLaunchKernelGGL((exa::fv::k_godunov<3, EulerPDE>), dim3(np), dim3(256), 0, st, in, out, dt_patch, dt_min, g,
                       dx[0], dx[1], dx[2], dt, gam, s->cfg.cfl);
  } else {
    hipLaunchKernelGGL((exa::fv::k_godunov<2, EulerFV2DPDE>), dim3(np), dim3(256), 0, st, in, out, dt_patch, dt_min,
                       g, dx[0], dx[1], dx[2], dt, gam, s->cfg.cfl);
  }
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

static int launch_fv_dt(const exa_fv_solver* s, hipStream_t st, int np, const double* in, double* dt_patch,
                        unsigned long long* dt_min, const double dx[3]) {
  const PatchGeom g = s->geom();
  if (s->cfg.dim == 3)
    hipLaunchKernelGGL((exa::fv::k_fv_stable_dt<3, EulerPDE>), dim3(np), dim3(256), 0, st, in, dt_patch, dt_min, g,
                       dx[0], dx[1], dx[2], s->cfg.gamma, s->cfg.cfl);
  else
    hipLaunchKernelGGL((exa::fv::k_fv_stable_dt<2, EulerFV2DPDE>), dim3(np), dim3(256), 0, st, in, dt_patch, dt_min,
                       g, dx[0], dx[1], dx[2], s->cfg.gamma, s->cfg.cfl);
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

extern "C" int exa_fv_solution_update(exa_fv_solver* s, double* luh_new, const double* luh, const double dx[3],
                                      double dt, double* max_admissible_dt) {
  EXA_ARG(s && luh_new && luh && dx, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int nbox = s->geom().ncells_box();
  std::vector<double> h(nbox * FV_V);
  EXA_TRY(s->a.ensure(sizeof(double) * nbox * FV_V));
  EXA_TRY(s->b.ensure(sizeof(double) * nbox * FV_V));
  EXA_TRY(s->st.ensure(sizeof(double) * 2));
  patch_ref_to_soa(luh, h.data(), nbox);
  EXA_CHECK(hipMemcpy(s->a.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  patch_ref_to_soa(luh_new, h.data(), nbox);
  EXA_CHECK(hipMemcpy(s->b.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(s->st.as<double>() + 1, &dt, sizeof(double), hipMemcpyHostToDevice));
  EXA_TRY(launch_fv_update(s, s->stream, 1, s->a.as<double>(), s->b.as<double>(), s->st.as<double>(), nullptr, dx,
                           s->st.as<double>() + 1));
  EXA_CHECK(hipStreamSynchronize(s->stream));
  EXA_CHECK(hipMemcpy(h.data(), s->b.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  patch_soa_to_ref(h.data(), luh_new, nbox);
  double adm = 0.0;
  EXA_CHECK(hipMemcpy(&adm, s->st.p, sizeof(double), hipMemcpyDeviceToHost));
  if (max_admissible_dt) *max_admissible_dt = adm;
  return EXA_OK;
}

extern "C" int exa_fv_stable_time_step_size(exa_fv_solver* s, const double* luh, const double dx[3], double* dt) {
  EXA_ARG(s && luh && dx && dt, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int nbox = s->geom().ncells_box();
  std::vector<double> h(nbox * FV_V);
  EXA_TRY(s->a.ensure(sizeof(double) * nbox * FV_V));
  EXA_TRY(s->st.ensure(sizeof(double) * 2));
  patch_ref_to_soa(luh, h.data(), nbox);
  EXA_CHECK(hipMemcpy(s->a.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  EXA_TRY(launch_fv_dt(s, s->stream, 1, s->a.as<double>(), s->st.as<double>(), nullptr, dx));
  EXA_CHECK(hipStreamSynchronize(s->stream));
  EXA_CHECK(hipMemcpy(dt, s->st.p, sizeof(double), hipMemcpyDeviceToHost));
  return exa_check_dt(*dt, "FV stableTimeStepSize");
}

template <class PDE>
__global__ void k_fv_rusanov_one(const double* q, double* out, int d, double gamma) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double f[FV_V];
    out[FV_V] = exa::fv::rusanov<PDE>(q, q + FV_V, d, gamma, f);
    for (int v = 0; v < FV_V; ++v) out[v] = f[v];
  }
}

extern "C" int exa_fv_riemann_solver(exa_fv_solver* s, double* fL, double* fR, const double* qL, const double* qR,
                                     int normalNonZero, double* s_max) {
  EXA_ARG(s && fL && fR && qL && qR, "null pointer");
  EXA_ARG(normalNonZero >= 0 && normalNonZero < s->cfg.dim, "normalNonZero");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  double hq[2 * FV_V], ho[FV_V + 1];
  std::memcpy(hq, qL, sizeof(double) * FV_V);
  std::memcpy(hq + FV_V, qR, sizeof(double) * FV_V);
  EXA_TRY(s->a.ensure(sizeof(hq)));
  EXA_TRY(s->b.ensure(sizeof(ho)));
  EXA_CHECK(hipMemcpy(s->a.p, hq, sizeof(hq), hipMemcpyHostToDevice));
  if (s->cfg.dim == 3)
    hipLaunchKernelGGL(k_fv_rusanov_one<EulerPDE>, dim3(1), dim3(64), 0, s->stream, s->a.as<double>(),
                       s->b.as<double>(), normalNonZero, s->cfg.gamma);
  else
    hipLaunchKernelGGL(k_fv_rusanov_one<EulerFV2DPDE>, dim3(1), dim3(64), 0, s->stream, s->a.as<double>(),
                       s->b.as<double>(), normalNonZero, s->cfg.gamma);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipStreamSynchronize(s->stream));
  EXA_CHECK(hipMemcpy(ho, s->b.p, sizeof(ho), hipMemcpyDeviceToHost));
  std::memcpy(fL, ho, sizeof(double) * FV_V);
  std::memcpy(fR, ho, sizeof(double) * FV_V);
  if (s_max) *s_max = ho[FV_V];
  return EXA_OK;
}

// ---------------------------------------------------------------------------
struct exa_fv_grid {
  exa_fv_solver* s = nullptr;
  int nx = 1, ny = 1, nz = 1, np = 0;
  double off[3] = {0, 0, 0}, width[3] = {1, 1, 1}, dx[3] = {1, 1, 1};
  DevBuf cur, nxt, nbr, state;
  int periodic[3] = {1, 1, 1};  // closed axes take the boundary values bc
  exa_bc bc{};
  hipStream_t stream = nullptr;
  bool dt_valid = false;  // a stable dt has been computed (sync checks it from then on)
  // per-phase kernel timing (0 ghost layers, 1 unused, 2 update, 3 dt)
  bool timing = false;
  std::vector<std::vector<hipEvent_t>> marks;  // 5 events per timed step
  double times[4] = {0, 0, 0, 0};
  long long counts[4] = {0, 0, 0, 0};
  double* dt_cur() const { return state.as<double>(); }
  unsigned long long* dt_next() const { return reinterpret_cast<unsigned long long*>(state.as<double>() + 1); }
  double* time() const { return state.as<double>() + 2; }
  unsigned long long* adm() const { return reinterpret_cast<unsigned long long*>(state.as<double>() + 3); }
};

__global__ void k_fv_dt_advance(double* dt_cur, unsigned long long* dt_next, double* time) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *time += *dt_cur;
    *dt_cur = __longlong_as_double((long long)*dt_next);
    *dt_next = (unsigned long long)__double_as_longlong(1.0e300);
  }
}

__global__ void k_fv_reset(unsigned long long* a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *a = (unsigned long long)__double_as_longlong(1.0e300);
}

__global__ void k_fv_take(double* dt_cur, unsigned long long* dt_next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *dt_cur = __longlong_as_double((long long)*dt_next);
    *dt_next = (unsigned long long)__double_as_longlong(1.0e300);
  }
}

// ---------------------------------------------------------------------------
// The boundary hooks of one patch (reference layout, nData per subcell).
static bool fv_bc_ok(const exa_bc* bc) { return bc->kind >= EXA_BC_OUTFLOW && bc->kind <= EXA_BC_EXACT; }

static int fv_hook(exa_fv_solver* s, exa::fv::FvHookArgs& A, double* dst, const double* src, bool dst_is_patch,
                   bool src_is_patch) {
  const exa::fv::PatchGeom geo = s->geom();
  const int nd = s->cfg.nvar + s->cfg.npar;
  const size_t patch = (size_t)geo.ncells_box() * nd;
  const size_t block = (size_t)(geo.dim == 3 ? geo.P * geo.P : geo.P) * geo.G * nd;
  const size_t nsrc = src_is_patch ? patch : block, ndst = dst_is_patch ? patch : block;
  A.dim = geo.dim;
  A.P = geo.P;
  A.G = geo.G;
  A.nd = nd;
  A.gamma = s->cfg.gamma;
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  EXA_TRY(s->a.ensure(sizeof(double) * nsrc));
  EXA_TRY(s->b.ensure(sizeof(double) * ndst));
  EXA_CHECK(hipMemcpyAsync(s->a.p, src, sizeof(double) * nsrc, hipMemcpyHostToDevice, s->stream));
  // the untouched entries of the destination keep the caller's values
  EXA_CHECK(hipMemcpyAsync(s->b.p, dst, sizeof(double) * ndst, hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(exa::fv::k_fv_patch_hook, dim3(1), dim3(256), 0, s->stream, s->b.as<double>(),
                     s->a.as<double>(), A);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipMemcpyAsync(dst, s->b.p, sizeof(double) * ndst, hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipStreamSynchronize(s->stream));
  return EXA_OK;
}

static int fv_position(const exa_fv_solver* s, const int pos[3], int* d, int* side) {
  int n = 0;
  for (int k = 0; k < 3; ++k)
    if (pos[k] != 0) {
      EXA_ARG(pos[k] == 1 || pos[k] == -1, "positions are -1, 0 or +1");
      EXA_ARG(k < s->cfg.dim, "position outside the patch's dimensions");
      *d = k;
      *side = pos[k] > 0;
      ++n;
    }
  EXA_ARG(n == 1, "exactly one non-zero position entry (a face)");
  return EXA_OK;
}

extern "C" int exa_fv_boundary_layer_extraction(exa_fv_solver* s, double* luhbnd, const double* luh,
                                                const int boundaryPosition[3]) {
  EXA_ARG(s && luhbnd && luh && boundaryPosition, "null pointer");
  exa::fv::FvHookArgs A{};
  A.op = 0;
  EXA_TRY(fv_position(s, boundaryPosition, &A.d, &A.side));
  return fv_hook(s, A, luhbnd, luh, false, true);
}

extern "C" int exa_fv_ghost_layer_filling_at_boundary(exa_fv_solver* s, double* luh, const double* luhbnd,
                                                      const int boundaryPosition[3]) {
  EXA_ARG(s && luhbnd && luh && boundaryPosition, "null pointer");
  exa::fv::FvHookArgs A{};
  A.op = 1;
  EXA_TRY(fv_position(s, boundaryPosition, &A.d, &A.side));
  return fv_hook(s, A, luh, luhbnd, true, false);
}

extern "C" int exa_fv_ghost_layer_filling(exa_fv_solver* s, double* luh, const double* luhNeighbour,
                                          const int neighbourPosition[3]) {
  EXA_ARG(s && luhNeighbour && luh && neighbourPosition, "null pointer");
  exa::fv::FvHookArgs A{};
  A.op = 2;
  EXA_TRY(fv_position(s, neighbourPosition, &A.d, &A.side));
  return fv_hook(s, A, luh, luhNeighbour, true, true);
}

extern "C" int exa_fv_boundary_conditions(exa_fv_solver* s, const exa_bc* bc, double* stateOut,
                                          const double* stateIn, const double centre[3], const double size[3],
                                          double t, double dt, int faceIndex, int normalNonZero) {
  (void)dt;
  EXA_ARG(s && bc && stateOut && stateIn && centre && size, "null pointer");
  EXA_ARG(fv_bc_ok(bc), "boundary condition kind");
  EXA_ARG(normalNonZero >= 0 && normalNonZero < s->cfg.dim && faceIndex / 2 == normalNonZero && faceIndex >= 0,
          "faceIndex must be 2 normalNonZero + 0 or 1");
  exa::fv::FvHookArgs A{};
  A.op = 3;
  A.d = normalNonZero;
  A.side = faceIndex % 2;
  A.bc = *bc;
  for (int k = 0; k < 3; ++k) {
    A.c[k] = k < s->cfg.dim ? centre[k] : 0.0;
    A.h[k] = k < s->cfg.dim ? size[k] : 0.0;
  }
  A.t = t;
  return fv_hook(s, A, stateOut, stateIn, false, false);
}

// Face neighbours nbr[(2d + side) np + p], wrapped on periodic axes and -1
// across the domain faces of the closed ones.
static hipError_t fv_grid_neighbours(exa_fv_grid* g) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int n[3] = {nx, ny, nz};
  std::vector<int> nbr(6 * g->np);
  for (int p = 0; p < g->np; ++p) {
    const int c[3] = {p % nx, (p / nx) % ny, p / (nx * ny)};
    for (int d = 0; d < 3; ++d)
      for (int side = 0; side < 2; ++side) {
        int q[3] = {c[0], c[1], c[2]};
        q[d] += side ? 1 : -1;
        const bool outside = q[d] < 0 || q[d] >= n[d];
        q[d] = (q[d] + n[d]) % n[d];
        nbr[(2 * d + side) * g->np + p] = outside && !g->periodic[d] ? -1 : (q[2] * ny + q[1]) * nx + q[0];
      }
  }
  return hipMemcpy(g->nbr.p, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice);
}

extern "C" int exa_fv_grid_create(exa_fv_solver* s, int nx, int ny, int nz, const double offset[3],
                                  const double width[3], exa_fv_grid** out) {
  EXA_ARG(s && out && nx > 0 && ny > 0 && nz > 0, "grid dimensions");
  EXA_ARG(s->cfg.dim == 3 || nz == 1, "2D grids have nz = 1");
  EXA_CHECK(hipSetDevice(s->device));
  exa_fv_grid* g = new exa_fv_grid();
  g->s = s;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->np = nx * ny * nz;
  for (int d = 0; d < 3; ++d) {
    g->off[d] = offset ? offset[d] : 0.0;
    g->width[d] = width ? width[d] : 1.0;
  }
  g->dx[0] = g->width[0] / nx;
  g->dx[1] = g->width[1] / ny;
  g->dx[2] = g->width[2] / nz;
  const int nbox = s->geom().ncells_box();
  const size_t psz = (size_t)nbox * FV_V;
  int r = EXA_OK;
  if ((r = g->cur.ensure(sizeof(double) * psz * g->np)) != EXA_OK ||
      (r = g->nxt.ensure(sizeof(double) * psz * g->np)) != EXA_OK ||
      (r = g->nbr.ensure(sizeof(int) * 6 * g->np)) != EXA_OK || (r = g->state.ensure(sizeof(double) * 8)) != EXA_OK) {
    delete g;
    return r;
  }
  hipError_t e = fv_grid_neighbours(g);
  double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const unsigned long long bb = exa_big_bits();
  std::memcpy(&st[1], &bb, 8);
  std::memcpy(&st[3], &bb, 8);
  if (e == hipSuccess) e = hipMemcpy(g->state.p, st, sizeof(st), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(g->cur.p, 0, sizeof(double) * psz * g->np);
  if (e == hipSuccess) e = hipMemset(g->nxt.p, 0, sizeof(double) * psz * g->np);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    exa_set_error("fv grid setup: %s", hipGetErrorString(e));
    delete g;
    return EXA_ERR_HIP;
  }
  *out = g;
  return EXA_OK;
}

static void fv_clear_marks(exa_fv_grid* g) {
  for (auto& m : g->marks)
    for (hipEvent_t e : m) (void)hipEventDestroy(e);
  g->marks.clear();
}

extern "C" int exa_fv_grid_set_boundary(exa_fv_grid* g, const int periodic[3], const exa_bc* bc) {
  EXA_ARG(g && periodic, "null pointer");
  const bool all = periodic[0] && periodic[1] && (periodic[2] || g->s->cfg.dim == 2);
  if (!all) EXA_ARG(bc && fv_bc_ok(bc), "a grid with a closed axis needs boundary values of a known kind");
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  for (int d = 0; d < 3; ++d) g->periodic[d] = (periodic[d] || d >= g->s->cfg.dim) ? 1 : 0;
  if (bc) g->bc = *bc;
  EXA_CHECK(fv_grid_neighbours(g));
  return EXA_OK;
}

extern "C" int exa_fv_grid_destroy(exa_fv_grid* g) {
  if (!g) return EXA_OK;
  (void)hipSetDevice(g->s->device);
  if (g->stream) {
    (void)hipStreamSynchronize(g->stream);
    (void)hipStreamDestroy(g->stream);
  }
  fv_clear_marks(g);
  delete g;
  return EXA_OK;
}

extern "C" int exa_fv_grid_enable_timing(exa_fv_grid* g, int enable) {
  EXA_ARG(g, "null pointer");
  g->timing = enable != 0;
  return EXA_OK;
}

extern "C" int exa_fv_grid_kernel_times(exa_fv_grid* g, double times[4], long long counts[4]) {
  EXA_ARG(g && times && counts, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  for (auto& m : g->marks)
    for (int p = 0; p < 4; ++p) {
      float ms = 0.f;
      EXA_CHECK(hipEventElapsedTime(&ms, m[p], m[p + 1]));
      g->times[p] += ms;
      g->counts[p] += 1;
    }
  fv_clear_marks(g);
  for (int p = 0; p < 4; ++p) {
    times[p] = g->times[p];
    counts[p] = g->counts[p];
    g->times[p] = 0.0;
    g->counts[p] = 0;
  }
  return EXA_OK;
}

static int fv_grid_transfer(exa_fv_grid* g, double* host, bool upload) {
  const int nbox = g->s->geom().ncells_box();
  const size_t psz = (size_t)nbox * FV_V;
  std::vector<double> buf(psz * g->np);
  if (upload) {
    for (int p = 0; p < g->np; ++p) patch_ref_to_soa(host + p * psz, buf.data() + p * psz, nbox);
    EXA_CHECK(hipMemcpy(g->cur.p, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
  } else {
    EXA_CHECK(hipStreamSynchronize(g->stream));
    EXA_CHECK(hipMemcpy(buf.data(), g->cur.p, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    for (int p = 0; p < g->np; ++p) patch_soa_to_ref(buf.data() + p * psz, host + p * psz, nbox);
  }
  return EXA_OK;
}

extern "C" int exa_fv_grid_upload(exa_fv_grid* g, const double* luh) {
  EXA_ARG(g && luh, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  return fv_grid_transfer(g, const_cast<double*>(luh), true);
}

extern "C" int exa_fv_grid_download(exa_fv_grid* g, double* luh) {
  EXA_ARG(g && luh, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  return fv_grid_transfer(g, luh, false);
}

extern "C" int exa_fv_grid_set_initial_conditions(exa_fv_grid* g, const exa_ic* ic) {
  EXA_ARG(g && ic, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  const PatchGeom geo = g->s->geom();
  if (g->s->cfg.dim == 3)
    hipLaunchKernelGGL(exa::fv::k_fv_initial_conditions<3>, dim3(g->np), dim3(256), 0, g->stream, g->cur.as<double>(),
                       *ic, g->s->cfg.gamma, geo, g->nx, g->ny, g->off[0], g->off[1], g->off[2], g->dx[0], g->dx[1],
                       g->dx[2]);
  else
    hipLaunchKernelGGL(exa::fv::k_fv_initial_conditions<2>, dim3(g->np), dim3(256), 0, g->stream, g->cur.as<double>(),
                       *ic, g->s->cfg.gamma, geo, g->nx, g->ny, g->off[0], g->off[1], g->off[2], g->dx[0], g->dx[1],
                       g->dx[2]);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_fv_grid_compute_dt(exa_fv_grid* g, double* dt) {
  EXA_ARG(g, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  hipLaunchKernelGGL(k_fv_reset, dim3(1), dim3(1), 0, g->stream, g->dt_next());
  EXA_TRY(launch_fv_dt(g->s, g->stream, g->np, g->cur.as<double>(), nullptr, g->dt_next(), g->dx));
  hipLaunchKernelGGL(k_fv_take, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_next());
  EXA_CHECK(hipGetLastError());
  double h = 0.0;
  EXA_CHECK(hipMemcpyAsync(&h, g->dt_cur(), sizeof(double), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  if (dt) *dt = h;
  g->dt_valid = true;
  return exa_check_dt(h, "exa_fv_grid_compute_dt");
}

extern "C" int exa_fv_grid_step(exa_fv_grid* g, int nsteps) {
  EXA_ARG(g && nsteps >= 0, "nsteps");
  EXA_CHECK(hipSetDevice(g->s->device));
  const PatchGeom geo = g->s->geom();
  exa::fv::FvBoundary fb;
  fb.bc = g->bc;
  for (int d = 0; d < 3; ++d) {
    fb.off[d] = g->off[d];
    fb.h[d] = g->dx[d];
  }
  fb.nx = g->nx;
  fb.ny = g->ny;
  fb.time = g->time();
  fb.gamma = g->s->cfg.gamma;
  for (int i = 0; i < nsteps; ++i) {
    hipEvent_t m[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    if (g->timing) {
      for (auto& e : m) EXA_CHECK(hipEventCreate(&e));
      g->marks.emplace_back(m, m + 5);
      EXA_CHECK(hipEventRecord(m[0], g->stream));
    }
    if (g->s->cfg.dim == 3)
      hipLaunchKernelGGL(exa::fv::k_fv_ghosts<3>, dim3(g->np), dim3(256), 0, g->stream, g->cur.as<double>(),
                         g->nbr.as<int>(), geo, g->np, fb);
    else
      hipLaunchKernelGGL(exa::fv::k_fv_ghosts<2>, dim3(g->np), dim3(256), 0, g->stream, g->cur.as<double>(),
                         g->nbr.as<int>(), geo, g->np, fb);
    EXA_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_fv_reset, dim3(1), dim3(1), 0, g->stream, g->adm());
    if (m[1]) EXA_CHECK(hipEventRecord(m[1], g->stream));
    if (m[2]) EXA_CHECK(hipEventRecord(m[2], g->stream));  // phase 1 is empty (one update kernel)
    EXA_TRY(launch_fv_update(g->s, g->stream, g->np, g->cur.as<double>(), g->nxt.as<double>(), nullptr, g->adm(),
                             g->dx, g->dt_cur()));
    if (m[3]) EXA_CHECK(hipEventRecord(m[3], g->stream));
    g->cur.swap(g->nxt);
    EXA_TRY(launch_fv_dt(g->s, g->stream, g->np, g->cur.as<double>(), nullptr, g->dt_next(), g->dx));
    if (m[4]) EXA_CHECK(hipEventRecord(m[4], g->stream));
    hipLaunchKernelGGL(k_fv_dt_advance, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_next(), g->time());
    EXA_CHECK(hipGetLastError());
  }
  return EXA_OK;
}

extern "C" int exa_fv_grid_sync(exa_fv_grid* g, double* time, double* next_dt, double* admissible_dt) {
  EXA_ARG(g, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  double st[4];
  EXA_CHECK(hipMemcpyAsync(st, g->state.p, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  EXA_CHECK(hipGetLastError());
  if (time) *time = st[2];
  if (next_dt) *next_dt = st[0];
  if (admissible_dt) *admissible_dt = st[3];
  return g->dt_valid ? exa_check_dt(st[0], "exa_fv_grid_sync") : EXA_OK;
}
