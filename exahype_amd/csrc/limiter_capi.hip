// C-ABI of the subcell limiter: per-cell drop-ins for the reference's
// projectOnFVLimiterSpace / projectOnDGSpace / DMP search.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "dg_tables_data.h"
#include "host_util.h"
#include "limiter_kernels.h"

using exa::lim::LimTables;

// DMP of one cell: the boundary min/max of its six faces, the decision, and
// the reference's bookkeeping of the arrays (block 0 <- cell min/max, block 1
// <- boundary min/max).  bmm: [2][6][nobs]; mm: [2][nobs].
__global__ void k_dmp_cell(const double* __restrict__ mm, double* __restrict__ bmm, int* __restrict__ ok, int nobs,
                           double relax, double scaling) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int good = 1;
  double* bmin = bmm;
  double* bmax = bmm + 6 * nobs;
  for (int v = 0; v < nobs; ++v) {
    double lo = bmin[v], hi = bmax[v];
    for (int i = 1; i < 6; ++i) {
      lo = bmin[i * nobs + v] < lo ? bmin[i * nobs + v] : lo;
      hi = hi < bmax[i * nobs + v] ? bmax[i * nobs + v] : hi;
    }
    double sd = (hi - lo) * scaling;
    sd = sd < relax ? relax : sd;
    if (mm[v] < lo - sd || mm[nobs + v] > hi + sd) good = 0;
    bmin[v] = mm[v];
    bmax[v] = mm[nobs + v];
    bmin[v + nobs] = lo;
    bmax[v + nobs] = hi;
  }
  *ok = good;
}

struct LimCache {
  std::mutex mu;
  DevBuf tab[EXA_MAX_ORDER + 1];
  bool ready[EXA_MAX_ORDER + 1] = {};
};
static LimCache g_lim[64];

static int lim_tables(int device, int N, const LimTables** out) {
  LimCache& c = g_lim[device];
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ready[N]) {
    LimTables t;
    if (!exa::lim::build_tables(N, EXA_TAB_gl_nodes[N], EXA_TAB_gl_weights[N], EXA_TAB_gll_nodes[N], t)) {
      exa_set_error("limiter projection matrix of order %d is singular", N);
      return EXA_ERR_HIP;
    }
    EXA_TRY(c.tab[N].ensure(sizeof(LimTables)));
    EXA_CHECK(hipMemcpy(c.tab[N].p, &t, sizeof(t), hipMemcpyHostToDevice));
    c.ready[N] = true;
  }
  *out = c.tab[N].as<LimTables>();
  return EXA_OK;
}

int exa_lim_tables_dev(int device, int N, const void** out) {
  const LimTables* t = nullptr;
  EXA_TRY(lim_tables(device, N, &t));
  *out = t;
  return EXA_OK;
}

// one variable's contraction chain (the kernels loop over the variables)
static size_t proj_lds_bytes(int B) {
  const int L = 2 * B - 1;
  return sizeof(double) * (size_t)(B * B * L + B * L * L + L * L * L);
}

// A projection launch, cell list or none (block b: cell b): the order-fixed
// all-variable kernels up to B = 9, the per-variable kernel at B = 10.
static int launch_projection(bool on_fv, int B, int nblocks, hipStream_t st, const double* in, double* out,
                             const LimTables* tab, int ghost, const int* cells) {
  bool done = false;
  auto go = [&](auto bb) -> int {
    constexpr int BB = decltype(bb)::value;
    const int bytes = exa::lim::proj_t_lds_doubles<BB>() * (int)sizeof(double);
    const void* f = on_fv ? reinterpret_cast<const void*>(&exa::lim::k_project_on_fv_t<BB>)
                          : reinterpret_cast<const void*>(&exa::lim::k_project_on_dg_t<BB>);
    EXA_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    if (on_fv)
      hipLaunchKernelGGL(exa::lim::k_project_on_fv_t<BB>, dim3(nblocks), dim3(256), bytes, st, in, out, tab, ghost,
                         cells);
    else
      hipLaunchKernelGGL(exa::lim::k_project_on_dg_t<BB>, dim3(nblocks), dim3(256), bytes, st, in, out, tab, ghost,
                         cells);
    done = true;
    return EXA_OK;
  };
  switch (B) {
    case 2: EXA_TRY(go(std::integral_constant<int, 2>{})); break;
    case 3: EXA_TRY(go(std::integral_constant<int, 3>{})); break;
    case 4: EXA_TRY(go(std::integral_constant<int, 4>{})); break;
    case 5: EXA_TRY(go(std::integral_constant<int, 5>{})); break;
    case 6: EXA_TRY(go(std::integral_constant<int, 6>{})); break;
    case 7: EXA_TRY(go(std::integral_constant<int, 7>{})); break;
    case 8: EXA_TRY(go(std::integral_constant<int, 8>{})); break;
    case 9: EXA_TRY(go(std::integral_constant<int, 9>{})); break;
    default: break;
  }
  if (!done) {
    const void* f = on_fv ? reinterpret_cast<const void*>(&exa::lim::k_project_on_fv)
                          : reinterpret_cast<const void*>(&exa::lim::k_project_on_dg);
    EXA_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)proj_lds_bytes(B)));
    if (on_fv)
      hipLaunchKernelGGL(exa::lim::k_project_on_fv, dim3(nblocks), dim3(256), proj_lds_bytes(B), st, in, out, tab, B,
                         ghost, cells);
    else
      hipLaunchKernelGGL(exa::lim::k_project_on_dg, dim3(nblocks), dim3(256), proj_lds_bytes(B), st, in, out, tab, B,
                         ghost, cells);
  }
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

struct LimScratch {
  DevBuf a, b, c, d;
  std::mutex mu;
};
static LimScratch g_ls[64];

static int check_order(int N) {
  if (N < 1 || N > EXA_MAX_ORDER) {
    exa_set_error("the subcell limiter supports orders 1..9");
    return EXA_ERR_UNSUPPORTED;
  }
  return EXA_OK;
}

extern "C" int exa_limiter_project_on_fv(int order, const double* luh, int ghost, double* lim) {
  EXA_ARG(luh && lim && ghost >= 0 && ghost <= 4, "arguments");
  EXA_TRY(check_order(order));
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  const LimTables* tab = nullptr;
  EXA_TRY(lim_tables(dev, order, &tab));
  const int B = order + 1, L = 2 * order + 1, S = L + 2 * ghost, V = exa::lim::V;
  LimScratch& sc = g_ls[dev];
  std::lock_guard<std::mutex> lk(sc.mu);
  std::vector<double> hin((size_t)V * B * B * B), hout((size_t)V * S * S * S, 0.0);
  for (int n = 0; n < B * B * B; ++n)
    for (int v = 0; v < V; ++v) hin[v * B * B * B + n] = luh[n * V + v];
  for (int n = 0; n < S * S * S; ++n)
    for (int v = 0; v < V; ++v) hout[v * S * S * S + n] = lim[n * V + v];
  EXA_TRY(sc.a.ensure(sizeof(double) * hin.size()));
  EXA_TRY(sc.b.ensure(sizeof(double) * hout.size()));
  EXA_CHECK(hipMemcpy(sc.a.p, hin.data(), sizeof(double) * hin.size(), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(sc.b.p, hout.data(), sizeof(double) * hout.size(), hipMemcpyHostToDevice));
  EXA_TRY(launch_projection(true, B, 1, nullptr, sc.a.as<double>(), sc.b.as<double>(), tab, ghost, nullptr));
  EXA_CHECK(hipDeviceSynchronize());
  EXA_CHECK(hipMemcpy(hout.data(), sc.b.p, sizeof(double) * hout.size(), hipMemcpyDeviceToHost));
  for (int n = 0; n < S * S * S; ++n)
    for (int v = 0; v < V; ++v) lim[n * V + v] = hout[v * S * S * S + n];
  (void)L;
  return EXA_OK;
}

extern "C" int exa_limiter_project_on_dg(int order, const double* lim, int ghost, double* luh) {
  EXA_ARG(luh && lim && ghost >= 0 && ghost <= 4, "arguments");
  EXA_TRY(check_order(order));
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  const LimTables* tab = nullptr;
  EXA_TRY(lim_tables(dev, order, &tab));
  const int B = order + 1, L = 2 * order + 1, S = L + 2 * ghost, V = exa::lim::V;
  LimScratch& sc = g_ls[dev];
  std::lock_guard<std::mutex> lk(sc.mu);
  std::vector<double> hin((size_t)V * S * S * S), hout((size_t)V * B * B * B);
  for (int n = 0; n < S * S * S; ++n)
    for (int v = 0; v < V; ++v) hin[v * S * S * S + n] = lim[n * V + v];
  EXA_TRY(sc.a.ensure(sizeof(double) * hin.size()));
  EXA_TRY(sc.b.ensure(sizeof(double) * hout.size()));
  EXA_CHECK(hipMemcpy(sc.a.p, hin.data(), sizeof(double) * hin.size(), hipMemcpyHostToDevice));
  EXA_TRY(launch_projection(false, B, 1, nullptr, sc.a.as<double>(), sc.b.as<double>(), tab, ghost, nullptr));
  EXA_CHECK(hipDeviceSynchronize());
  EXA_CHECK(hipMemcpy(hout.data(), sc.b.p, sizeof(double) * hout.size(), hipMemcpyDeviceToHost));
  for (int n = 0; n < B * B * B; ++n)
    for (int v = 0; v < V; ++v) luh[n * V + v] = hout[v * B * B * B + n];
  (void)L;
  return EXA_OK;
}

extern "C" int exa_limiter_dmp(int order, const double* luh, int nobs, double relax, double scaling, double* bmin,
                               double* bmax, int* admissible) {
  EXA_ARG(luh && bmin && bmax && admissible && nobs >= 1 && nobs <= 5, "arguments");
  EXA_TRY(check_order(order));
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  const LimTables* tab = nullptr;
  EXA_TRY(lim_tables(dev, order, &tab));
  const int B = order + 1, L = 2 * order + 1, V = exa::lim::V;
  LimScratch& sc = g_ls[dev];
  std::lock_guard<std::mutex> lk(sc.mu);
  std::vector<double> hin((size_t)V * B * B * B), hb(12 * nobs);
  for (int n = 0; n < B * B * B; ++n)
    for (int v = 0; v < V; ++v) hin[v * B * B * B + n] = luh[n * V + v];
  std::memcpy(hb.data(), bmin, sizeof(double) * 6 * nobs);
  std::memcpy(hb.data() + 6 * nobs, bmax, sizeof(double) * 6 * nobs);
  EXA_TRY(sc.a.ensure(sizeof(double) * hin.size()));
  EXA_TRY(sc.b.ensure(sizeof(double) * 2 * nobs));
  EXA_TRY(sc.c.ensure(sizeof(double) * hb.size()));
  EXA_TRY(sc.d.ensure(sizeof(int)));
  EXA_CHECK(hipMemcpy(sc.a.p, hin.data(), sizeof(double) * hin.size(), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(sc.c.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice));
  const size_t lds = proj_lds_bytes(B) + sizeof(double) * (2 * B * B * B + 2 * 4);
  EXA_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&exa::lim::k_cell_min_max),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(exa::lim::k_cell_min_max, dim3(1), dim3(256), lds, 0, sc.a.as<double>(), sc.b.as<double>(), tab,
                     B, nobs, 1);
  hipLaunchKernelGGL(k_dmp_cell, dim3(1), dim3(64), 0, 0, sc.b.as<double>(), sc.c.as<double>(), sc.d.as<int>(), nobs,
                     relax, scaling);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipDeviceSynchronize());
  EXA_CHECK(hipMemcpy(hb.data(), sc.c.p, sizeof(double) * hb.size(), hipMemcpyDeviceToHost));
  EXA_CHECK(hipMemcpy(admissible, sc.d.p, sizeof(int), hipMemcpyDeviceToHost));
  std::memcpy(bmin, hb.data(), sizeof(double) * 6 * nobs);
  std::memcpy(bmax, hb.data() + 6 * nobs, sizeof(double) * 6 * nobs);
  (void)L;
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Troubled-cell detection on a device grid and the batched FV projection of
// the troubled cells (the per-step limiter work of LimitingADERDGSolver,
// E/solvers/LimitingADERDGSolver.cpp:838-1048, on a single-level periodic
// grid): one launch per stage over all cells, or over the compacted list of
// troubled cells.
#include <map>

#include "grid_view.h"

namespace {

struct GridLimiter {
  int nobs = 0;
  bool have_snapshot = false;
  long long mm_new_version = -1;  // grid state the last check's extrema belong to
  DevBuf mm_prev, bmm, mm_new, ok, list, count;
};
std::mutex g_gl_mu;
std::map<exa_grid*, GridLimiter*> g_gl;

GridLimiter* grid_limiter(exa_grid* g) {
  std::lock_guard<std::mutex> lk(g_gl_mu);
  GridLimiter*& p = g_gl[g];
  if (!p) p = new GridLimiter();
  return p;
}

// bmm[cell] = min/max over the cell and its six face neighbours (periodic)
// of the snapshot min/max -- the boundary extrema the reference merges over
// the faces (mergeSolutionMinMaxOnFace) before its DMP.
__global__ void k_merge_neighbour_minmax(const double* __restrict__ mm, double* __restrict__ bmm, int nx, int ny,
                                         int nz, int nobs) {
  const long long nc = (long long)nx * ny * nz;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long long)nx * ny));
  long long nb[7];
  nb[0] = c;
  nb[1] = ((long long)iz * ny + iy) * nx + (ix + 1) % nx;
  nb[2] = ((long long)iz * ny + iy) * nx + (ix + nx - 1) % nx;
  nb[3] = ((long long)iz * ny + (iy + 1) % ny) * nx + ix;
  nb[4] = ((long long)iz * ny + (iy + ny - 1) % ny) * nx + ix;
  nb[5] = ((long long)((iz + 1) % nz) * ny + iy) * nx + ix;
  nb[6] = ((long long)((iz + nz - 1) % nz) * ny + iy) * nx + ix;
  for (int v = 0; v < nobs; ++v) {
    double lo = mm[(c * 2 + 0) * nobs + v], hi = mm[(c * 2 + 1) * nobs + v];
    for (int k = 1; k < 7; ++k) {
      const double a = mm[(nb[k] * 2 + 0) * nobs + v], b = mm[(nb[k] * 2 + 1) * nobs + v];
      lo = a < lo ? a : lo;
      hi = hi < b ? b : hi;
    }
    bmm[(c * 2 + 0) * nobs + v] = lo;
    bmm[(c * 2 + 1) * nobs + v] = hi;
  }
}

// Physical admissibility of the Euler state (density and energy minima
// positive, as the Euler demonstrators' isPhysicallyAdmissible) folded into
// the DMP flags; troubled cells appended to the list.
__global__ void k_flag_and_compact(const double* __restrict__ mm_new, int* __restrict__ ok, int* __restrict__ list,
                                   int* __restrict__ count, long long nc, int nobs) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int good = ok[c];
  if (nobs >= 1 && !(mm_new[(c * 2 + 0) * nobs + 0] > 0.0)) good = 0;
  if (nobs >= 5 && !(mm_new[(c * 2 + 0) * nobs + 4] > 0.0)) good = 0;
  ok[c] = good;
  if (!good) list[atomicAdd(count, 1)] = (int)c;
}

int minmax_all(const ExaGridView& v, const LimTables* tab, double* out, int nobs) {
  // one 64-lane workgroup per (cell, observable)
  EXA_ARG(v.nc * nobs < (1LL << 31), "too many cells for one launch");
  const dim3 grid((unsigned)(v.nc * nobs)), block(64);
  switch (v.order) {
    case 1: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<2>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 2: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<3>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 3: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<4>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 4: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<5>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 5: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<6>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 6: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<7>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 7: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<8>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 8: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<9>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    case 9: hipLaunchKernelGGL(exa::lim::k_cell_min_max_w<10>, grid, block, 0, v.stream, v.luh, out, tab, nobs, nullptr); break;
    default: exa_set_error("order %d", v.order); return EXA_ERR_UNSUPPORTED;
  }
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

}  // namespace

void exa_grid_limiter_release(exa_grid* g) {
  std::lock_guard<std::mutex> lk(g_gl_mu);
  auto it = g_gl.find(g);
  if (it != g_gl.end()) {
    delete it->second;
    g_gl.erase(it);
  }
}

extern "C" int exa_grid_limiter_snapshot(exa_grid* g, int nobs) {
  EXA_ARG(g && nobs >= 1 && nobs <= 5, "grid and 1..5 observables");
  ExaGridView v;
  EXA_TRY(exa_grid_view(g, &v));
  EXA_TRY(check_order(v.order));
  EXA_ARG(v.periodic, "the grid limiter merges neighbour extrema periodically: every axis must be periodic");
  EXA_CHECK(hipSetDevice(v.device));
  const LimTables* tab = nullptr;
  EXA_TRY(lim_tables(v.device, v.order, &tab));
  GridLimiter* L = grid_limiter(g);
  EXA_TRY(L->mm_prev.ensure(sizeof(double) * 2 * nobs * v.nc));
  EXA_TRY(L->bmm.ensure(sizeof(double) * 2 * nobs * v.nc));
  if (L->mm_new_version == v.version && L->nobs == nobs) {
    // the last check computed the extrema of this very state
    L->mm_prev.swap(L->mm_new);
    L->mm_new_version = -1;
  } else {
    EXA_TRY(minmax_all(v, tab, L->mm_prev.as<double>(), nobs));
  }
  hipLaunchKernelGGL(k_merge_neighbour_minmax, dim3((unsigned)((v.nc + 255) / 256)), dim3(256), 0, v.stream,
                     L->mm_prev.as<double>(), L->bmm.as<double>(), v.nx, v.ny, v.nz, nobs);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipStreamSynchronize(v.stream));
  L->nobs = nobs;
  L->have_snapshot = true;
  return EXA_OK;
}

extern "C" int exa_grid_limiter_check(exa_grid* g, double relax, double scaling, long long* ntroubled,
                                      int* flags) {
  EXA_ARG(g && ntroubled, "null pointer");
  GridLimiter* L = grid_limiter(g);
  EXA_ARG(L->have_snapshot, "exa_grid_limiter_snapshot must precede exa_grid_limiter_check");
  ExaGridView v;
  EXA_TRY(exa_grid_view(g, &v));
  EXA_CHECK(hipSetDevice(v.device));
  const LimTables* tab = nullptr;
  EXA_TRY(lim_tables(v.device, v.order, &tab));
  const int nobs = L->nobs;
  EXA_TRY(L->mm_new.ensure(sizeof(double) * 2 * nobs * v.nc));
  EXA_TRY(L->ok.ensure(sizeof(int) * v.nc));
  EXA_TRY(L->list.ensure(sizeof(int) * v.nc));
  EXA_TRY(L->count.ensure(sizeof(int)));
  EXA_TRY(minmax_all(v, tab, L->mm_new.as<double>(), nobs));
  const unsigned blocks = (unsigned)((v.nc + 255) / 256);
  hipLaunchKernelGGL(exa::lim::k_dmp, dim3(blocks), dim3(256), 0, v.stream, L->mm_new.as<double>(),
                     L->bmm.as<double>(), L->ok.as<int>(), nobs, (int)v.nc, relax, scaling);
  EXA_CHECK(hipMemsetAsync(L->count.p, 0, sizeof(int), v.stream));
  hipLaunchKernelGGL(k_flag_and_compact, dim3(blocks), dim3(256), 0, v.stream, L->mm_new.as<double>(),
                     L->ok.as<int>(), L->list.as<int>(), L->count.as<int>(), v.nc, nobs);
  EXA_CHECK(hipGetLastError());
  int n = 0;
  EXA_CHECK(hipMemcpyAsync(&n, L->count.p, sizeof(int), hipMemcpyDeviceToHost, v.stream));
  if (flags) {
    std::vector<int> ok(v.nc);
    EXA_CHECK(hipMemcpyAsync(ok.data(), L->ok.p, sizeof(int) * v.nc, hipMemcpyDeviceToHost, v.stream));
    EXA_CHECK(hipStreamSynchronize(v.stream));
    for (long long c = 0; c < v.nc; ++c) flags[c] = ok[c] ? 0 : 1;
  }
  EXA_CHECK(hipStreamSynchronize(v.stream));
  *ntroubled = n;
  L->mm_new_version = v.version;
  return EXA_OK;
}

extern "C" int exa_grid_limiter_project(exa_grid* g, int ghost, double* d_lim, long long capacity,
                                        int* cells) {
  EXA_ARG(g && d_lim && ghost >= 0 && ghost <= 4, "arguments");
  GridLimiter* L = grid_limiter(g);
  EXA_ARG(L->list.p != nullptr, "exa_grid_limiter_check must precede exa_grid_limiter_project");
  ExaGridView v;
  EXA_TRY(exa_grid_view(g, &v));
  EXA_CHECK(hipSetDevice(v.device));
  const LimTables* tab = nullptr;
  EXA_TRY(lim_tables(v.device, v.order, &tab));
  int n = 0;
  EXA_CHECK(hipMemcpyAsync(&n, L->count.p, sizeof(int), hipMemcpyDeviceToHost, v.stream));
  EXA_CHECK(hipStreamSynchronize(v.stream));
  const int B = v.order + 1, S = 2 * v.order + 1 + 2 * ghost;
  EXA_ARG((long long)n * exa::lim::V * S * S * S <= capacity, "d_lim holds fewer patches than troubled cells");
  if (n > 0) {
    EXA_TRY(launch_projection(true, B, n, v.stream, v.luh, d_lim, tab, ghost, L->list.as<int>()));
  }
  if (cells && n > 0) EXA_CHECK(hipMemcpyAsync(cells, L->list.p, sizeof(int) * n, hipMemcpyDeviceToHost, v.stream));
  EXA_CHECK(hipStreamSynchronize(v.stream));
  return EXA_OK;
}
