// C-ABI of the subcell limiter: per-cell drop-ins for the reference's
// projectOnFVLimiterSpace / projectOnDGSpace / DMP search.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
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

static size_t proj_lds_bytes(int B) {
  const int L = 2 * B - 1;
  return sizeof(double) * exa::lim::V * (size_t)(B * B * L + B * L * L + L * L * L);
}

struct LimScratch {
  DevBuf a, b, c, d;
  std::mutex mu;
};
static LimScratch g_ls[64];

static int check_order(int N) {
  if (N < 1 || N > 5) {
    exa_set_error("the subcell limiter supports orders 1..5 (LDS budget of one workgroup)");
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
  EXA_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&exa::lim::k_project_on_fv),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)proj_lds_bytes(B)));
  hipLaunchKernelGGL(exa::lim::k_project_on_fv, dim3(1), dim3(256), proj_lds_bytes(B), 0, sc.a.as<double>(),
                     sc.b.as<double>(), tab, B, ghost, nullptr);
  EXA_CHECK(hipGetLastError());
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
  EXA_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&exa::lim::k_project_on_dg),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)proj_lds_bytes(B)));
  hipLaunchKernelGGL(exa::lim::k_project_on_dg, dim3(1), dim3(256), proj_lds_bytes(B), 0, sc.a.as<double>(),
                     sc.b.as<double>(), tab, B, ghost, nullptr);
  EXA_CHECK(hipGetLastError());
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
  const size_t lds = proj_lds_bytes(B) + sizeof(double) * (3 * V * B * B * B + 2 * nobs * 4);
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
