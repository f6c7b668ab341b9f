// C-ABI of the ADER-DG AMR level transfers: per-call drop-ins for the
// reference's faceUnknowns{Prolongation,Restriction} and
// volumeUnknowns{Prolongation,Restriction}, plus batched forms over device
// arrays for whole lists of transfers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "amr_kernels.h"
#include "dg_tables_data.h"
#include "host_util.h"

using exa::amr::AmrTables;
using exa::amr::Transfer;

namespace {

struct AmrCache {
  std::mutex mu;
  DevBuf tab[EXA_MAX_ORDER + 1];
  bool ready[EXA_MAX_ORDER + 1] = {};
  DevBuf items, groups, a, b, c, d;  // staging of the per-call and batched forms
  hipEvent_t e0 = nullptr, e1 = nullptr;  // around the last batched launch
  bool lds_set = false;
  float last_ms = 0.f;
};
AmrCache g_amr[64];

int amr_tables(int dev, int N, const AmrTables** out) {
  AmrCache& c = g_amr[dev];
  if (!c.lds_set) {
    const int bytes = 2 * exa::amr::AMR_LDS_DOUBLES * (int)sizeof(double);
    hipError_t e = hipSuccess;
    auto setattr = [&](const void* f) {
      if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    };
#define EXA_AMR_ATTR(BB)                                                              \
    setattr(reinterpret_cast<const void*>(&exa::amr::k_amr_transfer<3, true, BB>));  \
    setattr(reinterpret_cast<const void*>(&exa::amr::k_amr_transfer<3, false, BB>)); \
    setattr(reinterpret_cast<const void*>(&exa::amr::k_amr_transfer<2, true, BB>));  \
    setattr(reinterpret_cast<const void*>(&exa::amr::k_amr_transfer<2, false, BB>));
    EXA_AMR_ATTR(1) EXA_AMR_ATTR(2) EXA_AMR_ATTR(3) EXA_AMR_ATTR(4) EXA_AMR_ATTR(5)
    EXA_AMR_ATTR(6) EXA_AMR_ATTR(7) EXA_AMR_ATTR(8) EXA_AMR_ATTR(9) EXA_AMR_ATTR(10)
#undef EXA_AMR_ATTR
    EXA_CHECK(e);
    c.lds_set = true;
  }
  if (!c.ready[N]) {
    AmrTables t{};
    for (int s = 0; s < 3; ++s)
      for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) t.P[s][i][j] = EXA_TAB_fineGridProjector1d[N][s][i][j];
    for (int i = 0; i <= N; ++i) t.w[i] = EXA_TAB_gl_weights[N][i];
    EXA_TRY(c.tab[N].ensure(sizeof(AmrTables)));
    EXA_CHECK(hipMemcpy(c.tab[N].p, &t, sizeof(t), hipMemcpyHostToDevice));
    c.ready[N] = true;
  }
  *out = c.tab[N].as<AmrTables>();
  return EXA_OK;
}

int check_transfer(const exa_amr_transfer& t, int dim) {
  EXA_ARG(t.levels >= 1 && t.levels <= 10, "levels (fine level - coarse level) must be in 1..10");
  int n = 1;
  for (int l = 0; l < t.levels; ++l) n *= 3;
  for (int a = 0; a < dim; ++a) EXA_ARG(t.index[a] >= 0 && t.index[a] < n, "sub-index outside [0, 3^levels)");
  EXA_ARG(t.fine >= 0 && t.coarse >= 0, "negative block offset");
  return EXA_OK;
}

// Launches one group per workgroup on the null stream; items/groups already on
// the device.
int launch(int dim, bool prolong, const AmrTables* tab, const Transfer* items, const int* groups, int ngroups,
           double* fine, double* coarse, int B, int nd) {
  if (ngroups == 0) return EXA_OK;
  const dim3 grid(ngroups), block(exa::amr::THREADS);
  // LDS: two buffers of one value chunk (as many values per node as fit)
  const int n = dim == 3 ? B * B * B : B * B;
  const size_t lds = 2 * sizeof(double) * (size_t)n * exa::amr::amr_chunk(B, dim, nd);
  bool ok = false;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, block, lds, 0, tab, items, groups, fine, coarse, nd);
    ok = true;
  };
  switch (B) {
#define EXA_AMR_CASE(BB)                                                                           \
  case BB:                                                                                         \
    if (dim == 3)                                                                                  \
      prolong ? go(exa::amr::k_amr_transfer<3, true, BB>) : go(exa::amr::k_amr_transfer<3, false, BB>); \
    else                                                                                           \
      prolong ? go(exa::amr::k_amr_transfer<2, true, BB>) : go(exa::amr::k_amr_transfer<2, false, BB>); \
    break;
    EXA_AMR_CASE(1) EXA_AMR_CASE(2) EXA_AMR_CASE(3) EXA_AMR_CASE(4) EXA_AMR_CASE(5)
    EXA_AMR_CASE(6) EXA_AMR_CASE(7) EXA_AMR_CASE(8) EXA_AMR_CASE(9) EXA_AMR_CASE(10)
#undef EXA_AMR_CASE
    default: break;
  }
  if (!ok) {
    exa_set_error("AMR transfer: order %d", B - 1);
    return EXA_ERR_UNSUPPORTED;
  }
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

// One transfer of host arrays through the device (per-call drop-ins).
int one_transfer(int dim, bool prolong, int order, int nd, double* fine, double* coarse, int levels,
                 const int* index) {
  EXA_ARG(order >= 0 && order <= EXA_MAX_ORDER, "order must be in 0..9");
  EXA_ARG(nd >= 1 && nd <= 64, "number of data per node must be in 1..64");
  exa_amr_transfer t{};
  t.levels = levels;
  for (int a = 0; a < dim; ++a) t.index[a] = index[a];
  EXA_TRY(check_transfer(t, dim));
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  AmrCache& c = g_amr[dev];
  std::lock_guard<std::mutex> lk(c.mu);
  const AmrTables* tab = nullptr;
  EXA_TRY(amr_tables(dev, order, &tab));
  const int B = order + 1;
  const size_t n = (size_t)(dim == 3 ? B * B * B : B * B) * nd;
  Transfer it{0, 0, levels, {t.index[0], t.index[1], t.index[2]}};
  const int grp[2] = {0, 1};
  EXA_TRY(c.items.ensure(sizeof(Transfer)));
  EXA_TRY(c.groups.ensure(sizeof(grp)));
  EXA_TRY(c.a.ensure(sizeof(double) * n));
  EXA_TRY(c.b.ensure(sizeof(double) * n));
  EXA_CHECK(hipMemcpy(c.items.p, &it, sizeof(it), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(c.groups.p, grp, sizeof(grp), hipMemcpyHostToDevice));
  // a = fine, b = coarse
  if (prolong)
    EXA_CHECK(hipMemcpy(c.b.p, coarse, sizeof(double) * n, hipMemcpyHostToDevice));
  else {
    EXA_CHECK(hipMemcpy(c.a.p, fine, sizeof(double) * n, hipMemcpyHostToDevice));
    EXA_CHECK(hipMemcpy(c.b.p, coarse, sizeof(double) * n, hipMemcpyHostToDevice));
  }
  EXA_TRY(launch(dim, prolong, tab, c.items.as<Transfer>(), c.groups.as<int>(), 1, c.a.as<double>(),
                 c.b.as<double>(), B, nd));
  EXA_CHECK(hipDeviceSynchronize());
  if (prolong)
    EXA_CHECK(hipMemcpy(fine, c.a.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  else
    EXA_CHECK(hipMemcpy(coarse, c.b.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return EXA_OK;
}

int batched(int dim, bool prolong, int order, int nd, double* d_fine, long long n_fine, double* d_coarse,
            long long n_coarse, const exa_amr_transfer* transfers, int count) {
  EXA_ARG(order >= 0 && order <= EXA_MAX_ORDER, "order must be in 0..9");
  EXA_ARG(nd >= 1 && nd <= 64, "number of data per node must be in 1..64");
  EXA_ARG(count >= 0 && (count == 0 || (transfers && d_fine && d_coarse)), "arguments");
  if (count == 0) return EXA_OK;
  const int B = order + 1;
  const long long blk = (long long)(dim == 3 ? B * B * B : B * B) * nd;
  std::vector<Transfer> items(count);
  std::vector<int> groups;
  groups.reserve(count + 1);
  for (int i = 0; i < count; ++i) {
    EXA_TRY(check_transfer(transfers[i], dim));
    const exa_amr_transfer& t = transfers[i];
    EXA_ARG(t.fine + blk <= n_fine && t.coarse + blk <= n_coarse, "transfer block beyond the end of its array");
    items[i] = Transfer{t.fine, t.coarse, t.levels, {t.index[0], t.index[1], dim == 3 ? t.index[2] : 0}};
    // prolongation: one group per transfer; restriction: runs of equal coarse offsets
    if (prolong || i == 0 || t.coarse != transfers[i - 1].coarse) groups.push_back(i);
  }
  groups.push_back(count);
  // Output blocks of different workgroups must not overlap: each fine block
  // is written by one prolongation, each coarse block by one run of restrictions.
  {
    std::vector<long long> outs;
    for (size_t g = 0; g + 1 < groups.size(); ++g)
      outs.push_back(prolong ? items[groups[g]].fine : items[groups[g]].coarse);
    std::sort(outs.begin(), outs.end());
    for (size_t i = 1; i < outs.size(); ++i)
      EXA_ARG(outs[i] >= outs[i - 1] + blk,
              prolong ? "two prolongations write overlapping fine blocks"
                      : "restrictions into one coarse block must be consecutive in the list, and coarse blocks "
                        "must not overlap");
  }
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  AmrCache& c = g_amr[dev];
  std::lock_guard<std::mutex> lk(c.mu);
  const AmrTables* tab = nullptr;
  EXA_TRY(amr_tables(dev, order, &tab));
  EXA_TRY(c.items.ensure(sizeof(Transfer) * items.size()));
  EXA_TRY(c.groups.ensure(sizeof(int) * groups.size()));
  EXA_CHECK(hipMemcpy(c.items.p, items.data(), sizeof(Transfer) * items.size(), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(c.groups.p, groups.data(), sizeof(int) * groups.size(), hipMemcpyHostToDevice));
  if (!c.e0) {
    EXA_CHECK(hipEventCreate(&c.e0));
    EXA_CHECK(hipEventCreate(&c.e1));
  }
  EXA_CHECK(hipEventRecord(c.e0, nullptr));
  EXA_TRY(launch(dim, prolong, tab, c.items.as<Transfer>(), c.groups.as<int>(), (int)groups.size() - 1, d_fine,
                 d_coarse, B, nd));
  EXA_CHECK(hipEventRecord(c.e1, nullptr));
  EXA_CHECK(hipDeviceSynchronize());
  float ms = 0.f;
  EXA_CHECK(hipEventElapsedTime(&ms, c.e0, c.e1));
  c.last_ms = ms;
  return EXA_OK;
}

}  // namespace

extern "C" int exa_amr_face_unknowns_prolongation(int order, int nvar, int npar, double* lQhbndFine,
                                                  double* lFhbndFine, const double* lQhbndCoarse,
                                                  const double* lFhbndCoarse, int coarseGridLevel, int fineGridLevel,
                                                  const int subfaceIndex[2]) {
  EXA_ARG(lQhbndFine && lFhbndFine && lQhbndCoarse && lFhbndCoarse && subfaceIndex, "null pointer");
  EXA_ARG(nvar >= 1 && npar >= 0, "nvar/npar");
  const int lv = fineGridLevel - coarseGridLevel;
  EXA_TRY(one_transfer(2, true, order, nvar + npar, lQhbndFine, const_cast<double*>(lQhbndCoarse), lv, subfaceIndex));
  return one_transfer(2, true, order, nvar, lFhbndFine, const_cast<double*>(lFhbndCoarse), lv, subfaceIndex);
}

extern "C" int exa_amr_face_unknowns_restriction(int order, int nvar, int npar, double* lQhbndCoarse,
                                                 double* lFhbndCoarse, const double* lQhbndFine,
                                                 const double* lFhbndFine, int coarseGridLevel, int fineGridLevel,
                                                 const int subfaceIndex[2]) {
  EXA_ARG(lQhbndFine && lFhbndFine && lQhbndCoarse && lFhbndCoarse && subfaceIndex, "null pointer");
  EXA_ARG(nvar >= 1 && npar >= 0, "nvar/npar");
  const int lv = fineGridLevel - coarseGridLevel;
  EXA_TRY(one_transfer(2, false, order, nvar + npar, const_cast<double*>(lQhbndFine), lQhbndCoarse, lv, subfaceIndex));
  return one_transfer(2, false, order, nvar, const_cast<double*>(lFhbndFine), lFhbndCoarse, lv, subfaceIndex);
}

extern "C" int exa_amr_volume_unknowns_prolongation(int order, int nvar, int npar, double* luhFine,
                                                    const double* luhCoarse, int coarseGridLevel, int fineGridLevel,
                                                    const int subcellIndex[3]) {
  EXA_ARG(luhFine && luhCoarse && subcellIndex, "null pointer");
  EXA_ARG(nvar >= 1 && npar >= 0, "nvar/npar");
  return one_transfer(3, true, order, nvar + npar, luhFine, const_cast<double*>(luhCoarse),
                      fineGridLevel - coarseGridLevel, subcellIndex);
}

extern "C" int exa_amr_volume_unknowns_restriction(int order, int nvar, int npar, double* luhCoarse,
                                                   const double* luhFine, int coarseGridLevel, int fineGridLevel,
                                                   const int subcellIndex[3]) {
  EXA_ARG(luhFine && luhCoarse && subcellIndex, "null pointer");
  EXA_ARG(nvar >= 1 && npar >= 0, "nvar/npar");
  return one_transfer(3, false, order, nvar + npar, const_cast<double*>(luhFine), luhCoarse,
                      fineGridLevel - coarseGridLevel, subcellIndex);
}

extern "C" int exa_amr_prolongation_batched(int dim, int order, int ndata, double* d_fine, long long n_fine,
                                            const double* d_coarse, long long n_coarse,
                                            const exa_amr_transfer* transfers, int count) {
  EXA_ARG(dim == 2 || dim == 3, "dim must be 2 (faces) or 3 (volumes)");
  return batched(dim, true, order, ndata, d_fine, n_fine, const_cast<double*>(d_coarse), n_coarse, transfers, count);
}

extern "C" int exa_amr_restriction_batched(int dim, int order, int ndata, double* d_coarse, long long n_coarse,
                                           const double* d_fine, long long n_fine,
                                           const exa_amr_transfer* transfers, int count) {
  EXA_ARG(dim == 2 || dim == 3, "dim must be 2 (faces) or 3 (volumes)");
  return batched(dim, false, order, ndata, const_cast<double*>(d_fine), n_fine, d_coarse, n_coarse, transfers, count);
}

extern "C" int exa_amr_last_kernel_ms(double* ms) {
  EXA_ARG(ms, "null pointer");
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_ARG(dev >= 0 && dev < 64, "device index");
  AmrCache& c = g_amr[dev];
  std::lock_guard<std::mutex> lk(c.mu);
  *ms = c.last_ms;
  return EXA_OK;
}
