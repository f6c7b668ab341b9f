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
  hipStream_t st = nullptr;               // the batched launches' stream (events on it time the kernel alone)
  bool lds_set = false;
  float last_ms = 0.f;
};
AmrCache g_amr[64];

void fill_tables(int N, AmrTables& t) {
  for (int s = 0; s < 3; ++s)
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N; ++j) t.P[s][i][j] = EXA_TAB_fineGridProjector1d[N][s][i][j];
  for (int i = 0; i <= N; ++i) t.w[i] = EXA_TAB_gl_weights[N][i];
  // level_op of one level, the same expressions
  for (int s = 0; s < 3; ++s)
    for (int o = 0; o <= N; ++o)
      for (int i = 0; i <= N; ++i) {
        t.Op[0][s][o][i] = t.P[s][i][o];
        t.Op[1][s][o][i] = t.w[i] * t.P[s][o][i] / (t.w[o] * 3.0);
      }
}

// Dynamic-LDS limit of every transfer kernel instance, once per device.
int amr_lds_attrs(AmrCache& c) {
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
  return EXA_OK;
}

int amr_tables(int dev, int N, const AmrTables** out) {
  AmrCache& c = g_amr[dev];
  EXA_TRY(amr_lds_attrs(c));
  if (!c.ready[N]) {
    AmrTables t{};
    fill_tables(N, t);
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

// Launches the groups on a stream (the null stream by default); items/groups
// already on the device.  ntransfers (all groups; < 0: not known): with small
// groups each workgroup takes several in turn, about 16 transfers, so that
// the register-staged kernel has the next block in flight; groups of 8 or
// more transfers get a workgroup each.
int launch(int dim, bool prolong, const AmrTables* tab, const Transfer* items, const int* groups, int ngroups,
           double* fine, double* coarse, int B, int nd, hipStream_t st = nullptr, long long ntransfers = -1,
           bool one_level = false) {
  if (ngroups == 0) return EXA_OK;
  long long nwg = ngroups;
  if (ntransfers >= 0 && ntransfers < 8LL * ngroups) {
    const long long want = (ntransfers + 15) / 16;
    nwg = want < 2048 ? 2048 : want;
    if (nwg > ngroups) nwg = ngroups;
  }
  const dim3 grid((unsigned)nwg), block(exa::amr::THREADS);
  // LDS: two buffers of one value chunk (as many values per node as fit)
  const int n = dim == 3 ? B * B * B : B * B;
  const size_t lds = 2 * sizeof(double) * (size_t)n * exa::amr::amr_chunk(B, dim, nd);
  bool ok = false;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, block, lds, st, tab, items, groups, fine, coarse, nd, ngroups, one_level ? 1 : 0);
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

// Validates a transfer list and fills items[count] and groups[*ngroups + 1]
// (group g = items groups[g] .. groups[g+1]-1); outs needs count entries.
// Prolongation: one group per transfer; restriction: runs of equal coarse
// offsets.  Output blocks of different groups must not overlap.
int build_list(int dim, bool prolong, long long blk, long long n_fine, long long n_coarse,
               const exa_amr_transfer* transfers, int count, Transfer* items, int* groups, long long* outs,
               int* ngroups) {
  int ng = 0;
  for (int i = 0; i < count; ++i) {
    EXA_TRY(check_transfer(transfers[i], dim));
    const exa_amr_transfer& t = transfers[i];
    EXA_ARG(t.fine + blk <= n_fine && t.coarse + blk <= n_coarse, "transfer block beyond the end of its array");
    items[i] = Transfer{t.fine, t.coarse, t.levels, {t.index[0], t.index[1], dim == 3 ? t.index[2] : 0}};
    if (prolong || i == 0 || t.coarse != transfers[i - 1].coarse) groups[ng++] = i;
  }
  groups[ng] = count;
  for (int g = 0; g < ng; ++g) outs[g] = prolong ? items[groups[g]].fine : items[groups[g]].coarse;
  std::sort(outs, outs + ng);
  for (int i = 1; i < ng; ++i)
    EXA_ARG(outs[i] >= outs[i - 1] + blk,
            prolong ? "two prolongations write overlapping fine blocks"
                    : "restrictions into one coarse block must be consecutive in the list, and coarse blocks "
                      "must not overlap");
  *ngroups = ng;
  return EXA_OK;
}

bool all_one_level(const exa_amr_transfer* transfers, int count) {
  for (int i = 0; i < count; ++i)
    if (transfers[i].levels != 1) return false;
  return true;
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
  std::vector<int> groups(count + 1);
  std::vector<long long> outs(count);
  int ng = 0;
  EXA_TRY(build_list(dim, prolong, blk, n_fine, n_coarse, transfers, count, items.data(), groups.data(),
                     outs.data(), &ng));
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  AmrCache& c = g_amr[dev];
  std::lock_guard<std::mutex> lk(c.mu);
  const AmrTables* tab = nullptr;
  EXA_TRY(amr_tables(dev, order, &tab));
  EXA_TRY(c.items.ensure(sizeof(Transfer) * count));
  EXA_TRY(c.groups.ensure(sizeof(int) * (ng + 1)));
  EXA_CHECK(hipMemcpy(c.items.p, items.data(), sizeof(Transfer) * count, hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(c.groups.p, groups.data(), sizeof(int) * (ng + 1), hipMemcpyHostToDevice));
  if (!c.e0) {
    EXA_CHECK(hipEventCreate(&c.e0));
    EXA_CHECK(hipEventCreate(&c.e1));
    EXA_CHECK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
  }
  // the caller's data may come from any stream: start after all of it
  EXA_CHECK(hipDeviceSynchronize());
  EXA_CHECK(hipEventRecord(c.e0, c.st));
  EXA_TRY(launch(dim, prolong, tab, c.items.as<Transfer>(), c.groups.as<int>(), ng, d_fine, d_coarse, B, nd, c.st,
                 count, all_one_level(transfers, count)));
  EXA_CHECK(hipEventRecord(c.e1, c.st));
  EXA_CHECK(hipStreamSynchronize(c.st));
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

// Device tables of one order and a transfer launch on a stream, for units that
// keep their own transfer lists on the device (the adaptive-mesh step).
int exa_amr_tables_dev(int dev, int order, const void** tab) {
  AmrCache& c = g_amr[dev];
  std::lock_guard<std::mutex> lk(c.mu);
  const AmrTables* t = nullptr;
  EXA_TRY(amr_tables(dev, order, &t));
  *tab = t;
  return EXA_OK;
}

int exa_amr_enqueue(int dim, bool prolong, int order, const void* tab, const void* items, const int* groups,
                    int ngroups, double* fine, double* coarse, int nd, hipStream_t st) {
  return launch(dim, prolong, static_cast<const AmrTables*>(tab), static_cast<const Transfer*>(items), groups,
                ngroups, fine, coarse, order + 1, nd, st);
}

// ---------------------------------------------------------------------------
// Allocation-free batched restriction.  For callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The operator tables and the transfer list are
// staged in `scratch`, a fixed file-scope device array of
// EXA_AMR_NOALLOC_MAX_TRANSFERS transfers (about 590 KB) that every GPU the
// library runs on holds until the process ends.  exa_init takes that array's
// address on each device once (exa_amr_noalloc_setup); host code only ever
// uses that address, never the array's name.  Calls are serialised by a
// process-wide lock; the array keeps the last call's tables and list, and two
// processes calling at once on one GPU would race on it.
namespace {
struct NoAllocScratch {
  AmrTables tab;
  Transfer items[EXA_AMR_NOALLOC_MAX_TRANSFERS];
  int groups[EXA_AMR_NOALLOC_MAX_TRANSFERS + 1];
};
}  // namespace

static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

namespace {
NoAllocScratch* g_scratch[64];  // per device: that device's copy of `scratch`
std::mutex g_noalloc_mu;
// host staging of the list (static: the call allocates no host memory either)
Transfer h_items[EXA_AMR_NOALLOC_MAX_TRANSFERS];
int h_groups[EXA_AMR_NOALLOC_MAX_TRANSFERS + 1];
long long h_outs[EXA_AMR_NOALLOC_MAX_TRANSFERS];
}  // namespace

// Called by exa_init once per device, with that device current.
int exa_amr_noalloc_setup(int dev) {
  if (dev < 0 || dev >= 64) {
    exa_set_error("device index %d", dev);
    return EXA_ERR_ARG;
  }
  NoAllocScratch* ptr = nullptr;
  EXA_CHECK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
  AmrCache& c = g_amr[dev];
  std::lock_guard<std::mutex> lk(c.mu);
  EXA_TRY(amr_lds_attrs(c));
  std::lock_guard<std::mutex> lk2(g_noalloc_mu);
  g_scratch[dev] = ptr;
  return EXA_OK;
}

extern "C" int exa_amr_restriction_batched_noalloc(int dim, int order, int ndata, double* d_coarse,
                                                   long long n_coarse, const double* d_fine, long long n_fine,
                                                   const exa_amr_transfer* transfers, int count) {
  EXA_ARG(dim == 2 || dim == 3, "dim must be 2 (faces) or 3 (volumes)");
  EXA_ARG(order >= 0 && order <= EXA_MAX_ORDER, "order must be in 0..9");
  EXA_ARG(ndata >= 1 && ndata <= 64, "number of data per node must be in 1..64");
  EXA_ARG(count >= 0 && (count == 0 || (transfers && d_fine && d_coarse)), "arguments");
  if (count > EXA_AMR_NOALLOC_MAX_TRANSFERS) {
    exa_set_error("invalid argument: %d transfers exceed the %d of the allocation-free staging area", count,
                  EXA_AMR_NOALLOC_MAX_TRANSFERS);
    return EXA_ERR_ARG;
  }
  if (count == 0) return EXA_OK;
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_ARG(dev >= 0 && dev < 64, "device index");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoAllocScratch* sp = g_scratch[dev];
  EXA_ARG(sp, "exa_init has not run on the current device (it sets up the staging area)");
  const int B = order + 1;
  const long long blk = (long long)(dim == 3 ? B * B * B : B * B) * ndata;
  int ng = 0;
  EXA_TRY(build_list(dim, false, blk, n_fine, n_coarse, transfers, count, h_items, h_groups, h_outs, &ng));
  AmrTables t{};
  fill_tables(order, t);
  EXA_CHECK(hipMemcpy(&sp->tab, &t, sizeof(t), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(sp->items, h_items, sizeof(Transfer) * count, hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(sp->groups, h_groups, sizeof(int) * (ng + 1), hipMemcpyHostToDevice));
  EXA_TRY(launch(dim, false, &sp->tab, sp->items, sp->groups, ng, const_cast<double*>(d_fine), d_coarse, B, ndata,
                 nullptr, count, all_one_level(transfers, count)));
  EXA_CHECK(hipDeviceSynchronize());
  return EXA_OK;
}
