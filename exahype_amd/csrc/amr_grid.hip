// C-ABI of the limiting ADER-DG step on a static adaptive mesh
// (exa_amr_grid_*): the mesh is built on the host once, every stage of a step
// is one launch over all leaves (or over a device-built list of limiter
// patches), and a step is replayed as one hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "amr_grid_kernels.h"
#include "amr_kernels.h"
#include "host_util.h"
#include "solver_internal.h"

using namespace exa;
using namespace exa::amrg;

hipError_t exa_amr_grid_init(const DGTables* tab) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), tab, sizeof(DGTables) * (EXA_MAX_ORDER + 1));
}

namespace {

// ---------------------------------------------------------------------------
// the static mesh (host)

struct MeshBuild {
  int n0 = 0, levels = 1;
  double c[3] = {0, 0, 0}, r = 0, w[3] = {0, 0, 0};
  std::vector<std::unordered_set<long long>> refined;

  long long nl(int l) const {
    long long n = n0;
    for (int i = 0; i < l; ++i) n *= 3;
    return n;
  }
  long long key(int l, long long x, long long y, long long z) const {
    const long long n = nl(l);
    return (z * n + y) * n + x;
  }
  bool is_refined(int l, long long x, long long y, long long z) const {
    return l < levels && refined[l].count(key(l, x, y, z)) != 0;
  }
  bool present(int l, long long x, long long y, long long z) const {
    return l == 0 || is_refined(l - 1, x / 3, y / 3, z / 3);
  }
  std::vector<std::array<int, 3>> present_cells(int l) const {
    std::vector<std::array<int, 3>> out;
    if (l == 0) {
      for (int k = 0; k < n0; ++k)
        for (int j = 0; j < n0; ++j)
          for (int i = 0; i < n0; ++i) out.push_back({i, j, k});
      return out;
    }
    const long long n = nl(l - 1);
    for (long long kk : refined[l - 1]) {
      const int x = (int)(kk % n), y = (int)((kk / n) % n), z = (int)(kk / (n * n));
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int e = 0; e < 3; ++e) out.push_back({3 * x + e, 3 * y + b, 3 * z + a});
    }
    return out;
  }
  // the cell box meets the shell radius - w <= |x - c| <= radius + w
  bool meets_shell(int l, const std::array<int, 3>& q) const {
    const double h = 1.0 / (double)nl(l);
    double dn = 0.0, df = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double lo = q[a] * h, hi = lo + h;
      const double near = std::min(std::max(c[a], lo), hi);
      const double far = std::fabs(lo - c[a]) > std::fabs(hi - c[a]) ? lo : hi;
      dn += (near - c[a]) * (near - c[a]);
      df += (far - c[a]) * (far - c[a]);
    }
    return std::sqrt(dn) <= r + w[l] && std::sqrt(df) >= r - w[l];
  }
  std::array<int, 3> wrap(int l, std::array<long long, 3> q) const {
    const long long n = nl(l);
    return {(int)(((q[0] % n) + n) % n), (int)(((q[1] % n) + n) % n), (int)(((q[2] % n) + n) % n)};
  }
  bool refine_by_shell() {
    bool changed = false;
    for (int l = 0; l + 1 < levels; ++l)
      for (const auto& q : present_cells(l))
        if (!is_refined(l, q[0], q[1], q[2]) && meets_shell(l, q)) {
          refined[l].insert(key(l, q[0], q[1], q[2]));
          changed = true;
        }
    return changed;
  }
  // no leaf may have a face neighbour leaf two levels finer
  bool balance() {
    bool changed = false;
    for (int l = 0; l + 2 < levels; ++l)
      for (const auto& q : present_cells(l)) {
        if (is_refined(l, q[0], q[1], q[2])) continue;
        bool hit = false;
        for (int d = 0; d < 3 && !hit; ++d)
          for (int s = 0; s < 2 && !hit; ++s) {
            std::array<long long, 3> nb{q[0], q[1], q[2]};
            nb[d] += s ? 1 : -1;
            const auto w3 = wrap(l, nb);
            if (!is_refined(l, w3[0], w3[1], w3[2])) continue;
            const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
            for (int a = 0; a < 3 && !hit; ++a)
              for (int b = 0; b < 3 && !hit; ++b) {
                long long ch[3];
                ch[d] = 3LL * w3[d] + (s ? 0 : 2);
                ch[t0] = 3LL * w3[t0] + b;
                ch[t1] = 3LL * w3[t1] + a;
                if (is_refined(l + 1, ch[0], ch[1], ch[2])) hit = true;
              }
          }
        if (hit) {
          refined[l].insert(key(l, q[0], q[1], q[2]));
          changed = true;
        }
      }
    return changed;
  }
};

struct Leaf {
  int l, x, y, z;
};

// Refinement by the shell and 2:1 balance until nothing changes; the leaves
// ordered by (level, z, y, x).
void build_mesh(const exa_amr_config& cfg, MeshBuild& m, std::vector<Leaf>& leaves) {
  m.n0 = cfg.base_cells;
  m.levels = cfg.levels;
  for (int a = 0; a < 3; ++a) m.c[a] = cfg.centre[a];
  m.r = cfg.radius;
  m.w[0] = cfg.width[0];
  m.w[1] = cfg.width[1];
  m.refined.assign(cfg.levels, {});
  for (bool changed = true; changed;) {
    const bool a = m.refine_by_shell();
    const bool b = m.balance();
    changed = a || b;
  }
  leaves.clear();
  for (int l = 0; l < m.levels; ++l)
    for (const auto& q : m.present_cells(l))
      if (!m.is_refined(l, q[0], q[1], q[2])) leaves.push_back({l, q[0], q[1], q[2]});
  std::sort(leaves.begin(), leaves.end(), [](const Leaf& a, const Leaf& b) {
    if (a.l != b.l) return a.l < b.l;
    if (a.z != b.z) return a.z < b.z;
    if (a.y != b.y) return a.y < b.y;
    return a.x < b.x;
  });
}

int check_mesh_cfg(const exa_amr_config* cfg) {
  EXA_ARG(cfg->levels >= 1 && cfg->levels <= 3, "levels must be 1..3");
  EXA_ARG(cfg->base_cells >= 3 && cfg->base_cells <= 243, "base_cells must be 3..243");
  return EXA_OK;
}

}  // namespace

extern "C" int exa_amr_mesh_cells(const exa_amr_config* cfg, int* cells, long long capacity, long long* n) {
  EXA_ARG(cfg && n, "null pointer");
  EXA_TRY(check_mesh_cfg(cfg));
  MeshBuild m;
  std::vector<Leaf> leaves;
  build_mesh(*cfg, m, leaves);
  *n = (long long)leaves.size();
  if (cells)
    for (long long q = 0; q < std::min<long long>(capacity, *n); ++q) {
      cells[4 * q + 0] = leaves[q].l;
      cells[4 * q + 1] = leaves[q].x;
      cells[4 * q + 2] = leaves[q].y;
      cells[4 * q + 3] = leaves[q].z;
    }
  return EXA_OK;
}

// ---------------------------------------------------------------------------

struct exa_amr_grid {
  exa_solver* s = nullptr;
  exa_amr_config cfg{};
  int device = 0, N = 0, B = 0, P = 0, S = 0, levels = 1, finest = 0;
  hipStream_t stream = nullptr;
  std::vector<Leaf> leaves;
  long long nc = 0, first_fine = 0, nslots = 0, nfaces = 0, cap = 0;
  int n_prol_groups = 0, n_restr_groups = 0;
  long long nanc_slots = 0, ndesc = 0;
  std::vector<double> h;
  // device
  DevBuf geo, level, hbuf, inst, luh, luh_prev, lduh, qf, ff, iters, iter_sum;
  DevBuf face_kind, face_ref, anc_kids, nbr6, faces;
  DevBuf prol_items, prol_groups, restr_items, restr_groups;
  DevBuf status, status_old, fw, fw_tmp, cure, slot_of, new_slot, fv_done;
  DevBuf fmin, fmax, fmin_tmp, fmax_tmp, mm, lmm;
  DevBuf pool[2], stack, top, lists, counts, irregular, cnt;
  const void* amr_tab = nullptr;
  const void* lim_tab = nullptr;
  PredFn pred = nullptr;
  int pred_threads = 0;
  size_t pred_lds = 0;
  int par = 0;  // pool holding the patches at the start of the next step
  bool use_graph = true, timing = false, dt_valid = false;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[6];
  std::vector<hipEvent_t> all_ev;
  long long timed_steps = 0;

  int* list_cells(int k) { return lists.as<int>() + (size_t)(2 * k) * cap; }
  int* list_slots(int k) { return lists.as<int>() + (size_t)(2 * k + 1) * cap; }
  int* count(int k) { return counts.as<int>() + k; }
  InstParams* ip() { return inst.as<InstParams>(); }
  exa::Ens ens() {
    exa::Ens e{};
    e.cell_inst = level.as<int>();
    e.inst = ip();
    return e;
  }
};

namespace {

unsigned blocks_for(long long n, int t = 256) { return (unsigned)std::max<long long>(1, (n + t - 1) / t); }

void release_graphs(exa_amr_grid* g) {
  for (int i = 0; i < 2; ++i) {
    if (g->gexec[i]) (void)hipGraphExecDestroy(g->gexec[i]);
    if (g->graph[i]) (void)hipGraphDestroy(g->graph[i]);
    g->gexec[i] = nullptr;
    g->graph[i] = nullptr;
  }
}

void clear_events(exa_amr_grid* g) {
  for (hipEvent_t e : g->all_ev) (void)hipEventDestroy(e);
  g->all_ev.clear();
  for (auto& v : g->ev) v.clear();
  g->timed_steps = 0;
}

// Patch-size dispatch of the list kernels (P = 2N+1).
template <class F>
int with_patch(int P, F f) {
  switch (P) {
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 9: f(std::integral_constant<int, 9>{}); break;
    case 11: f(std::integral_constant<int, 11>{}); break;
    case 13: f(std::integral_constant<int, 13>{}); break;
    case 15: f(std::integral_constant<int, 15>{}); break;
    default: exa_set_error("limiter patch of %d subcells", P); return EXA_ERR_UNSUPPORTED;
  }
  return EXA_OK;
}

template <class F>
int with_order(int N, F f) {
  switch (N) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: exa_set_error("order %d (the adaptive mesh takes orders 1..7)", N); return EXA_ERR_UNSUPPORTED;
  }
  return EXA_OK;
}

size_t proj_lds(int B) {
  const int L = 2 * B - 1;
  return sizeof(double) * (size_t)V * (B * B * L + B * L * L);  // lim::proj_t_lds_doubles<B>
}
size_t muscl_lds(int P) { return sizeof(double) * V * (size_t)P * P * P; }

// ---- the limiter's starting state (setInitialConditions, then the status
// after them: admissibility only, LimitingADERDGSolver.cpp:906-930)
// only_if: a device flag that must be set for any work to be done (null: always)
int enqueue_minmax_all(exa_amr_grid* g, double* out, const int* only_if = nullptr) {
  const auto* tab = static_cast<const lim::LimTables*>(g->lim_tab);
  return with_order(g->N, [&](auto n) {
    constexpr int NN = decltype(n)::value;
    hipLaunchKernelGGL(lim::k_cell_min_max_w<NN + 1>, dim3((unsigned)(g->nc * NOBS)), dim3(64), 0, g->stream,
                       g->luh.as<double>(), out, tab, NOBS, only_if);
  });
}

__global__ void k_limiter_start(const double* __restrict__ mm, double* __restrict__ fmin, double* __restrict__ fmax,
                                int* __restrict__ status, int* __restrict__ fw, int* __restrict__ cure,
                                int* __restrict__ slot_of, const int* __restrict__ level, int finest, long long nc,
                                int steps_till_cured, int* __restrict__ any_troubled) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  for (int i = 0; i < 6; ++i)
    for (int v = 0; v < NOBS; ++v) {
      fmin[(q * 6 + i) * NOBS + v] = mm[(q * 2 + 0) * NOBS + v];
      fmax[(q * 6 + i) * NOBS + v] = mm[(q * 2 + 1) * NOBS + v];
    }
  int s = ST_OK;
  if (level[q] == finest && !(mm[(q * 2 + 0) * NOBS + 0] > 0.0 && mm[(q * 2 + 0) * NOBS + 4] > 0.0)) {
    s = ST_TROUBLED;
    cure[q] = steps_till_cured + 1;
    atomicOr(any_troubled, 1);
  } else {
    cure[q] = 0;
  }
  status[q] = s;
  for (int i = 0; i < 6; ++i) fw[q * 6 + i] = s;
  slot_of[q] = -1;
}

__global__ void k_stack_init(int* __restrict__ stack, int* __restrict__ top, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap) stack[i] = cap - 1 - i;
  if (i == 0) *top = cap;
}

int enqueue_spreading(exa_amr_grid* g, int rounds) {
  const long long nf = g->nc - g->first_fine;
  for (int r = 0; r < rounds; ++r) {
    hipLaunchKernelGGL(k_merge_meta, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->fw.as<int>(),
                       g->fw_tmp.as<int>(), nullptr, nullptr, nullptr, nullptr, g->face_kind.as<int>(),
                       g->face_ref.as<int>(), g->anc_kids.as<int>(), g->level.as<int>(), g->finest, g->first_fine,
                       g->nc, 0, g->irregular.as<int>());
    hipLaunchKernelGGL(k_spread, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->fw_tmp.as<int>(), g->fw.as<int>(),
                       g->status.as<int>(), g->first_fine, g->nc, g->irregular.as<int>());
  }
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

int enqueue_build(exa_amr_grid* g, int mode, int k, const int* st_old, const int* only_if) {
  const long long nf = g->nc - g->first_fine;
  hipLaunchKernelGGL(k_build_list, dim3(blocks_for(nf)), dim3(256), 0, g->stream, mode, st_old, g->status.as<int>(),
                     g->slot_of.as<int>(), g->new_slot.as<int>(), g->fv_done.as<int>(), g->first_fine, g->nc,
                     g->list_cells(k), g->list_slots(k), g->count(k), only_if);
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

int enqueue_project_fv(exa_amr_grid* g, const double* src, double* pool, int k) {
  EXA_TRY(with_patch(g->P, [&](auto p) {
    constexpr int PP = decltype(p)::value;
    hipLaunchKernelGGL(k_project_fv_list<PP>, dim3((unsigned)g->cap), dim3(256), proj_lds(g->B), g->stream, src, pool,
                       g->list_cells(k), g->list_slots(k), g->count(k),
                       static_cast<const lim::LimTables*>(g->lim_tab));
  }));
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

// ic: the FV initial conditions of the patches; nullptr: projections of the solution
int limiter_start(exa_amr_grid* g, const exa_ic* ic) {
  EXA_TRY(enqueue_minmax_all(g, g->mm.as<double>()));
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, g->stream, g->irregular.as<int>(), 1LL, 0);
  hipLaunchKernelGGL(k_limiter_start, dim3(blocks_for(g->nc)), dim3(256), 0, g->stream, g->mm.as<double>(),
                     g->fmin.as<double>(), g->fmax.as<double>(), g->status.as<int>(), g->fw.as<int>(),
                     g->cure.as<int>(), g->slot_of.as<int>(), g->level.as<int>(), g->finest, g->nc,
                     g->cfg.steps_till_cured, g->irregular.as<int>());
  EXA_TRY(enqueue_spreading(g, 5));
  hipLaunchKernelGGL(k_stack_init, dim3(blocks_for(g->cap)), dim3(256), 0, g->stream, g->stack.as<int>(),
                     g->top.as<int>(), (int)g->cap);
  const long long nf = g->nc - g->first_fine;
  hipLaunchKernelGGL(k_slots_alloc, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->status.as<int>(),
                     g->slot_of.as<int>(), g->new_slot.as<int>(), g->stack.as<int>(), g->top.as<int>(),
                     g->first_fine, g->nc, g->cnt.as<unsigned long long>());
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, g->stream, g->counts.as<int>(), 4LL, 0);
  EXA_TRY(enqueue_build(g, LIST_NEW_SLOT, LIST_NEW_SLOT, g->status.as<int>(), nullptr));
  if (ic) {
    const double gamma = exa_solver_cfg(g->s)->gamma;
    EXA_TRY(with_patch(g->P, [&](auto p) {
      constexpr int PP = decltype(p)::value;
      hipLaunchKernelGGL(k_amr_fv_ic<PP>, dim3((unsigned)g->cap), dim3(256), 0, g->stream,
                         g->pool[g->par].as<double>(), g->list_cells(LIST_NEW_SLOT), g->list_slots(LIST_NEW_SLOT),
                         g->count(LIST_NEW_SLOT), g->geo.as<int4>(), g->hbuf.as<double>(), *ic, gamma);
    }));
    EXA_CHECK(hipGetLastError());
  } else {
    EXA_TRY(enqueue_project_fv(g, g->luh.as<double>(), g->pool[g->par].as<double>(), LIST_NEW_SLOT));
  }
  EXA_CHECK(hipStreamSynchronize(g->stream));
  release_graphs(g);
  return EXA_OK;
}

int enqueue_dt(exa_amr_grid* g) {
  const double cfl = exa_solver_cfg(g->s)->cfl;
  EXA_TRY(with_order(g->N, [&](auto n) {
    constexpr int NN = decltype(n)::value;
    hipLaunchKernelGGL(k_amr_dt<NN>, dim3((unsigned)g->nc), dim3(Ord<NN>::NT), 0, g->stream, g->luh.as<double>(),
                       g->status.as<int>(), g->level.as<int>(), g->finest, g->ip(), cfl);
  }));
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

// ---- one step
int enqueue_step(exa_amr_grid* g, int par, bool timed) {
  const int N = g->N, B = g->B, P = g->P;
  const long long nc = g->nc, nf = nc - g->first_fine;
  double* pool_old = g->pool[par].as<double>();
  double* pool_new = g->pool[par ^ 1].as<double>();
  const auto* ltab = static_cast<const lim::LimTables*>(g->lim_tab);
  int* irr = g->irregular.as<int>();
  unsigned long long* cnt = g->cnt.as<unsigned long long>();
  const double gamma = exa_solver_cfg(g->s)->gamma;
  const double hf = g->h[g->finest];
  hipEvent_t e0 = nullptr;
  int phase = -1;
  auto mark = [&](int next) -> int {
    if (!timed) return EXA_OK;
    hipEvent_t e = nullptr;
    EXA_CHECK(hipEventCreate(&e));
    g->all_ev.push_back(e);
    EXA_CHECK(hipEventRecord(e, g->stream));
    if (phase >= 0) g->ev[phase].emplace_back(e0, e);
    e0 = e;
    phase = next;
    return EXA_OK;
  };
  EXA_TRY(mark(0));
  // start of the step: flags, the statuses the merges and the switch read
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, g->stream, irr, 1LL, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, g->stream, g->counts.as<int>(), 4LL, 0);
  hipLaunchKernelGGL(k_copy_int, dim3(blocks_for(nc)), dim3(256), 0, g->stream, g->status_old.as<int>(),
                     g->status.as<int>(), nc);
  hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(nc)), dim3(256), 0, g->stream, g->fv_done.as<int>(), nc, 0);
  // Prediction
  exa::PredictorDebug nodbg;
  nodbg.lQi = nodbg.lFi = nodbg.lQhi = nodbg.lFhi = nullptr;
  hipLaunchKernelGGL(g->pred, dim3((unsigned)nc), dim3(g->pred_threads), g->pred_lds, g->stream, g->luh.as<double>(),
                     g->lduh.as<double>(), g->qf.as<double>(), g->ff.as<double>(), g->iters.as<int>(), nullptr, 1.0,
                     1.0, 1.0, gamma, exa_solver_cfg(g->s)->fp64_sweeps, g->iter_sum.as<unsigned long long>(), nodbg,
                     g->ens());
  EXA_CHECK(hipGetLastError());
  EXA_TRY(mark(1));
  // Sending: face data of the virtual cells (one value per node and call)
  const long long B2 = (long long)B * B;
  if (g->nanc_slots > 0) {
    const long long off = (nc * 6 + g->ndesc) * V * B2, n = g->nanc_slots * V * B2;
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(256), 0, g->stream, g->qf.as<double>() + off, n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(256), 0, g->stream, g->ff.as<double>() + off, n, 0.0);
    EXA_TRY(exa_amr_enqueue(2, false, N, g->amr_tab, g->restr_items.p, g->restr_groups.as<int>(), g->n_restr_groups,
                            g->qf.as<double>(), g->qf.as<double>(), 1, g->stream));
    EXA_TRY(exa_amr_enqueue(2, false, N, g->amr_tab, g->restr_items.p, g->restr_groups.as<int>(), g->n_restr_groups,
                            g->ff.as<double>(), g->ff.as<double>(), 1, g->stream));
  }
  if (g->ndesc > 0) {
    EXA_TRY(exa_amr_enqueue(2, true, N, g->amr_tab, g->prol_items.p, g->prol_groups.as<int>(), g->n_prol_groups,
                            g->qf.as<double>(), g->qf.as<double>(), 1, g->stream));
    EXA_TRY(exa_amr_enqueue(2, true, N, g->amr_tab, g->prol_items.p, g->prol_groups.as<int>(), g->n_prol_groups,
                            g->ff.as<double>(), g->ff.as<double>(), 1, g->stream));
  }
  // Merging: limiter status and DMP extrema of every face
  hipLaunchKernelGGL(k_merge_meta, dim3(blocks_for(nc)), dim3(256), 0, g->stream, g->fw.as<int>(), g->fw_tmp.as<int>(),
                     g->fmin.as<double>(), g->fmax.as<double>(), g->fmin_tmp.as<double>(), g->fmax_tmp.as<double>(),
                     g->face_kind.as<int>(), g->face_ref.as<int>(), g->anc_kids.as<int>(), g->level.as<int>(),
                     g->finest, 0LL, nc, 1, nullptr);
  hipLaunchKernelGGL(k_copy_int, dim3(blocks_for(nc * 6)), dim3(256), 0, g->stream, g->fw.as<int>(),
                     g->fw_tmp.as<int>(), nc * 6);
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(nc * 6 * NOBS)), dim3(256), 0, g->stream, g->fmin.as<double>(),
                     g->fmin_tmp.as<double>(), nc * 6 * NOBS);
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(nc * 6 * NOBS)), dim3(256), 0, g->stream, g->fmax.as<double>(),
                     g->fmax_tmp.as<double>(), nc * 6 * NOBS);
  EXA_CHECK(hipGetLastError());
  EXA_TRY(mark(2));
  EXA_TRY(with_order(N, [&](auto n) {
    constexpr int NN = decltype(n)::value;
    hipLaunchKernelGGL(k_riemann_faces<NN>, dim3(blocks_for(g->nfaces * 64)), dim3(256), 0, g->stream,
                       g->qf.as<double>(), g->ff.as<double>(), g->faces.as<RFace>(), g->nfaces, g->status_old.as<int>(),
                       gamma, cnt);
  }));
  EXA_CHECK(hipGetLastError());
  EXA_TRY(mark(3));
  // Update: every leaf by ADER-DG; the FV range of the finest level is then
  // overwritten from its patches
  const long long nd = nc * V * (long long)B * B * B;
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(nd)), dim3(256), 0, g->stream, g->luh_prev.as<double>(),
                     g->luh.as<double>(), nd);
  EXA_TRY(with_order(N, [&](auto n) {
    constexpr int NN = decltype(n)::value;
    hipLaunchKernelGGL((k_update<NN, UPD_SURFACE | UPD_UPDATE>), dim3((unsigned)nc), dim3(Ord<NN>::NT), 0, g->stream,
                       g->luh.as<double>(), g->lduh.as<double>(), g->ff.as<double>(), nullptr, nullptr, 1.0, 1.0, 1.0,
                       gamma, exa_solver_cfg(g->s)->cfl, g->ens());
  }));
  EXA_CHECK(hipGetLastError());
  EXA_TRY(mark(4));
  const unsigned capb = (unsigned)g->cap;
  auto fv_pass = [&](int k, const int* st) -> int {
    EXA_TRY(with_patch(P, [&](auto p) {
      constexpr int PP = decltype(p)::value;
      hipLaunchKernelGGL(k_ghost_fill<PP>, dim3(capb), dim3(256), 0, g->stream, pool_old, g->list_cells(k),
                         g->list_slots(k), g->count(k), g->slot_of.as<int>(), st, g->nbr6.as<int>(), cnt);
      hipLaunchKernelGGL(k_muscl_list<PP>, dim3(capb), dim3(256), muscl_lds(PP), g->stream, pool_old, pool_new,
                         g->list_slots(k), g->count(k), hf, g->ip(), g->finest, gamma);
    }));
    EXA_TRY(with_patch(P, [&](auto p) {
      constexpr int PP = decltype(p)::value;
      hipLaunchKernelGGL(k_project_dg_list<PP>, dim3(capb), dim3(256), proj_lds(B), g->stream, pool_new,
                         g->luh.as<double>(), g->list_cells(k), g->list_slots(k), g->count(k), ltab,
                         g->fv_done.as<int>(), cnt);
    }));
    EXA_CHECK(hipGetLastError());
    return EXA_OK;
  };
  auto patch_minmax = [&](int k) -> int {
    return with_patch(P, [&](auto p) {
      constexpr int PP = decltype(p)::value;
      hipLaunchKernelGGL(k_patch_min_max<PP>, dim3(capb), dim3(256), 0, g->stream, pool_new, g->list_cells(k),
                         g->list_slots(k), g->count(k), g->lmm.as<double>());
    });
  };
  EXA_TRY(enqueue_build(g, LIST_FV_OLD, LIST_FV_OLD, g->status_old.as<int>(), nullptr));
  EXA_TRY(fv_pass(LIST_FV_OLD, g->status_old.as<int>()));
  // DMP, admissibility, new status
  EXA_TRY(enqueue_minmax_all(g, g->mm.as<double>()));
  EXA_TRY(patch_minmax(LIST_FV_OLD));
  hipLaunchKernelGGL(k_status, dim3(blocks_for(nc)), dim3(256), 0, g->stream, g->mm.as<double>(), g->lmm.as<double>(),
                     g->fmin.as<double>(), g->fmax.as<double>(), g->status_old.as<int>(), g->status.as<int>(),
                     g->fw.as<int>(), g->cure.as<int>(), g->level.as<int>(), g->finest, nc, g->cfg.dmp_relaxation,
                     g->cfg.dmp_scaling, g->cfg.steps_till_cured, irr, cnt);
  EXA_CHECK(hipGetLastError());
  // irregular change: spreading, then patches, rollback and recomputation
  EXA_TRY(enqueue_spreading(g, 5));
  hipLaunchKernelGGL(k_cure_reset, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->status.as<int>(),
                     g->cure.as<int>(), g->first_fine, nc, g->cfg.steps_till_cured, irr);
  hipLaunchKernelGGL(k_slots_free, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->status.as<int>(),
                     g->slot_of.as<int>(), g->stack.as<int>(), g->top.as<int>(), g->first_fine, nc);
  hipLaunchKernelGGL(k_slots_alloc, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->status.as<int>(),
                     g->slot_of.as<int>(), g->new_slot.as<int>(), g->stack.as<int>(), g->top.as<int>(),
                     g->first_fine, nc, cnt);
  EXA_TRY(enqueue_build(g, LIST_NEW_SLOT, LIST_NEW_SLOT, g->status_old.as<int>(), irr));
  EXA_TRY(enqueue_project_fv(g, g->luh_prev.as<double>(), pool_old, LIST_NEW_SLOT));
  EXA_TRY(enqueue_build(g, LIST_FV_NEW, LIST_FV_NEW, g->status_old.as<int>(), irr));
  hipLaunchKernelGGL(k_rollback_list, dim3(capb), dim3(256), 0, g->stream, g->luh.as<double>(),
                     g->luh_prev.as<double>(), g->list_cells(LIST_FV_NEW), g->count(LIST_FV_NEW),
                     g->fv_done.as<int>(), B * B * B);
  EXA_TRY(fv_pass(LIST_FV_NEW, g->status.as<int>()));
  // the patches of cells ADER-DG advanced hold their new solution
  EXA_TRY(enqueue_build(g, LIST_DG_PATCH, LIST_DG_PATCH, g->status_old.as<int>(), nullptr));
  EXA_TRY(enqueue_project_fv(g, g->luh.as<double>(), pool_new, LIST_DG_PATCH));
  // after a recomputation the extrema of every cell start afresh
  EXA_TRY(enqueue_minmax_all(g, g->mm.as<double>(), irr));
  EXA_TRY(patch_minmax(LIST_FV_NEW));
  hipLaunchKernelGGL(k_minmax_reset, dim3(blocks_for(nc)), dim3(256), 0, g->stream, g->mm.as<double>(),
                     g->lmm.as<double>(), g->fmin.as<double>(), g->fmax.as<double>(), g->status.as<int>(),
                     g->level.as<int>(), g->finest, nc, irr);
  EXA_CHECK(hipGetLastError());
  EXA_TRY(mark(5));
  // TimeStepSizeComputation: the ADER-DG time step of every leaf
  EXA_TRY(enqueue_dt(g));
  hipLaunchKernelGGL(k_step_end, dim3(blocks_for(nf)), dim3(256), 0, g->stream, g->ip(), g->levels, irr,
                     g->status.as<int>(), g->nbr6.as<int>(), g->first_fine, nc, cnt);
  EXA_CHECK(hipGetLastError());
  EXA_TRY(mark(-1));
  return EXA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" int exa_amr_grid_create(exa_solver* s, const exa_amr_config* cfg, exa_amr_grid** out) {
  EXA_ARG(s && cfg && out, "null pointer");
  EXA_TRY(check_mesh_cfg(cfg));
  EXA_ARG(cfg->dmp_relaxation >= 0.0 && cfg->dmp_scaling >= 0.0 && cfg->steps_till_cured >= 0,
          "DMP parameters and steps-till-cured must not be negative");
  const exa_aderdg_config& sc = *exa_solver_cfg(s);
  EXA_ARG(sc.nvar == 5 && sc.npar == 0, "5 variables, no parameters");
  auto g = new exa_amr_grid();
  std::unique_ptr<exa_amr_grid> guard(g);
  g->s = s;
  g->cfg = *cfg;
  g->device = exa_solver_device(s);
  g->N = sc.order;
  g->B = g->N + 1;
  g->P = 2 * g->N + 1;
  g->S = g->P + 2 * GHOST;
  g->levels = cfg->levels;
  g->finest = cfg->levels - 1;
  EXA_CHECK(hipSetDevice(g->device));
  EXA_TRY(exa_solver_predictor(s, &g->pred, &g->pred_threads, &g->pred_lds));
  // mesh
  MeshBuild m;
  build_mesh(*cfg, m, g->leaves);
  g->nc = (long long)g->leaves.size();
  EXA_ARG(g->nc < (1LL << 31) / 6, "too many cells");
  g->first_fine = 0;
  while (g->first_fine < g->nc && g->leaves[g->first_fine].l != g->finest) ++g->first_fine;
  std::vector<std::unordered_map<long long, int>> index(g->levels);
  for (long long q = 0; q < g->nc; ++q) {
    const Leaf& L = g->leaves[q];
    index[L.l][m.key(L.l, L.x, L.y, L.z)] = (int)q;
  }
  g->h.resize(g->levels);
  for (int l = 0; l < g->levels; ++l) g->h[l] = 1.0 / (double)m.nl(l);
  // faces
  const long long nc = g->nc;
  std::vector<int> face_kind(nc * 6), face_ref(nc * 6), nbr6(nc * 6, -1), anc_kids;
  std::vector<RFace> faces;
  struct Desc { long long q; int f; long long r; int fr, s0, s1; };
  struct Anc { long long q; int f; int kids[9]; int sub[9][2]; };
  std::vector<Desc> desc;
  std::vector<Anc> anc;
  for (long long q = 0; q < nc; ++q) {
    const Leaf& L = g->leaves[q];
    for (int d = 0; d < 3; ++d)
      for (int s2 = 0; s2 < 2; ++s2) {
        const int f = 2 * d + s2;
        std::array<long long, 3> nb{L.x, L.y, L.z};
        nb[d] += s2 ? 1 : -1;
        const auto w3 = m.wrap(L.l, nb);
        const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
        if (m.present(L.l, w3[0], w3[1], w3[2])) {
          if (!m.is_refined(L.l, w3[0], w3[1], w3[2])) {
            const int r = index[L.l].at(m.key(L.l, w3[0], w3[1], w3[2]));
            face_kind[q * 6 + f] = FK_SAME;
            face_ref[q * 6 + f] = r;
            if (L.l == g->finest) nbr6[q * 6 + f] = r;
            if (s2 == 1) {
              const bool fin = L.l == g->finest;
              faces.push_back(RFace{q * 6 + 2 * d + 1, (long long)r * 6 + 2 * d, d, fin ? (int)q : -1, fin ? r : -1, 0});
            }
          } else {
            Anc a{};
            a.q = q;
            a.f = f;
            int k = 0;
            for (int aa = 0; aa < 3; ++aa)
              for (int bb = 0; bb < 3; ++bb) {
                long long ch[3];
                ch[d] = 3LL * w3[d] + (s2 ? 0 : 2);
                ch[t0] = 3LL * w3[t0] + bb;
                ch[t1] = 3LL * w3[t1] + aa;
                const auto it = index[L.l + 1].find(m.key(L.l + 1, ch[0], ch[1], ch[2]));
                if (it == index[L.l + 1].end()) {
                  exa_set_error("mesh not 2:1 balanced at a refined face");
                  return EXA_ERR_ARG;
                }
                a.kids[k] = it->second;
                a.sub[k][0] = (int)(ch[t0] - 3LL * w3[t0]);
                a.sub[k][1] = (int)(ch[t1] - 3LL * w3[t1]);
                ++k;
              }
            face_kind[q * 6 + f] = FK_FINER;
            face_ref[q * 6 + f] = (int)anc.size();
            anc.push_back(a);
          }
        } else {
          const long long px = w3[0] / 3, py = w3[1] / 3, pz = w3[2] / 3;
          const auto it = index[L.l - 1].find(m.key(L.l - 1, px, py, pz));
          if (it == index[L.l - 1].end()) {
            exa_set_error("mesh not 2:1 balanced at a coarse face");
            return EXA_ERR_ARG;
          }
          const long long par[3] = {px, py, pz};
          face_kind[q * 6 + f] = FK_COARSER;
          face_ref[q * 6 + f] = it->second;
          desc.push_back(Desc{q, f, it->second, f ^ 1, (int)(w3[t0] - 3 * par[t0]), (int)(w3[t1] - 3 * par[t1])});
        }
      }
  }
  g->ndesc = (long long)desc.size();
  g->nanc_slots = (long long)anc.size();
  g->nslots = g->ndesc + g->nanc_slots;
  const int B2 = g->B * g->B;
  std::vector<exa::amr::Transfer> prol, restr;
  std::vector<int> prol_g, restr_g;
  for (long long j = 0; j < g->ndesc; ++j) {
    const Desc& D = desc[j];
    const long long slot = nc * 6 + j;
    const int d = D.f / 2;
    if (D.f & 1)
      faces.push_back(RFace{D.q * 6 + D.f, slot, d, -1, -1, 0});
    else
      faces.push_back(RFace{slot, D.q * 6 + D.f, d, -1, -1, 0});
    for (int v = 0; v < V; ++v) {
      prol_g.push_back((int)prol.size());
      prol.push_back(exa::amr::Transfer{(slot * V + v) * B2, ((D.r * 6 + D.fr) * V + v) * B2, 1, {D.s0, D.s1, 0}});
    }
  }
  prol_g.push_back((int)prol.size());
  for (long long j = 0; j < g->nanc_slots; ++j) {
    const Anc& A = anc[j];
    const long long slot = nc * 6 + g->ndesc + j;
    const int d = A.f / 2;
    if (A.f & 1)
      faces.push_back(RFace{A.q * 6 + A.f, slot, d, -1, -1, 0});
    else
      faces.push_back(RFace{slot, A.q * 6 + A.f, d, -1, -1, 0});
    for (int v = 0; v < V; ++v) {
      restr_g.push_back((int)restr.size());
      for (int k = 0; k < 9; ++k)
        restr.push_back(exa::amr::Transfer{((long long)(A.kids[k] * 6LL + (A.f ^ 1)) * V + v) * B2,
                                           (slot * V + v) * B2, 1, {A.sub[k][0], A.sub[k][1], 0}});
    }
    for (int k = 0; k < 9; ++k) anc_kids.push_back(A.kids[k]);
  }
  restr_g.push_back((int)restr.size());
  g->n_prol_groups = (int)prol_g.size() - 1;
  g->n_restr_groups = (int)restr_g.size() - 1;
  g->nfaces = (long long)faces.size();
  // limiter patch pool
  const long long nfine = nc - g->first_fine;
  const size_t patch_bytes = sizeof(double) * V * (size_t)g->S * g->S * g->S;
  const int B3 = g->B * B2;
  const size_t cell_bytes = sizeof(double) * (size_t)V * (3 * B3 + 12 * B2) + sizeof(double) * 4 * 6 * NOBS +
                            sizeof(double) * 4 * NOBS + sizeof(int) * 32;
  size_t free_b = 0, total_b = 0;
  EXA_CHECK(hipMemGetInfo(&free_b, &total_b));
  long long cap = cfg->patch_capacity > 0 ? cfg->patch_capacity : std::max<long long>(nfine, 1);
  cap = std::min<long long>(cap, std::max<long long>(nfine, 1));
  if (cfg->patch_capacity <= 0) {
    // default: every finest cell, or as many as two pools fit into 60% of the
    // memory the mesh leaves free
    const size_t base = cell_bytes * (size_t)nc + sizeof(double) * V * B2 * g->nslots * 2;
    const size_t room = free_b > base ? (size_t)((free_b - base) * 0.6) : 0;
    cap = std::max<long long>(1, std::min<long long>(cap, (long long)(room / (2 * patch_bytes))));
  }
  const size_t need = cell_bytes * (size_t)nc + 2 * patch_bytes * (size_t)cap + sizeof(double) * V * B2 * g->nslots * 2;
  if (need > free_b) {
    exa_set_error("the adaptive mesh needs %.1f GB of device memory, %.1f GB are free", need / 1e9, free_b / 1e9);
    return EXA_ERR_NOMEM;
  }
  g->cap = cap;
  // device arrays
  std::vector<int4> geo(nc);
  std::vector<int> lev(nc);
  for (long long q = 0; q < nc; ++q) {
    const Leaf& L = g->leaves[q];
    geo[q] = make_int4(L.l, L.x, L.y, L.z);
    lev[q] = L.l;
  }
  auto up = [&](DevBuf& b, const void* src, size_t bytes) -> int {
    EXA_TRY(b.ensure(std::max<size_t>(bytes, 16)));
    if (bytes) EXA_CHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return EXA_OK;
  };
  EXA_TRY(up(g->geo, geo.data(), sizeof(int4) * nc));
  EXA_TRY(up(g->level, lev.data(), sizeof(int) * nc));
  EXA_TRY(up(g->hbuf, g->h.data(), sizeof(double) * g->levels));
  std::vector<InstParams> ip(g->levels);
  for (int l = 0; l < g->levels; ++l) {
    ip[l] = InstParams{};
    ip[l].dx[0] = ip[l].dx[1] = ip[l].dx[2] = g->h[l];
    ip[l].gamma = sc.gamma;
    ip[l].dt_cur = 0.0;
    const double big = 1.0e300;
    std::memcpy(&ip[l].dt_next, &big, sizeof(big));
    ip[l].time = 0.0;
  }
  EXA_TRY(up(g->inst, ip.data(), sizeof(InstParams) * g->levels));
  const size_t cellv = sizeof(double) * (size_t)nc * V * B3;
  const size_t facev = sizeof(double) * (size_t)(nc * 6 + g->nslots) * V * B2;
  EXA_TRY(g->luh.ensure(cellv));
  EXA_TRY(g->luh_prev.ensure(cellv));
  EXA_TRY(g->lduh.ensure(cellv));
  EXA_TRY(g->qf.ensure(facev));
  EXA_TRY(g->ff.ensure(facev));
  EXA_CHECK(hipMemset(g->qf.p, 0, facev));
  EXA_CHECK(hipMemset(g->ff.p, 0, facev));
  EXA_TRY(g->iters.ensure(sizeof(int) * nc));
  EXA_TRY(g->iter_sum.ensure(sizeof(unsigned long long)));
  EXA_CHECK(hipMemset(g->iter_sum.p, 0, sizeof(unsigned long long)));
  EXA_TRY(up(g->face_kind, face_kind.data(), sizeof(int) * nc * 6));
  EXA_TRY(up(g->face_ref, face_ref.data(), sizeof(int) * nc * 6));
  EXA_TRY(up(g->anc_kids, anc_kids.data(), sizeof(int) * anc_kids.size()));
  EXA_TRY(up(g->nbr6, nbr6.data(), sizeof(int) * nc * 6));
  EXA_TRY(up(g->faces, faces.data(), sizeof(RFace) * faces.size()));
  EXA_TRY(up(g->prol_items, prol.data(), sizeof(exa::amr::Transfer) * prol.size()));
  EXA_TRY(up(g->prol_groups, prol_g.data(), sizeof(int) * prol_g.size()));
  EXA_TRY(up(g->restr_items, restr.data(), sizeof(exa::amr::Transfer) * restr.size()));
  EXA_TRY(up(g->restr_groups, restr_g.data(), sizeof(int) * restr_g.size()));
  for (DevBuf* b : {&g->status, &g->status_old, &g->cure, &g->slot_of, &g->new_slot, &g->fv_done})
    EXA_TRY(b->ensure(sizeof(int) * nc));
  EXA_TRY(g->fw.ensure(sizeof(int) * nc * 6));
  EXA_TRY(g->fw_tmp.ensure(sizeof(int) * nc * 6));
  for (DevBuf* b : {&g->fmin, &g->fmax, &g->fmin_tmp, &g->fmax_tmp}) EXA_TRY(b->ensure(sizeof(double) * nc * 6 * NOBS));
  EXA_TRY(g->mm.ensure(sizeof(double) * nc * 2 * NOBS));
  EXA_TRY(g->lmm.ensure(sizeof(double) * nc * 2 * NOBS));
  EXA_TRY(g->pool[0].ensure(patch_bytes * cap));
  EXA_TRY(g->pool[1].ensure(patch_bytes * cap));
  EXA_TRY(g->stack.ensure(sizeof(int) * cap));
  EXA_TRY(g->top.ensure(sizeof(int)));
  EXA_TRY(g->lists.ensure(sizeof(int) * 8 * cap));
  EXA_TRY(g->counts.ensure(sizeof(int) * 4));
  EXA_TRY(g->irregular.ensure(sizeof(int)));
  EXA_TRY(g->cnt.ensure(sizeof(unsigned long long) * NCNT));
  EXA_CHECK(hipMemset(g->cnt.p, 0, sizeof(unsigned long long) * NCNT));
  EXA_CHECK(hipMemset(g->status.p, 0, sizeof(int) * nc));
  EXA_CHECK(hipMemset(g->slot_of.p, 0xff, sizeof(int) * nc));
  EXA_TRY(exa_amr_tables_dev(g->device, g->N, &g->amr_tab));
  EXA_TRY(exa_lim_tables_dev(g->device, g->N, &g->lim_tab));
  // LDS limits of the list kernels
  hipError_t e = hipSuccess;
  EXA_TRY(with_patch(g->P, [&](auto p) {
    constexpr int PP = decltype(p)::value;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_project_fv_list<PP>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)proj_lds(g->B));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_project_dg_list<PP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)proj_lds(g->B));
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_muscl_list<PP>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)muscl_lds(PP));
  }));
  EXA_CHECK(e);
  EXA_CHECK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  *out = guard.release();
  return EXA_OK;
}

extern "C" int exa_amr_grid_destroy(exa_amr_grid* g) {
  if (!g) return EXA_OK;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  release_graphs(g);
  clear_events(g);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return EXA_OK;
}

extern "C" int exa_amr_grid_num_cells(const exa_amr_grid* g, long long* n) {
  EXA_ARG(g && n, "null pointer");
  *n = g->nc;
  return EXA_OK;
}

extern "C" int exa_amr_grid_cells(const exa_amr_grid* g, int* cells) {
  EXA_ARG(g && cells, "null pointer");
  for (long long q = 0; q < g->nc; ++q) {
    cells[4 * q + 0] = g->leaves[q].l;
    cells[4 * q + 1] = g->leaves[q].x;
    cells[4 * q + 2] = g->leaves[q].y;
    cells[4 * q + 3] = g->leaves[q].z;
  }
  return EXA_OK;
}

extern "C" int exa_amr_grid_upload(exa_amr_grid* g, const double* luh) {
  EXA_ARG(g && luh, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  const int B3 = g->B * g->B * g->B;
  std::vector<double> h((size_t)g->nc * V * B3);
  for (long long q = 0; q < g->nc; ++q)
    for (int n = 0; n < B3; ++n)
      for (int v = 0; v < V; ++v) h[(size_t)q * V * B3 + v * B3 + n] = luh[((size_t)q * B3 + n) * V + v];
  EXA_CHECK(hipStreamSynchronize(g->stream));
  EXA_CHECK(hipMemcpy(g->luh.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  g->dt_valid = false;
  return limiter_start(g, nullptr);
}

extern "C" int exa_amr_grid_download(exa_amr_grid* g, double* luh) {
  EXA_ARG(g && luh, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  const int B3 = g->B * g->B * g->B;
  std::vector<double> h((size_t)g->nc * V * B3);
  EXA_CHECK(hipMemcpy(h.data(), g->luh.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  for (long long q = 0; q < g->nc; ++q)
    for (int n = 0; n < B3; ++n)
      for (int v = 0; v < V; ++v) luh[((size_t)q * B3 + n) * V + v] = h[(size_t)q * V * B3 + v * B3 + n];
  return EXA_OK;
}

extern "C" int exa_amr_grid_set_initial_conditions(exa_amr_grid* g, const exa_ic* ic) {
  EXA_ARG(g && ic, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  const double gamma = exa_solver_cfg(g->s)->gamma;
  EXA_TRY(with_order(g->N, [&](auto n) {
    constexpr int NN = decltype(n)::value;
    hipLaunchKernelGGL(k_amr_ic<NN>, dim3((unsigned)g->nc), dim3(Ord<NN>::NT), 0, g->stream, g->luh.as<double>(),
                       g->geo.as<int4>(), g->hbuf.as<double>(), *ic, gamma);
  }));
  EXA_CHECK(hipGetLastError());
  g->dt_valid = false;
  return limiter_start(g, ic);
}

extern "C" int exa_amr_grid_compute_dt(exa_amr_grid* g, double* dt) {
  EXA_ARG(g, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  hipLaunchKernelGGL(k_dt_reset_all, dim3(1), dim3(64), 0, g->stream, g->ip(), g->levels);
  EXA_TRY(enqueue_dt(g));
  hipLaunchKernelGGL(k_dt_take_all, dim3(1), dim3(1), 0, g->stream, g->ip(), g->levels);
  EXA_CHECK(hipGetLastError());
  InstParams p0;
  EXA_CHECK(hipMemcpyAsync(&p0, g->inst.p, sizeof(p0), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  g->dt_valid = true;
  if (dt) *dt = p0.dt_cur;
  return exa_check_dt(p0.dt_cur, "exa_amr_grid_compute_dt");
}

extern "C" int exa_amr_grid_use_graph(exa_amr_grid* g, int enable) {
  EXA_ARG(g, "null pointer");
  g->use_graph = enable != 0;
  return EXA_OK;
}

extern "C" int exa_amr_grid_enable_timing(exa_amr_grid* g, int enable) {
  EXA_ARG(g, "null pointer");
  g->timing = enable != 0;
  return EXA_OK;
}

extern "C" int exa_amr_grid_step(exa_amr_grid* g, int nsteps) {
  EXA_ARG(g && nsteps >= 0, "nsteps");
  EXA_ARG(g->dt_valid, "exa_amr_grid_compute_dt must run after setting the solution");
  EXA_CHECK(hipSetDevice(g->device));
  for (int i = 0; i < nsteps; ++i) {
    if (g->use_graph && !g->timing) {
      const int p = g->par;
      if (!g->gexec[p]) {
        EXA_CHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        const int r = enqueue_step(g, p, false);
        hipGraph_t graph = nullptr;
        const hipError_t he = hipStreamEndCapture(g->stream, &graph);
        if (r != EXA_OK) return r;
        EXA_CHECK(he);
        g->graph[p] = graph;
        EXA_CHECK(hipGraphInstantiate(&g->gexec[p], graph, nullptr, nullptr, 0));
      }
      EXA_CHECK(hipGraphLaunch(g->gexec[p], g->stream));
    } else {
      EXA_TRY(enqueue_step(g, g->par, g->timing));
      if (g->timing) ++g->timed_steps;
    }
    g->par ^= 1;
  }
  return EXA_OK;
}

extern "C" int exa_amr_grid_sync(exa_amr_grid* g, double* time, double* next_dt) {
  EXA_ARG(g, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  InstParams p0;
  EXA_CHECK(hipMemcpy(&p0, g->inst.p, sizeof(p0), hipMemcpyDeviceToHost));
  unsigned long long c[NCNT];
  EXA_CHECK(hipMemcpy(c, g->cnt.p, sizeof(c), hipMemcpyDeviceToHost));
  if (time) *time = p0.time;
  if (next_dt) *next_dt = p0.dt_cur;
  if (c[CNT_PATCH_OVERFLOW]) {
    exa_set_error("exa_amr_grid_sync: %llu cells of the limiter domain found no free limiter patch (capacity %lld)",
                  c[CNT_PATCH_OVERFLOW], g->cap);
    return EXA_ERR_NOMEM;
  }
  return g->dt_valid ? exa_check_dt(p0.dt_cur, "exa_amr_grid_sync") : EXA_OK;
}

extern "C" int exa_amr_grid_limiter_status(exa_amr_grid* g, int* status) {
  EXA_ARG(g && status, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  EXA_CHECK(hipMemcpy(status, g->status.p, sizeof(int) * g->nc, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_amr_grid_patches(exa_amr_grid* g, double* patches, int* cells, long long capacity, long long* n) {
  EXA_ARG(g && n, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  std::vector<int> slot(g->nc);
  EXA_CHECK(hipMemcpy(slot.data(), g->slot_of.p, sizeof(int) * g->nc, hipMemcpyDeviceToHost));
  const int S = g->S, P = g->P, S3 = S * S * S;
  std::vector<double> buf((size_t)V * S3);
  long long k = 0;
  for (long long q = 0; q < g->nc; ++q) {
    if (slot[q] < 0) continue;
    if (k < capacity && patches && cells) {
      EXA_CHECK(hipMemcpy(buf.data(), g->pool[g->par].as<double>() + (size_t)slot[q] * V * S3,
                          sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
      double* o = patches + (size_t)k * V * P * P * P;
      for (int z = 0; z < P; ++z)
        for (int y = 0; y < P; ++y)
          for (int x = 0; x < P; ++x)
            for (int v = 0; v < V; ++v)
              o[((z * P + y) * P + x) * V + v] =
                  buf[(size_t)v * S3 + ((z + GHOST) * S + y + GHOST) * S + x + GHOST];
      cells[k] = (int)q;
    }
    ++k;
  }
  *n = k;
  return EXA_OK;
}

extern "C" int exa_amr_grid_counters(exa_amr_grid* g, long long counts[8]) {
  EXA_ARG(g && counts, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  unsigned long long c[NCNT];
  EXA_CHECK(hipMemcpy(c, g->cnt.p, sizeof(c), hipMemcpyDeviceToHost));
  for (int i = 0; i < NCNT; ++i) counts[i] = (long long)c[i];
  return EXA_OK;
}

extern "C" int exa_amr_grid_phase_times(exa_amr_grid* g, double ms[6], long long* steps) {
  EXA_ARG(g && ms, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  for (int i = 0; i < 6; ++i) {
    double t = 0.0;
    for (auto& p : g->ev[i]) {
      float f = 0.f;
      EXA_CHECK(hipEventElapsedTime(&f, p.first, p.second));
      t += f;
    }
    ms[i] = t;
  }
  if (steps) *steps = g->timed_steps;
  return EXA_OK;
}

extern "C" int exa_amr_grid_total_picard_iterations(exa_amr_grid* g, long long* total) {
  EXA_ARG(g && total, "null pointer");
  EXA_CHECK(hipSetDevice(g->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  unsigned long long t = 0;
  EXA_CHECK(hipMemcpy(&t, g->iter_sum.p, sizeof(t), hipMemcpyDeviceToHost));
  *total = (long long)t;
  return EXA_OK;
}
