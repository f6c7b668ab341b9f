// Kernels of the generic PDE arms: linear advection in flux and in
// non-conservative form, Euler with the test non-conservative product.
#include <cstring>
#include <mutex>

#include "aderdg_pde_kernels.h"
#include "predictor_tu.h"

namespace {
template <class PDE, int N>
PdePredFn pred(bool debug) {
  return debug ? &exa::k_predictor_pde<N, PDE, true> : &exa::k_predictor_pde<N, PDE, false>;
}
template <class PDE, int N>
PdeUpdateFn upd(int mode) {
  constexpr int FULL = exa::UPD_SURFACE | exa::UPD_UPDATE | exa::UPD_DT;
  if (mode == exa::UPD_DT) return &exa::k_update<N, exa::UPD_DT, PDE>;
  if (mode == FULL) return &exa::k_update<N, FULL, PDE>;
  if (mode == (FULL | exa::UPD_LINEAR)) return &exa::k_update<N, FULL | exa::UPD_LINEAR, PDE>;
  return nullptr;
}
template <class PDE, int N>
PdePredFn lin_pred(bool debug, bool ps) {
  if (ps) return debug ? &exa::k_predictor_linear<N, PDE, true, true> : &exa::k_predictor_linear<N, PDE, false, true>;
  return debug ? &exa::k_predictor_linear<N, PDE, true> : &exa::k_predictor_linear<N, PDE, false>;
}

// F(PDE, N) for the linear family: the PDEs with a coefficient matrix, orders 1..9
#define EXA_LINEAR_DISPATCH(pde, order, F)                                              \
  do {                                                                                  \
    switch (pde) {                                                                      \
      case EXA_PDE_ADVECTION_NCP: EXA_LINEAR_ORDERS(exa::PdeAdvectionNcp, order, F); break; \
      case EXA_PDE_EULER_NCP_TEST: EXA_LINEAR_ORDERS(exa::PdeEulerNcpTest, order, F); break; \
      default: break;                                                                   \
    }                                                                                   \
  } while (0)
#define EXA_LINEAR_ORDERS(P, order, F) \
  switch (order) {                     \
    case 1: F(P, 1); break;            \
    case 2: F(P, 2); break;            \
    case 3: F(P, 3); break;            \
    case 4: F(P, 4); break;            \
    case 5: F(P, 5); break;            \
    case 6: F(P, 6); break;            \
    case 7: F(P, 7); break;            \
    case 8: F(P, 8); break;            \
    case 9: F(P, 9); break;            \
    default: break;                    \
  }

// F(PDE, N) for the PDE of an id and an order 1..7
#define EXA_PDE_DISPATCH(pde, order, F)                                     \
  do {                                                                      \
    switch (pde) {                                                          \
      case EXA_PDE_ADVECTION: EXA_PDE_ORDERS(exa::PdeAdvection, order, F); break; \
      case EXA_PDE_ADVECTION_NCP: EXA_PDE_ORDERS(exa::PdeAdvectionNcp, order, F); break; \
      case EXA_PDE_EULER_NCP_TEST: EXA_PDE_ORDERS(exa::PdeEulerNcpTest, order, F); break; \
      default: break;                                                       \
    }                                                                       \
  } while (0)
#define EXA_PDE_ORDERS(P, order, F)          \
  switch (order) {                           \
    case 1: F(P, 1); break;                  \
    case 2: F(P, 2); break;                  \
    case 3: F(P, 3); break;                  \
    case 4: F(P, 4); break;                  \
    case 5: F(P, 5); break;                  \
    case 6: F(P, 6); break;                  \
    case 7: F(P, 7); break;                  \
    default: break;                          \
  }
}  // namespace

// ---- PDEs registered by a project (exahype_amd_pde.h) ----
namespace {
struct UserSlot {
  bool used = false;
  char name[64] = {0};
  exa_pde_kernels k{};
};
UserSlot g_user[EXA_PDE_USER_SLOTS];
std::mutex g_user_mu;

const exa_pde_kernels* user_kernels(int pde, int order) {
  const int i = pde - EXA_PDE_USER_FIRST;
  if (i < 0 || i >= EXA_PDE_USER_SLOTS) return nullptr;
  std::lock_guard<std::mutex> lk(g_user_mu);
  if (!g_user[i].used || order < 1 || order > g_user[i].k.max_order) return nullptr;
  return &g_user[i].k;
}
}  // namespace

bool exa_pde_user(int pde) {
  const int i = pde - EXA_PDE_USER_FIRST;
  if (i < 0 || i >= EXA_PDE_USER_SLOTS) return false;
  std::lock_guard<std::mutex> lk(g_user_mu);
  return g_user[i].used;
}

int exa_pde_user_max_order(int pde) {
  const int i = pde - EXA_PDE_USER_FIRST;
  if (i < 0 || i >= EXA_PDE_USER_SLOTS) return 0;
  std::lock_guard<std::mutex> lk(g_user_mu);
  return g_user[i].used ? g_user[i].k.max_order : 0;
}

int exa_pde_user_add(const char* name, const exa_pde_kernels* k) {
  std::lock_guard<std::mutex> lk(g_user_mu);
  int free_slot = -1;
  for (int i = 0; i < EXA_PDE_USER_SLOTS; ++i) {
    if (g_user[i].used && std::strncmp(g_user[i].name, name, sizeof(g_user[i].name)) == 0) {
      g_user[i].k = *k;
      return i;
    }
    if (!g_user[i].used && free_slot < 0) free_slot = i;
  }
  if (free_slot < 0) return -1;
  g_user[free_slot].used = true;
  std::strncpy(g_user[free_slot].name, name, sizeof(g_user[free_slot].name) - 1);
  g_user[free_slot].k = *k;
  return free_slot;
}

int exa_pde_user_find(const char* name) {
  std::lock_guard<std::mutex> lk(g_user_mu);
  for (int i = 0; i < EXA_PDE_USER_SLOTS; ++i)
    if (g_user[i].used && std::strncmp(g_user[i].name, name, sizeof(g_user[i].name)) == 0) return i;
  return -1;
}

int exa_pde_user_init_all(const void* tables) {
  std::lock_guard<std::mutex> lk(g_user_mu);
  for (int i = 0; i < EXA_PDE_USER_SLOTS; ++i)
    if (g_user[i].used && g_user[i].k.init(tables) != 0) return -1;
  return 0;
}

int exa_pde_user_init(int pde, const void* tables) {
  const int i = pde - EXA_PDE_USER_FIRST;
  std::lock_guard<std::mutex> lk(g_user_mu);
  if (i < 0 || i >= EXA_PDE_USER_SLOTS || !g_user[i].used) return -1;
  return g_user[i].k.init(tables);
}

const void* exa_pde_user_kernel(int pde, int order, int which) {
  const exa_pde_kernels* k = user_kernels(pde, order);
  if (!k) return nullptr;
  switch (which) {
    case EXA_USER_IC: return k->initial_conditions[order];
    case EXA_USER_BOUNDARY_FACES: return k->boundary_faces[order];
    case EXA_USER_BOUNDARY_REF: return k->boundary_ref[order];
    default: return nullptr;
  }
}

bool exa_pde_generic(int pde) {
  return pde == EXA_PDE_ADVECTION || pde == EXA_PDE_ADVECTION_NCP || pde == EXA_PDE_EULER_NCP_TEST ||
         exa_pde_user(pde);
}

PdePredFn exa_pde_predictor(int pde, int order, bool debug) {
  if (const exa_pde_kernels* k = user_kernels(pde, order))
    return reinterpret_cast<PdePredFn>(const_cast<void*>(k->predictor[order][debug ? 1 : 0]));
  PdePredFn f = nullptr;
#define F_(P, N) f = pred<P, N>(debug)
  EXA_PDE_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

int exa_pde_predictor_lds_bytes(int order) {
  switch (order) {
    case 1: return exa::PdeOrd<1>::LDS_DOUBLES * (int)sizeof(double);
    case 2: return exa::PdeOrd<2>::LDS_DOUBLES * (int)sizeof(double);
    case 3: return exa::PdeOrd<3>::LDS_DOUBLES * (int)sizeof(double);
    case 4: return exa::PdeOrd<4>::LDS_DOUBLES * (int)sizeof(double);
    case 5: return exa::PdeOrd<5>::LDS_DOUBLES * (int)sizeof(double);
    case 6: return exa::PdeOrd<6>::LDS_DOUBLES * (int)sizeof(double);
    case 7: return exa::PdeOrd<7>::LDS_DOUBLES * (int)sizeof(double);
    default: return 0;
  }
}

PdeRiemannFn exa_pde_riemann(int pde, int order) {
  if (const exa_pde_kernels* k = user_kernels(pde, order))
    return reinterpret_cast<PdeRiemannFn>(const_cast<void*>(k->riemann[order]));
  PdeRiemannFn f = nullptr;
#define F_(P, N) f = &exa::k_riemann_pde<N, P>
  EXA_PDE_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

PdeRiemannListFn exa_pde_riemann_list(int pde, int order) {
  if (const exa_pde_kernels* k = user_kernels(pde, order))
    return reinterpret_cast<PdeRiemannListFn>(const_cast<void*>(k->riemann_list[order]));
  PdeRiemannListFn f = nullptr;
#define F_(P, N) f = &exa::k_riemann_list_pde<N, P>
  EXA_PDE_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

PdeUpdateFn exa_pde_update(int pde, int order, int mode) {
  if (const exa_pde_kernels* k = user_kernels(pde, order)) {
    constexpr int FULL = exa::UPD_SURFACE | exa::UPD_UPDATE | exa::UPD_DT;
    const void* f = mode == exa::UPD_DT ? k->update_dt[order] : (mode == FULL ? k->update_full[order] : nullptr);
    return reinterpret_cast<PdeUpdateFn>(const_cast<void*>(f));
  }
  PdeUpdateFn f = nullptr;
#define F_(P, N) f = upd<P, N>(mode)
  EXA_PDE_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

bool exa_pde_linear_supported(int pde) { return pde == EXA_PDE_ADVECTION_NCP || pde == EXA_PDE_EULER_NCP_TEST; }

PdePredFn exa_pde_linear_predictor(int pde, int order, bool debug, bool point_source) {
  PdePredFn f = nullptr;
#define F_(P, N) f = lin_pred<P, N>(debug, point_source)
  EXA_LINEAR_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

int exa_pde_linear_lds_bytes(int order) {
  const int B = order + 1;
  return 4 * exa::V * B * B * B * (int)sizeof(double);
}

PdeRiemannFn exa_pde_linear_riemann(int pde, int order) {
  PdeRiemannFn f = nullptr;
#define F_(P, N) f = &exa::k_riemann_linear<N, P>
  EXA_LINEAR_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

PdeUpdateFn exa_pde_linear_update(int pde, int order, int mode) {
  PdeUpdateFn f = nullptr;
#define F_(P, N) f = upd<P, N>(mode)
  EXA_LINEAR_DISPATCH(pde, order, F_);
#undef F_
  return f;
}

PointSourceFn exa_pde_point_source(int order) {
  switch (order) {
    case 1: return &exa::k_point_source<1>;
    case 2: return &exa::k_point_source<2>;
    case 3: return &exa::k_point_source<3>;
    case 4: return &exa::k_point_source<4>;
    case 5: return &exa::k_point_source<5>;
    case 6: return &exa::k_point_source<6>;
    case 7: return &exa::k_point_source<7>;
    case 8: return &exa::k_point_source<8>;
    case 9: return &exa::k_point_source<9>;
    default: return nullptr;
  }
}

LinearVolumeFn exa_pde_linear_volume(int order) {
  switch (order) {
    case 1: return &exa::k_volume_integral_linear<1>;
    case 2: return &exa::k_volume_integral_linear<2>;
    case 3: return &exa::k_volume_integral_linear<3>;
    case 4: return &exa::k_volume_integral_linear<4>;
    case 5: return &exa::k_volume_integral_linear<5>;
    case 6: return &exa::k_volume_integral_linear<6>;
    case 7: return &exa::k_volume_integral_linear<7>;
    case 8: return &exa::k_volume_integral_linear<8>;
    case 9: return &exa::k_volume_integral_linear<9>;
    default: return nullptr;
  }
}

hipError_t exa_pde_init(const DGTables* tab) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), tab, sizeof(DGTables) * (EXA_MAX_ORDER + 1));
  if (e != hipSuccess) return e;
  const int pdes[3] = {EXA_PDE_ADVECTION, EXA_PDE_ADVECTION_NCP, EXA_PDE_EULER_NCP_TEST};
  for (int pde : pdes)
    for (int order = 1; order <= exa::PDE_MAX_ORDER; ++order)
      for (int dbg = 0; dbg < 2; ++dbg) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(exa_pde_predictor(pde, order, dbg)),
                                hipFuncAttributeMaxDynamicSharedMemorySize, exa_pde_predictor_lds_bytes(order));
        if (e != hipSuccess) return e;
      }
  const int lin[2] = {EXA_PDE_ADVECTION_NCP, EXA_PDE_EULER_NCP_TEST};
  for (int pde : lin)
    for (int order = 1; order <= exa::LINEAR_MAX_ORDER; ++order)
      for (int dbg = 0; dbg < 2; ++dbg)
        for (int ps = 0; ps < 2; ++ps) {
          e = hipFuncSetAttribute(reinterpret_cast<const void*>(exa_pde_linear_predictor(pde, order, dbg, ps)),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, exa_pde_linear_lds_bytes(order));
          if (e != hipSuccess) return e;
        }
  return hipSuccess;
}
