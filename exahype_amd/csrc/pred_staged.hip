// Predictor kernels with the space-time iterate staged in global memory:
// the only form of orders 8 and 9, selectable (EXA_CONTRACT_STAGED) at 5..7;
// FP64, FP64 with the gravity source, and mixed (an FP32 phase).
#include "aderdg_kernels.h"
#include "predictor_tu.h"

namespace {
template <int N>
StagedFn pick(bool debug, bool src, bool mixed) {
  if (src) return debug ? &exa::k_predictor_staged<N, true, true> : &exa::k_predictor_staged<N, false, true>;
  if (mixed)
    return debug ? &exa::k_predictor_staged<N, true, false, true> : &exa::k_predictor_staged<N, false, false, true>;
  return debug ? &exa::k_predictor_staged<N, true> : &exa::k_predictor_staged<N, false>;
}
}  // namespace

StagedFn exa_pred_staged_kernel(int order, bool debug, bool src, bool mixed) {
  if (src && mixed) return nullptr;
  switch (order) {
    case 5: return pick<5>(debug, src, mixed);
    case 6: return pick<6>(debug, src, mixed);
    case 7: return pick<7>(debug, src, mixed);
    case 8: return pick<8>(debug, src, mixed);
    case 9: return pick<9>(debug, src, mixed);
    default: return nullptr;
  }
}

int exa_pred_staged_lds_bytes(int order) {
  switch (order) {
    case 5: return exa::Ord<5>::LDS_STAGED * (int)sizeof(double);
    case 6: return exa::Ord<6>::LDS_STAGED * (int)sizeof(double);
    case 7: return exa::Ord<7>::LDS_STAGED * (int)sizeof(double);
    case 8: return exa::Ord<8>::LDS_STAGED * (int)sizeof(double);
    case 9: return exa::Ord<9>::LDS_STAGED * (int)sizeof(double);
    default: return 0;
  }
}

hipError_t exa_pred_staged_init(const DGTables* tab, const DGTablesF* tabf) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), tab, sizeof(DGTables) * (EXA_MAX_ORDER + 1));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), tabf, sizeof(DGTablesF) * (EXA_MAX_ORDER + 1));
  if (e != hipSuccess) return e;
  for (int order = EXA_STAGED_MIN_ORDER; order <= EXA_MAX_ORDER; ++order)
    for (int dbg = 0; dbg < 2; ++dbg)
      for (int var = 0; var < 3; ++var) {  // FP64, FP64 with the gravity source, mixed
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(exa_pred_staged_kernel(order, dbg, var == 1, var == 2)),
                                hipFuncAttributeMaxDynamicSharedMemorySize, exa_pred_staged_lds_bytes(order));
        if (e != hipSuccess) return e;
      }
  return hipSuccess;
}
