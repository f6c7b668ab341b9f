// Predictor kernels of orders 8 and 9 (iterate staged in global memory).
#include "aderdg_kernels.h"
#include "predictor_tu.h"

StagedFn exa_pred_staged_kernel(int order, bool debug) {
  switch (order) {
    case 8: return debug ? &exa::k_predictor_staged<8, true> : &exa::k_predictor_staged<8, false>;
    case 9: return debug ? &exa::k_predictor_staged<9, true> : &exa::k_predictor_staged<9, false>;
    default: return nullptr;
  }
}

int exa_pred_staged_lds_bytes(int order) {
  switch (order) {
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
  for (int order = 8; order <= 9; ++order)
    for (int dbg = 0; dbg < 2; ++dbg) {
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(exa_pred_staged_kernel(order, dbg)),
                              hipFuncAttributeMaxDynamicSharedMemorySize, exa_pred_staged_lds_bytes(order));
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}
