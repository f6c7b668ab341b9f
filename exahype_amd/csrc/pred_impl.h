// Body of one per-order predictor translation unit; EXA_PRED_ORDER is set by
// the including pred_n<N>.hip.
#include "aderdg_kernels.h"
#include "predictor_tu.h"

#define EXA_CAT2(a, b) a##b
#define EXA_CAT(a, b) EXA_CAT2(a, b)

namespace {
constexpr int NP = EXA_PRED_ORDER;

template <bool DEBUG, int WPE, bool NODE>
PredFn by_prec(int prec) {
  switch (prec) {
    case EXA_PREC_FP32: return &exa::k_predictor<NP, DEBUG, exa::PRED_FP32, WPE, NODE>;
    case EXA_PREC_FP32_MFMA: return &exa::k_predictor<NP, DEBUG, exa::PRED_FP32_MFMA, WPE, NODE>;
    case EXA_PREC_BF16_MFMA: return &exa::k_predictor<NP, DEBUG, exa::PRED_BF16_MFMA, WPE, NODE>;
    default: return &exa::k_predictor<NP, DEBUG, exa::PRED_FP64, WPE, NODE>;
  }
}
}  // namespace

PredFn EXA_CAT(exa_pred_kernel_, EXA_PRED_ORDER)(bool debug, int prec, int wpe, bool node) {
  if (debug) return node ? by_prec<true, 1, true>(prec) : by_prec<true, 1, false>(prec);
  constexpr int W2 = NP <= 5 ? 2 : 1;
  if (wpe == 2) return node ? by_prec<false, W2, true>(prec) : by_prec<false, W2, false>(prec);
  return node ? by_prec<false, 1, true>(prec) : by_prec<false, 1, false>(prec);
}

hipError_t EXA_CAT(exa_pred_init_, EXA_PRED_ORDER)(const DGTables* tab, const DGTablesF* tabf) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), tab, sizeof(DGTables) * (EXA_MAX_ORDER + 1));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), tabf, sizeof(DGTablesF) * (EXA_MAX_ORDER + 1));
  if (e != hipSuccess) return e;
  const int bytes = exa::Ord<NP>::LDS_DOUBLES * (int)sizeof(double);
  const int precs[4] = {EXA_PREC_FP64, EXA_PREC_FP32, EXA_PREC_FP32_MFMA, EXA_PREC_BF16_MFMA};
  for (int dbg = 0; dbg < 2; ++dbg)
    for (int w = 1; w <= 2; ++w)
      for (int node = 0; node < 2; ++node)
        for (int p : precs) {
          e = hipFuncSetAttribute(
              reinterpret_cast<const void*>(EXA_CAT(exa_pred_kernel_, EXA_PRED_ORDER)(dbg, p, w, node)),
              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
          if (e != hipSuccess) return e;
        }
  return hipSuccess;
}
