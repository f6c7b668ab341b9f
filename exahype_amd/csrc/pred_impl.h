// Body of one per-order predictor translation unit; EXA_PRED_ORDER is set by
// the including pred_n<N>.hip.
#include "aderdg_kernels.h"
#include "predictor_tu.h"

#define EXA_CAT2(a, b) a##b
#define EXA_CAT(a, b) EXA_CAT2(a, b)

namespace {
constexpr int NP = EXA_PRED_ORDER;

template <bool DEBUG, int WPE, int FORM>
PredFn by_prec(int prec) {
  switch (prec) {
    case EXA_PREC_FP32: return &exa::k_predictor<NP, DEBUG, exa::PRED_FP32, WPE, FORM>;
    case EXA_PREC_FP32_MFMA: return &exa::k_predictor<NP, DEBUG, exa::PRED_FP32_MFMA, WPE, FORM>;
    case EXA_PREC_BF16_MFMA: return &exa::k_predictor<NP, DEBUG, exa::PRED_BF16_MFMA, WPE, FORM>;
    default: return &exa::k_predictor<NP, DEBUG, exa::PRED_FP64, WPE, FORM>;
  }
}

template <bool DEBUG, int WPE>
PredFn by_form(int prec, int form) {
  switch (form) {
    case exa::FORM_NODE: return by_prec<DEBUG, WPE, exa::FORM_NODE>(prec);
    case exa::FORM_PENCIL_PAIR: return by_prec<DEBUG, WPE, exa::FORM_PENCIL_PAIR>(prec);
    default: return by_prec<DEBUG, WPE, exa::FORM_PENCIL>(prec);
  }
}

// FP32 sweeps, the uniform-cell variant
template <int WPE>
PredFn uskip_by_form(int form) {
  switch (form) {
    case exa::FORM_NODE: return &exa::k_predictor<NP, false, exa::PRED_FP32, WPE, exa::FORM_NODE, false, true>;
    case exa::FORM_PENCIL_PAIR:
      return &exa::k_predictor<NP, false, exa::PRED_FP32, WPE, exa::FORM_PENCIL_PAIR, false, true>;
    default: return &exa::k_predictor<NP, false, exa::PRED_FP32, WPE, exa::FORM_PENCIL, false, true>;
  }
}
}  // namespace

PredFn EXA_CAT(exa_pred_kernel_, EXA_PRED_ORDER)(bool debug, int prec, int wpe, int form, bool uskip) {
  constexpr int W2 = NP <= 5 ? 2 : 1;
  if (uskip && !debug && prec == EXA_PREC_FP32) return wpe == 2 ? uskip_by_form<W2>(form) : uskip_by_form<1>(form);
  if (debug) return by_form<true, 1>(prec, form);
  if (wpe == 2) return by_form<false, W2>(prec, form);
  return by_form<false, 1>(prec, form);
}

// The gravity-source variant: FP64, pencil owners, one wave per SIMD.
PredFn EXA_CAT(exa_pred_kernel_src_, EXA_PRED_ORDER)(bool debug) {
  if (debug) return &exa::k_predictor<NP, true, exa::PRED_FP64, 1, exa::FORM_PENCIL, true>;
  return &exa::k_predictor<NP, false, exa::PRED_FP64, 1, exa::FORM_PENCIL, true>;
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
      for (int form = 0; form < 3; ++form)
        for (int p : precs) {
          e = hipFuncSetAttribute(
              reinterpret_cast<const void*>(EXA_CAT(exa_pred_kernel_, EXA_PRED_ORDER)(dbg, p, w, form)),
              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
          if (e != hipSuccess) return e;
        }
  for (int w = 1; w <= 2; ++w)
    for (int form = 0; form < 3; ++form) {
      e = hipFuncSetAttribute(
          reinterpret_cast<const void*>(EXA_CAT(exa_pred_kernel_, EXA_PRED_ORDER)(false, EXA_PREC_FP32, w, form, true)),
          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
      if (e != hipSuccess) return e;
    }
  for (int dbg = 0; dbg < 2; ++dbg) {
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(EXA_CAT(exa_pred_kernel_src_, EXA_PRED_ORDER)(dbg)),
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
