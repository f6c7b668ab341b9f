// Interface of the per-order translation units that instantiate the
// predictor kernels (pred_n<N>.hip): compiling the orders in parallel units
// keeps the build short.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace exa {
struct PredictorDebug;
struct Ens;
}  // namespace exa

using PredFn = void (*)(const double*, double*, double*, double*, int*, const double*, double, double, double, double,
                        int, unsigned long long*, exa::PredictorDebug, exa::Ens);

// Kernel of one variant: per-cell debug outputs (one wave per SIMD only),
// precision EXA_PREC_*, waves per SIMD (2 is built for N <= 5), contraction
// form exa::FORM_* (pencil owners, node owners, pencil owners two slices at
// a time); uskip: with FP32 sweeps, the variant in which cells uniform to
// FP32 resolution take the FP64 sweeps alone.  exa_pred_kernel_src_N: the
// gravity-source variant (FP64, pencil owners, one wave per SIMD).
#define EXA_PRED_TU_DECL(N)                                                   \
  PredFn exa_pred_kernel_##N(bool debug, int prec, int wpe, int form, bool uskip = false); \
  PredFn exa_pred_kernel_src_##N(bool debug);                                 \
  hipError_t exa_pred_init_##N(const DGTables* tab, const DGTablesF* tabf);
EXA_PRED_TU_DECL(1)
EXA_PRED_TU_DECL(2)
EXA_PRED_TU_DECL(3)
EXA_PRED_TU_DECL(4)
EXA_PRED_TU_DECL(5)
EXA_PRED_TU_DECL(6)
EXA_PRED_TU_DECL(7)
#undef EXA_PRED_TU_DECL

// Staged predictor (pred_staged.hip; orders 5..9, the only form of 8 and 9):
// the space-time iterate lives in a staging slot of global memory per
// workgroup (2 B^4 V doubles), the grid is persistent (one workgroup per
// slot, cells strided by the grid size).
#define EXA_STAGED_MIN_ORDER 5
using StagedFn = void (*)(const double*, double*, double*, double*, int*, const double*, double, double, double,
                          double, long long, double*, unsigned long long*, exa::PredictorDebug, exa::Ens, int);
// src: the gravity source (FP64); mixed: an FP32 phase before the last argument's FP64 sweeps
StagedFn exa_pred_staged_kernel(int order, bool debug, bool src = false, bool mixed = false);
int exa_pred_staged_lds_bytes(int order);
hipError_t exa_pred_staged_init(const DGTables* tab, const DGTablesF* tabf);

// Kernels of the generic PDE arms (pde_tu.hip): EXA_PDE_ADVECTION,
// EXA_PDE_ADVECTION_NCP, EXA_PDE_EULER_NCP_TEST at orders 1..7; nullptr for
// any other PDE or order.
using PdePredFn = void (*)(const double*, double*, double*, double*, int*, const double*, double, double, double,
                           double, unsigned long long*, exa::PredictorDebug, exa::Ens);
using PdeRiemannFn = void (*)(const double*, double*, const int*, int, double, long long, long long, exa::Ens);
using PdeRiemannListFn = void (*)(const double*, double*, const int*, long long, double, exa::Ens);
using PdeUpdateFn = void (*)(double*, double*, const double*, const double*, unsigned long long*, double, double,
                             double, double, double, exa::Ens);
// pointSource of one cell (k_point_source) at orders 1..9
using PointSourceFn = void (*)(double, double, double, double, double, double, double, double, double*, exa::Ens);
PointSourceFn exa_pde_point_source(int order);
bool exa_pde_generic(int pde);
PdePredFn exa_pde_predictor(int pde, int order, bool debug);
int exa_pde_predictor_lds_bytes(int order);
PdeRiemannFn exa_pde_riemann(int pde, int order);
PdeRiemannListFn exa_pde_riemann_list(int pde, int order);
// modes with UPD_DT (the only ones that depend on the PDE): UPD_DT and
// UPD_SURFACE | UPD_UPDATE | UPD_DT
PdeUpdateFn exa_pde_update(int pde, int order, int mode);
// The linear kernel family (Cauchy-Kowalevski predictor with the linear
// volume integral, coefficient-matrix Riemann, fluctuation-form surface
// integral via UPD_LINEAR): EXA_PDE_ADVECTION_NCP and EXA_PDE_EULER_NCP_TEST
// at orders 1..9; nullptr otherwise.
using LinearVolumeFn = void (*)(const double*, double*);
bool exa_pde_linear_supported(int pde);
PdePredFn exa_pde_linear_predictor(int pde, int order, bool debug, bool point_source = false);
int exa_pde_linear_lds_bytes(int order);
PdeRiemannFn exa_pde_linear_riemann(int pde, int order);
// modes UPD_DT and UPD_SURFACE | UPD_UPDATE | UPD_DT | UPD_LINEAR
PdeUpdateFn exa_pde_linear_update(int pde, int order, int mode);
LinearVolumeFn exa_pde_linear_volume(int order);
hipError_t exa_pde_init(const DGTables* tab);
// Registered PDEs (exahype_amd_pde.h): ids EXA_PDE_USER_FIRST + slot.  Their
// predictor, Riemann and update kernels come through the exa_pde_* getters
// above; the rest through exa_pde_user_kernel.
bool exa_pde_user(int pde);
int exa_pde_user_max_order(int pde);
int exa_pde_user_add(const char* name, const exa_pde_kernels* k);  // slot, or -1 when full
int exa_pde_user_find(const char* name);                            // slot, or -1
int exa_pde_user_init_all(const void* tables);                      // every slot, current device
int exa_pde_user_init(int pde, const void* tables);
enum { EXA_USER_IC = 0, EXA_USER_BOUNDARY_FACES = 1, EXA_USER_BOUNDARY_REF = 2 };
const void* exa_pde_user_kernel(int pde, int order, int which);
