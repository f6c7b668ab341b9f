// Internal interfaces between the library's translation units: what the
// adaptive-mesh step (amr_grid.hip) takes from the solver, the AMR transfer
// and the limiter units.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "predictor_tu.h"

// capi.hip: the register-form Euler predictor a solver launches (orders 1..7)
int exa_solver_predictor(const exa_solver* s, PredFn* fn, int* threads, size_t* lds);
const exa_aderdg_config* exa_solver_cfg(const exa_solver* s);
int exa_solver_device(const exa_solver* s);

// amr_capi.hip: device tables of one order; transfers of a device-held list
// (items: exa::amr::Transfer, groups as in the batched entry points) on a stream
int exa_amr_tables_dev(int dev, int order, const void** tab);
int exa_amr_enqueue(int dim, bool prolong, int order, const void* tab, const void* items, const int* groups,
                    int ngroups, double* fine, double* coarse, int nd, hipStream_t st);

// limiter_capi.hip: projection / min-max tables of one order (exa::lim::LimTables)
int exa_lim_tables_dev(int device, int N, const void** out);

// amr_grid.hip: uploads the DG tables its kernels read (called by exa_init)
hipError_t exa_amr_grid_init(const DGTables* tab);
