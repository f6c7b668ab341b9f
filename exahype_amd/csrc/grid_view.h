// Internal view of a device grid for the units that add kernels over it
// (the limiter), without exposing the grid's definition.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/exahype_amd.h"

struct ExaGridView {
  double* luh;        // [cell][v][node], cells ordered (z,y,x)
  long long nc;
  int nx, ny, nz;
  int order;
  int device;
  hipStream_t stream;
  long long version;  // changes whenever luh changes
  bool periodic;      // every axis periodic (exa_grid_set_boundary not used)
};

int exa_grid_view(exa_grid* g, ExaGridView* v);
// Frees whatever the limiter unit keeps for grid g (called by exa_grid_destroy).
void exa_grid_limiter_release(exa_grid* g);
