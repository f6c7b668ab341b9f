// Predictor kernels of order 3.
#define EXA_PRED_ORDER 3
#include "pred_impl.h"
