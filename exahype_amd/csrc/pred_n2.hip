// Predictor kernels of order 2.
#define EXA_PRED_ORDER 2
#include "pred_impl.h"
