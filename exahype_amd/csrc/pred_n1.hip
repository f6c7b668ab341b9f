// Predictor kernels of order 1.
#define EXA_PRED_ORDER 1
#include "pred_impl.h"
