// Predictor kernels of order 6.
#define EXA_PRED_ORDER 6
#include "pred_impl.h"
