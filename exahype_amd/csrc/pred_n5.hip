// Predictor kernels of order 5.
#define EXA_PRED_ORDER 5
#include "pred_impl.h"
