// Predictor kernels of order 7.
#define EXA_PRED_ORDER 7
#include "pred_impl.h"
