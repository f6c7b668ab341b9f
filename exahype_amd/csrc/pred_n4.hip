// Predictor kernels of order 4.
#define EXA_PRED_ORDER 4
#include "pred_impl.h"
