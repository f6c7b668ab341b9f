// Operand types and conversions of the matrix-core (MFMA) paths.
#pragma once
#include <hip/hip_runtime.h>

namespace exa {

typedef short exa_s4 __attribute__((ext_vector_type(4)));  // four BF16 bit patterns
typedef float exa_f4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bit pattern, round to nearest even (finite inputs).
__device__ inline short f32_to_bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

}  // namespace exa
