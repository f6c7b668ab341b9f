// Device memory helpers of the C ABI: allocation, host<->device copies and a
// streaming device copy kernel.  They let a caller without its own HIP
// allocator (a host code in another language, the tests) drive the batched
// entry points, which take device pointers; the copy kernel is the HBM
// reference of the benchmarks (bytes moved per second by a plain stream).
#include <hip/hip_runtime.h>

#include "host_util.h"

namespace {

// 32 bytes per lane per iteration, grid-stride; the tail (bytes % 32) is
// copied byte by byte by the first lanes.
__global__ void __launch_bounds__(256) k_stream_copy(double4* __restrict__ dst, const double4* __restrict__ src,
                                                     long long n4, unsigned char* __restrict__ dtail,
                                                     const unsigned char* __restrict__ stail, int ntail) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

}  // namespace

extern "C" int exa_device_malloc(void** ptr, long long bytes) {
  EXA_ARG(ptr && bytes > 0, "pointer and positive size");
  EXA_CHECK(hipMalloc(ptr, (size_t)bytes));
  return EXA_OK;
}

extern "C" int exa_device_free(void* ptr) {
  if (!ptr) return EXA_OK;
  EXA_CHECK(hipFree(ptr));
  return EXA_OK;
}

extern "C" int exa_memcpy_htod(void* dst, const void* src, long long bytes) {
  EXA_ARG(dst && src && bytes >= 0, "arguments");
  if (bytes) EXA_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return EXA_OK;
}

extern "C" int exa_memcpy_dtoh(void* dst, const void* src, long long bytes) {
  EXA_ARG(dst && src && bytes >= 0, "arguments");
  if (bytes) EXA_CHECK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return EXA_OK;
}

extern "C" int exa_device_copy(void* dst, const void* src, long long bytes, float* ms) {
  EXA_ARG(dst && src && bytes >= 0, "arguments");
  EXA_ARG(((uintptr_t)dst % 32) == 0 && ((uintptr_t)src % 32) == 0, "pointers must be 32-byte aligned");
  const long long n4 = bytes / 32;
  const int ntail = (int)(bytes % 32);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EXA_CHECK(hipEventCreate(&e0));
  EXA_CHECK(hipEventCreate(&e1));
  int dev = 0, ncu = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const long long want = (n4 + 255) / 256;
  const unsigned grid = (unsigned)(want < 1 ? 1 : (want > 16LL * ncu ? 16LL * ncu : want));
  EXA_CHECK(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, nullptr, static_cast<double4*>(dst),
                     static_cast<const double4*>(src), n4, static_cast<unsigned char*>(dst) + 32 * n4,
                     static_cast<const unsigned char*>(src) + 32 * n4, ntail);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipEventRecord(e1, nullptr));
  EXA_CHECK(hipEventSynchronize(e1));
  float t = 0.f;
  EXA_CHECK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Self-test of the matrix-core operand layouts the predictor's low-precision
// contraction relies on: D = A * B for 16x16 row-major A, B with the lane
// mapping A[l%16][4(l/16)+j], B[4(l/16)+j][l%16], D[4(l/16)+i][l%16]
// (16x16x16 BF16) and A[l%16][l/16 + 4s], B[l/16 + 4s][l%16] over four K-steps
// (16x16x4 FP32).
#include "mfma_util.h"

namespace {
__global__ void __launch_bounds__(64) k_mfma_selftest(int kind, const float* __restrict__ A,
                                                      const float* __restrict__ Bm, float* __restrict__ D) {
  const int l = threadIdx.x, r = l & 15, q = l >> 4;
  exa::exa_f4 acc = {0.f, 0.f, 0.f, 0.f};
  if (kind == 1) {
    exa::exa_s4 a, b;
    for (int j = 0; j < 4; ++j) {
      a[j] = exa::f32_to_bf16_rne(A[r * 16 + 4 * q + j]);
      b[j] = exa::f32_to_bf16_rne(Bm[(4 * q + j) * 16 + r]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  } else {
    for (int s = 0; s < 4; ++s)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[r * 16 + q + 4 * s], Bm[(q + 4 * s) * 16 + r], acc, 0, 0, 0);
  }
  for (int i = 0; i < 4; ++i) D[(4 * q + i) * 16 + r] = acc[i];
}
}  // namespace

extern "C" int exa_selftest_mfma(int kind, const float* A, const float* Bm, float* D) {
  EXA_ARG(A && Bm && D && (kind == 0 || kind == 1), "arguments");
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  float* d = nullptr;
  EXA_CHECK(hipMalloc(&d, 3 * 256 * sizeof(float)));
  hipError_t e = hipMemcpy(d, A, 256 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + 256, Bm, 256 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, kind, d, d + 256, d + 512);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(D, d + 512, 256 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) {
    exa_set_error("mfma self-test: %s", hipGetErrorString(e));
    return EXA_ERR_HIP;
  }
  return EXA_OK;
}
