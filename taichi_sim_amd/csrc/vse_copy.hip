// vse_copy.hip -- plain device-to-device copy kernel: the achievable-HBM
// reference for the roofline and the known-size kernel that calibrates the
// rocprof FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>

#include "../../include/vse.h"

namespace {

// One 16-byte element per thread: the pattern that reaches the highest
// sustained HBM throughput on this part (a grid-stride loop does not).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}

}  // namespace

extern "C" int vse_copy_bandwidth(size_t bytes, int iters, float *gbs) {
  if (bytes < 4096 || iters < 1 || !gbs) return VSE_ERR_ARG;
  const size_t n4 = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = VSE_OK;
  if (hipMalloc(&a, n4 * sizeof(float4)) != hipSuccess || hipMalloc(&b, n4 * sizeof(float4)) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = VSE_ERR_HIP;
  }
  if (rc == VSE_OK && hipMemsetAsync(a, 0, n4 * sizeof(float4), st) != hipSuccess) rc = VSE_ERR_HIP;
  const unsigned blocks = (unsigned)((n4 + 255) / 256);
  if (rc == VSE_OK) {
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a, b, n4);   // warm-up
    if (hipEventRecord(e0, st) != hipSuccess) rc = VSE_ERR_HIP;
    for (int i = 0; i < iters && rc == VSE_OK; i++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a, b, n4);
    if (rc == VSE_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                         hipEventSynchronize(e1) != hipSuccess))
      rc = VSE_ERR_HIP;
  }
  if (rc == VSE_OK) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = VSE_ERR_HIP;
    else *gbs = (float)(2.0 * n4 * sizeof(float4) * iters / (ms * 1e-3) / 1e9);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}
