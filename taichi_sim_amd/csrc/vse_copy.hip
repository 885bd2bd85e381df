// vse_copy.hip -- plain device-to-device copy kernel: the achievable-HBM
// reference for the roofline and the known-size kernel that calibrates the
// rocprof FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>

#include "../../include/vse.h"

namespace {

// Grid-stride copy with four independent 16-byte loads in flight per thread.
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    const float4 a0 = src[i], a1 = src[i + stride], a2 = src[i + 2 * stride], a3 = src[i + 3 * stride];
    dst[i] = a0;
    dst[i + stride] = a1;
    dst[i + 2 * stride] = a2;
    dst[i + 3 * stride] = a3;
  }
  for (; i < n4; i += stride) dst[i] = src[i];
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
  const int blocks = 256 * 32;   // 32 workgroups per CU on 256 CUs, grid-stride loop
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
