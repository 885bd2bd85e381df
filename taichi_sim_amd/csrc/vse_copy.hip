// vse_copy.hip -- plain device-to-device copy kernel: the achievable-HBM
// reference for the roofline and the known-size kernel that calibrates the
// rocprof FETCH_SIZE / WRITE_SIZE counters; and a stream-mix kernel that
// measures what the device sustains at a stencil pass's read/write mix.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/vse.h"

namespace {

// One 16-byte element per thread: the pattern that reaches the highest
// sustained HBM throughput on this part (a grid-stride loop does not).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}

// The read/write mix of a stencil pass without its arithmetic: every thread
// loads element i of all T arrays first (the first n_read of them read-only),
// then writes the other T - n_read back in place.  T is a compile-time count
// so that every load is in flight together and the values stay in registers.
template <int T>
__global__ __launch_bounds__(256) void k_streams(float *const *__restrict__ a, int n_read, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v[T];
#pragma unroll
  for (int k = 0; k < T; k++) v[k] = a[k][i];
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < T; k++)
    if (k < n_read) s += v[k];
#pragma unroll
  for (int k = 0; k < T; k++)
    if (k >= n_read) a[k][i] = v[k] + s;
}

template <int T>
void launch_streams(float *const *a, int n_read, int total, size_t n, hipStream_t st) {
  if constexpr (T <= VSE_STREAMS_MAX) {
    if (total == T) hipLaunchKernelGGL(k_streams<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n_read, n);
    else launch_streams<T + 1>(a, n_read, total, n, st);
  }
}

}  // namespace

extern "C" int vse_stream_bandwidth(size_t n, int n_read, int n_rw, int iters, float *gbs) {
  if (n < 1024 || n_read < 0 || n_rw < 1 || n_read + n_rw > VSE_STREAMS_MAX || iters < 1 || !gbs) return VSE_ERR_ARG;
  const int na = n_read + n_rw;
  std::vector<float *> arr(na, nullptr);
  float **d_ptrs = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = VSE_OK;
  for (int k = 0; k < na && rc == VSE_OK; k++) {
    if (hipMalloc(&arr[k], n * sizeof(float)) != hipSuccess) rc = VSE_ERR_HIP;
  }
  if (rc == VSE_OK && (hipMalloc(&d_ptrs, na * sizeof(float *)) != hipSuccess ||
                       hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                       hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = VSE_ERR_HIP;
  for (int k = 0; k < na && rc == VSE_OK; k++)
    if (hipMemsetAsync(arr[k], 0, n * sizeof(float), st) != hipSuccess) rc = VSE_ERR_HIP;
  if (rc == VSE_OK && (hipMemcpyAsync(d_ptrs, arr.data(), na * sizeof(float *), hipMemcpyHostToDevice, st) != hipSuccess ||
                       hipStreamSynchronize(st) != hipSuccess))
    rc = VSE_ERR_HIP;
  if (rc == VSE_OK) {
    launch_streams<1>(d_ptrs, n_read, na, n, st);   // warm-up
    if (hipEventRecord(e0, st) != hipSuccess) rc = VSE_ERR_HIP;
    for (int i = 0; i < iters && rc == VSE_OK; i++) launch_streams<1>(d_ptrs, n_read, na, n, st);
    if (rc == VSE_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                         hipEventSynchronize(e1) != hipSuccess))
      rc = VSE_ERR_HIP;
  }
  if (rc == VSE_OK) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = VSE_ERR_HIP;
    else *gbs = (float)((double)n * sizeof(float) * (n_read + 2.0 * n_rw) * iters / (ms * 1e-3) / 1e9);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (d_ptrs) (void)hipFree(d_ptrs);
  for (float *q : arr)
    if (q) (void)hipFree(q);
  return rc;
}

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
