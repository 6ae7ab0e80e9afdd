// Calibration for HBM traffic counters: a plain device copy of known size.
// rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this program gives the counter
// reading per byte actually moved, which scales the same counters read on the
// skeleton kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_kernel(const double4 *__restrict__ a, double4 *__restrict__ b,
                            size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

#define CHK(x)                                                              \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
      return 1;                                                             \
    }                                                                       \
  } while (0)

int main() {
  const size_t bytes = 256ull << 20;  // 256 MiB read, 256 MiB written
  const size_t n = bytes / sizeof(double4);
  double4 *a, *b;
  CHK(hipMalloc(&a, bytes));
  CHK(hipMalloc(&b, bytes));
  CHK(hipMemset(a, 1, bytes));
  CHK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  for (int it = 0; it < 3; ++it) {
    CHK(hipEventRecord(e0));
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, 0, a, b, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    printf("copy %zu MiB: %.3f ms, %.1f GB/s (read + write)\n", bytes >> 20, ms,
           2.0 * bytes / ms / 1e6);
  }
  CHK(hipFree(a));
  CHK(hipFree(b));
  return 0;
}
