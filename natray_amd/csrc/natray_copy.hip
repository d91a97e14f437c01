// natray_copy: a plain device-to-device copy of a known size, the reference point for HBM
// traffic counters.
//
//   natray_copy [MiB=256] [launches=5]
//
// Each launch reads and writes exactly MiB mebibytes with 16-byte loads and stores, one pass
// over the data and no reuse, so a profiler's FETCH_SIZE and WRITE_SIZE for one launch can be
// compared with MiB x 1024 KiB to get the counters' scale on this GPU. It prints one line
// with the bytes per launch and the achieved GB/s.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                   \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 256;
  const int launches = argc > 2 ? atoi(argv[2]) : 5;
  if (mib == 0 || mib > 65536 || launches <= 0 || launches > 1000) {
    fprintf(stderr, "usage: %s [MiB 1..65536] [launches 1..1000]\n", argv[0]);
    return 2;
  }
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  CHK(hipMalloc(&a, bytes));
  CHK(hipMalloc(&b, bytes));
  CHK(hipMemset(a, 0, bytes));
  CHK(hipMemset(b, 0, bytes));
  int cus = 0;
  CHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const unsigned grid = (unsigned)cus * 16u;
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  CHK(hipGetLastError());
  CHK(hipEventRecord(e0, 0));
  for (int k = 0; k < launches; ++k) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  CHK(hipGetLastError());
  CHK(hipEventRecord(e1, 0));
  CHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  CHK(hipEventElapsedTime(&ms, e0, e1));
  printf("natray_copy: %zu bytes read and %zu written per launch, %d timed launches, %.1f GB/s (read + write)\n", bytes,
         bytes, launches, 2.0 * (double)bytes * launches / (ms * 1e-3) / 1e9);
  fflush(stdout);
  CHK(hipFree(a));
  CHK(hipFree(b));
  return 0;
}
