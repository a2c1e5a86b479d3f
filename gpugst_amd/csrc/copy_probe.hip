// copy_probe.hip -- a plain device-to-device copy of a known number of bytes.
// Used only to calibrate the HBM byte counters (FETCH_SIZE / WRITE_SIZE) that
// rocprofv3 reports for the engine kernel: the ratio counter / true bytes on
// this copy is the scale factor applied to the engine's readings.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char** argv) {
  const size_t bytes = (argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)1 << 30) & ~(size_t)15;
  const size_t n = bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "alloc failed\n");
    return 1;
  }
  (void)hipMemset(a, 1, bytes);
  (void)hipMemset(b, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int rep = 0; rep < 3; rep++) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(copy_u4, dim3(4096), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy %zu bytes: %.3f ms, %.1f GB/s read+write\n", bytes, ms, 2.0 * bytes / (ms * 1e6));
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
