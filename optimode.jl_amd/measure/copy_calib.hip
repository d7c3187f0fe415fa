// Calibration kernel for rocprofv3 byte counters: reads and writes exactly
// 256 MiB (float4 copy), so FETCH_SIZE / WRITE_SIZE per dispatch can be scaled to bytes.
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ void copy_f4(const float4* __restrict__ a, float4* __restrict__ b, long n) {
  for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) b[i] = a[i];
}
int main() {
  const long bytes = 256l << 20, n = bytes / 16;
  float4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  (void)hipMemset(a, 0, bytes);
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(copy_f4, dim3(4096), dim3(256), 0, 0, a, b, n);
  (void)hipDeviceSynchronize();
  printf("copy_f4: %ld bytes read and %ld written per dispatch\n", bytes, bytes);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
