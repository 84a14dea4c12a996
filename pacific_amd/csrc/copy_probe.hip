// copy_probe.hip -- a plain device-to-device copy of known size.
//
// Calibrates the HBM counters (FETCH_SIZE, WRITE_SIZE) used to report the
// force kernel's traffic: a copy of B bytes must fetch B and write B bytes, so
// the ratio of counted to true bytes scales the kernel's counters.  Also
// prints the achieved copy bandwidth (read + write) as a practical ceiling.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_probe(const double4* __restrict__ src,
                                                  double4* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 256;
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int reps = 5;
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(copy_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  printf("copy_probe bytes=%zu best_ms=%.4f read+write_GBs=%.1f\n", bytes, best,
         2.0 * bytes / (best * 1e-3) / 1e9);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
