// Plain device copy of a known number of bytes, run under rocprofv3 --pmc to
// calibrate FETCH_SIZE / WRITE_SIZE against a kernel whose HBM traffic is
// exactly known (read n bytes, write n bytes).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copyKernel(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

int main(int argc, char** argv) {
  const size_t bytes = (argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1ull << 30);
  const size_t n = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(double4)) != hipSuccess || hipMalloc(&b, n * sizeof(double4)) != hipSuccess) return 1;
  if (hipMemset(a, 1, n * sizeof(double4)) != hipSuccess) return 1;
  for (int rep = 0; rep < 3; rep++) copyKernel<<<4096, 256>>>(a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  std::printf("copied %zu bytes per launch, 3 launches\n", n * sizeof(double4));
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
