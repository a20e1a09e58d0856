// hbm_copy: a plain device-to-device copy kernel of known size, used to
// calibrate the FETCH_SIZE / WRITE_SIZE counters (bytes the counters report
// per byte actually moved) and to measure attainable HBM bandwidth.
//   hbm_copy [MiB per buffer = 1024] [repetitions = 10] [bytes per lane access = 16 | 8]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

template <typename T>
__global__ void copy_kernel(const T *__restrict__ a, T *__restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int main(int argc, char **argv) {
  const long long mib = (argc > 1) ? atoll(argv[1]) : 1024;
  const int reps = (argc > 2) ? atoi(argv[2]) : 10;
  const int width = (argc > 3) ? atoi(argv[3]) : 16;
  if (width != 16 && width != 8) {
    fprintf(stderr, "bytes per lane access must be 16 or 8\n");
    return 1;
  }
  const long long bytes = mib << 20, n = bytes / width;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  const int blocks = 256 * 8 * 4;
  auto launch = [&]() {
    if (width == 16)
      hipLaunchKernelGGL(copy_kernel<double2>, dim3(blocks), dim3(256), 0, 0, a, b, n);
    else
      hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks), dim3(256), 0, 0,
                         reinterpret_cast<const double *>(a), reinterpret_cast<double *>(b), n);
  };
  launch();
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  if (hipEventRecord(e0, 0) != hipSuccess) return 1;
  for (int r = 0; r < reps; ++r) launch();
  if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
  const double per = ms / reps;
  printf("copy %lld MiB (%d-byte accesses): %.4f ms per launch, %.1f GB/s (read + write), %lld bytes read and %lld written per launch\n",
         mib, width, per, 2.0 * bytes / (per * 1e-3) / 1e9, bytes, bytes);
  hipFree(a);
  hipFree(b);
  return 0;
}
