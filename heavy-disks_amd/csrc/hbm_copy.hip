// hbm_copy: a plain device-to-device copy kernel of known size, used to
// calibrate the FETCH_SIZE / WRITE_SIZE counters (bytes the counters report
// per byte actually moved) and to measure attainable HBM bandwidth.
//   hbm_copy [MiB per buffer = 1024] [repetitions = 10] [bytes per lane access = 16 | 8]
//            [variant = 0] [workgroups = 8192]
// variant 0: grid-stride copy, one access per lane per iteration;
//         1: the same with four independent accesses per lane in flight;
//         2: variant 0 with nontemporal loads and stores;
//         3: variant 1 with nontemporal loads and stores;
//         4: read only (each lane sums its elements; one store per lane);
//         5: write only (fill).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef double d2v __attribute__((ext_vector_type(2)));

template <typename T, bool NT>
__device__ __forceinline__ T ld(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <typename T, bool NT>
__device__ __forceinline__ void st(T v, T *p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename T, int U, bool NT>
__global__ void copy_kernel(const T *__restrict__ a, T *__restrict__ b, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = ld<T, NT>(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) st<T, NT>(v[u], b + i + u * stride);
  }
  for (; i < n; i += stride) st<T, NT>(ld<T, NT>(a + i), b + i);
}

template <typename T>
__global__ void read_kernel(const T *__restrict__ a, T *__restrict__ out, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  T acc0 = a[0] * 0.0, acc1 = acc0, acc2 = acc0, acc3 = acc0;
  long long i = t;
  for (; i + 3 * stride < n; i += 4 * stride) {
    acc0 += a[i];
    acc1 += a[i + stride];
    acc2 += a[i + 2 * stride];
    acc3 += a[i + 3 * stride];
  }
  for (; i < n; i += stride) acc0 += a[i];
  out[t] = acc0 + acc1 + acc2 + acc3;
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ b, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = v;
}

int main(int argc, char **argv) {
  const long long mib = (argc > 1) ? atoll(argv[1]) : 1024;
  const int reps = (argc > 2) ? atoi(argv[2]) : 10;
  const int width = (argc > 3) ? atoi(argv[3]) : 16;
  const int variant = (argc > 4) ? atoi(argv[4]) : 0;
  const int blocks = (argc > 5) ? atoi(argv[5]) : 256 * 8 * 4;
  if (width != 16 && width != 8) {
    fprintf(stderr, "bytes per lane access must be 16 or 8\n");
    return 1;
  }
  if (variant < 0 || variant > 5 || blocks < 1 || blocks > (1 << 20) || mib < 1 || reps < 1) {
    fprintf(stderr, "variant 0..5, workgroups 1..2^20, MiB and repetitions >= 1\n");
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
  // the read-only variant stores one value per lane of the grid: blocks x 256
  // elements, far below the buffer's n
  if ((long long)blocks * 256 > n) {
    fprintf(stderr, "buffer too small for the grid\n");
    return 1;
  }
  auto launch = [&]() {
    if (width == 16) {
      const d2v *x = reinterpret_cast<const d2v *>(a);
      d2v *y = reinterpret_cast<d2v *>(b);
      switch (variant) {
        case 0: hipLaunchKernelGGL((copy_kernel<d2v, 1, false>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 1: hipLaunchKernelGGL((copy_kernel<d2v, 4, false>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 2: hipLaunchKernelGGL((copy_kernel<d2v, 1, true>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 3: hipLaunchKernelGGL((copy_kernel<d2v, 4, true>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 4: hipLaunchKernelGGL(read_kernel<d2v>, dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        default: hipLaunchKernelGGL(fill_kernel<d2v>, dim3(blocks), dim3(256), 0, 0, y, n, d2v{1.0, 2.0});
      }
    } else {
      const double *x = reinterpret_cast<const double *>(a);
      double *y = reinterpret_cast<double *>(b);
      switch (variant) {
        case 0: hipLaunchKernelGGL((copy_kernel<double, 1, false>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 1: hipLaunchKernelGGL((copy_kernel<double, 4, false>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 2: hipLaunchKernelGGL((copy_kernel<double, 1, true>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 3: hipLaunchKernelGGL((copy_kernel<double, 4, true>), dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        case 4: hipLaunchKernelGGL(read_kernel<double>, dim3(blocks), dim3(256), 0, 0, x, y, n); break;
        default: hipLaunchKernelGGL(fill_kernel<double>, dim3(blocks), dim3(256), 0, 0, y, n, 1.0);
      }
    }
  };
  launch();
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  if (hipEventRecord(e0, 0) != hipSuccess) return 1;
  for (int r = 0; r < reps; ++r) launch();
  if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
  const double per = ms / reps;
  const long long rd = variant == 5 ? 0 : bytes, wr = variant == 4 ? 0 : bytes;
  printf("variant %d, %d workgroups, %lld MiB (%d-byte accesses): %.4f ms per launch, %.1f GB/s "
         "(read + write), %lld bytes read and %lld written per launch\n",
         variant, blocks, mib, width, per, (double)(rd + wr) / (per * 1e-3) / 1e9, rd, wr);
  hipFree(a);
  hipFree(b);
  return 0;
}
