// A plain device copy of known size, for the HBM measurements (measurement
// infrastructure, no part of the library): the bandwidth a streaming copy
// reaches on the box, and the bytes that rocprofv3's FETCH_SIZE / WRITE_SIZE
// counters report for it, which calibrate those counters for the element
// kernel's traffic.
//
// usage: hbm_copy [MiB per buffer, default 1024] [repetitions, default 10]
// prints one JSON line: best and mean GB/s (read + write bytes / kernel time).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(call)                                                                            \
  do                                                                                        \
  {                                                                                         \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
    {                                                                                       \
      fprintf(stderr, "hbm_copy: %s failed: %s\n", #call, hipGetErrorString(e_));           \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

// 16 bytes per lane per iteration, grid-stride
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char **argv)
{
  const long long mib = argc > 1 ? atoll(argv[1]) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 10;
  if (mib <= 0 || mib > 65536 || reps <= 0 || reps > 1000)
  {
    fprintf(stderr, "hbm_copy: bad arguments\n");
    return 2;
  }
  const size_t bytes = (size_t)mib << 20;
  const long long n = (long long)(bytes / sizeof(double2));
  double2 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 1, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  int dev = 0, ncu = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)(ncu * 32);  // 8 waves of 4 per CU, several blocks per CU
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);  // warm-up
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  double best = 0.0, sum = 0.0;
  for (int r = 0; r < reps; r++)
  {
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    sum += gbs;
    if (gbs > best) best = gbs;
  }
  printf("{\"kernel\": \"k_copy\", \"bytes_read\": %zu, \"bytes_written\": %zu, \"launches\": %d, \"best_GB_s\": %.1f, "
         "\"mean_GB_s\": %.1f}\n",
         bytes, bytes, reps + 1, best, sum / reps);
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
