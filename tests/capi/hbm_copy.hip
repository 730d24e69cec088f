// A plain device copy of known size, for the HBM measurements (measurement
// infrastructure, no part of the library): the bandwidth a streaming copy
// reaches on the box, and the bytes that rocprofv3's FETCH_SIZE / WRITE_SIZE
// counters report for it, which calibrate those counters for the element
// kernel's traffic.
//
// usage: hbm_copy [MiB per buffer, default 1024] [repetitions, default 10]
// prints one JSON line: the best launch shape's best and mean GB/s (read +
// write bytes / kernel time), each shape timed over the repetitions.
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

// four 16-byte loads in flight per lane before the stores (n a multiple of 4 * stride, checked by the host)
__global__ __launch_bounds__(256) void k_copy4(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n)
{
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i + 3 * stride < n; i += 4 * stride)
  {
    const double2 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
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
  // launch shapes: the plain grid-stride copy at 32 blocks of 256 per CU, and
  // the four-loads-in-flight copy at 8 and 16 blocks per CU; the best is reported
  struct Shape { int unroll, per_cu; };
  const Shape shapes[3] = {{1, 32}, {4, 8}, {4, 16}};
  double best = 0.0, best_mean = 0.0;
  int best_shape = 0;
  for (int sh = 0; sh < 3; sh++)
  {
    const unsigned grid = (unsigned)(ncu * shapes[sh].per_cu);
    const long long stride = (long long)grid * 256;
    if (shapes[sh].unroll == 4 && n % (4 * stride) != 0) continue;  // k_copy4 covers whole rounds only
    auto launch = [&]() {
      if (shapes[sh].unroll == 4) hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(256), 0, 0, a, b, n);
      else hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    };
    launch();  // warm-up
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    double sb = 0.0, sum = 0.0;
    for (int r = 0; r < reps; r++)
    {
      CK(hipEventRecord(e0, 0));
      launch();
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      const double gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
      sum += gbs;
      if (gbs > sb) sb = gbs;
    }
    if (sb > best)
    {
      best = sb;
      best_mean = sum / reps;
      best_shape = sh;
    }
  }
  printf("{\"kernel\": \"%s\", \"blocks_per_cu\": %d, \"bytes_read\": %zu, \"bytes_written\": %zu, "
         "\"launches_per_shape\": %d, \"best_GB_s\": %.1f, \"mean_GB_s\": %.1f}\n",
         shapes[best_shape].unroll == 4 ? "k_copy4" : "k_copy", shapes[best_shape].per_cu, bytes, bytes, reps + 1, best,
         best_mean);
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
