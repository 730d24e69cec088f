// Mixed-precision element pass: grid kernels over the device functions of
// dnlg_elem_f32.hpp (compiled with FMA contraction).
#include "dnlg_elem_f32.hpp"

namespace dnl {

__global__ __launch_bounds__(128) void k_elem_hex8_f32(ElemArgs32 a)
{
  __shared__ float lx[48 * 128];  // each lane's element node rows, 24 KB: 3 waves per SIMD
  if ((a.check_running && !a.st->running) || a.st->halted) return;
  const long long e = (long long)blockIdx.x * 128 + threadIdx.x;
  if (e >= a.ne) return;
  elem_hex8_f32_body<128>(a, e, lx + threadIdx.x);
}

__global__ __launch_bounds__(128) void k_elem_quad4ax_f32(ElemArgs32 a)
{
  if ((a.check_running && !a.st->running) || a.st->halted) return;
  const long long e = (long long)blockIdx.x * 128 + threadIdx.x;
  if (e >= a.ne) return;
  elem_quad4ax_f32_body(a, e);
}

void launch_elem_f32(int dim, const ElemArgs32 &a, hipStream_t s)
{
  if (a.ne <= 0) return;
  const unsigned grid = (unsigned)((a.ne + 127) / 128);
  if (dim == 3) hipLaunchKernelGGL(k_elem_hex8_f32, dim3(grid), dim3(128), 0, s, a);
  else hipLaunchKernelGGL(k_elem_quad4ax_f32, dim3(grid), dim3(128), 0, s, a);
}

}  // namespace dnl
