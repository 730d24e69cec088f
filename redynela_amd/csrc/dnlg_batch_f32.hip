// Persistent batched solver, mixed form: FP32 element pass (one element per
// lane), FP64 nodal state.  Compiled with -ffp-contract=fast-honor-pragmas;
// the FP64 helpers shared with dnlg_elem_f64.hpp switch contraction off in
// their bodies.
#include "dnlg_batch.hpp"
#include "dnlg_elem_f32.hpp"

namespace dnl {

struct ElemPassF32
{
  static __device__ __forceinline__ void run(const ElemArgs &, const ElemArgs32 &a32, long long e0, long long e1,
                                             double (*)[256])
  {
    for (long long e = e0 + threadIdx.x; e < e1; e += 256) elem_quad4ax_f32_body(a32, e);
    __syncthreads();
  }
};

void launch_batch_f32(const BatchArgs &B, int nbars, hipStream_t s)
{
  hipLaunchKernelGGL((k_batch_quad4<ElemPassF32, true>), dim3(nbars), dim3(256), 0, s, B);
}

}  // namespace dnl
