// Persistent batched solver, FP64 form (compiled with -ffp-contract=off).
#include "dnlg_batch.hpp"

namespace dnl {

void launch_batch_f64(const BatchArgs &B, int nbars, hipStream_t s)
{
  hipLaunchKernelGGL((k_batch_quad4<ElemPassF64, false>), dim3(nbars), dim3(256), 0, s, B);
}

}  // namespace dnl
