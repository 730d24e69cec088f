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
                                             QuadLds<256> &)
  {
    __shared__ float lx[20 * 256];
    for (long long e = e0 + threadIdx.x; e < e1; e += 256) elem_quad4ax_f32_body<256>(a32, e, lx + threadIdx.x);
    __syncthreads();
  }
};

void launch_batch_f32(const BatchArgs &B, int nbars, hipStream_t s)
{
  hipLaunchKernelGGL((k_batch_quad4<ElemPassF32, true>), dim3(nbars), dim3(256), 0, s, B, B.order, B.bars, B.bar_ea,
                     B.bar_ea32);
}

// stepped engine, mixed form: one element per lane, 64 elements per block.
// The element's node rows sit in LDS and the pass is held to 4 waves per
// SIMD (128 VGPRs, 12 B spilled per lane): 1,050 -> 1,065 bars/s on the
// 512-bar batch against 3 waves with the rows in registers, same box (3
// waves with the rows in LDS: 1,036)
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_bstep_elem32(BatchArgs B, const int2 *__restrict__ blk,
                                                    const BarDesc *__restrict__ bars,
                                                    const ElemArgs32 *__restrict__ bea32, int preamble)
{
  __shared__ float lx[20 * NT];  // each lane's element node rows
  const int2 bb = blk[blockIdx.x];
  StepState *st = B.states + bb.x;
  if (!st->running || st->status) return;
  const BarDesc &d = bars[bb.x];
  const ElemArgs32 &a = bea32[bb.x];
  const long long e_end = d.elem0 + d.ne, e = (long long)bb.y + threadIdx.x;
  if (e < e_end) elem_quad4ax_f32_body<NT>(a, e, lx + threadIdx.x);
  if (preamble || (st->increment % st->freq) == 0)
  {
    // Courant length in FP64 on the FP64 coordinates, as in the FP64 form
    const double l = e < e_end ? quad_length(B.ea.x, B.ea.conn + e * 4) / st->cs : __builtin_inf();
    courant_contrib(l, B.crit + bb.x);
  }
}

// tile 0: 64 elements per block; tile 1: 256 (the block tables match)
void launch_bstep_elem_f32(const BatchArgs &B, const int2 *blk, int nblk, int preamble, int tile, hipStream_t s)
{
  if (tile) hipLaunchKernelGGL(k_bstep_elem32<256>, dim3(nblk), dim3(256), 0, s, B, blk, B.bars, B.bar_ea32, preamble);
  else hipLaunchKernelGGL(k_bstep_elem32<64>, dim3(nblk), dim3(64), 0, s, B, blk, B.bars, B.bar_ea32, preamble);
}

}  // namespace dnl
