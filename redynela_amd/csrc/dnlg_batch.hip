// Batched solver, FP64 form (compiled with -ffp-contract=off): the persistent
// kernel's FP64 instance, and the stepped engine's element pass, per-bar time
// step and nodal phase (the last two serve both forms).
#include "dnlg_batch.hpp"

namespace dnl {

// ---------------------------------------------------------------------------
// Stepped engine: every bar advances by one increment per group of three
// launches (element pass over all running bars, per-bar time step, nodal
// phase over all active bars), each bar with its own time step.  The order of
// operations inside a bar is that of ExplicitSolverCH::solve, so a bar stays
// bit-identical to its single-domain run.  Block tables map a block to
// (bar, first element / node); blocks never straddle two bars.
// ---------------------------------------------------------------------------

// per-bar scalar phase.  preamble: computeTimeStep(True) after the preamble's
// element pass.  Otherwise, for each running bar: computeTimeStep (Courant
// minimum from this increment's element pass when increment % freq == 0),
// initStep, and -- ahead of the nodal phase, which reads neither t nor the
// running flag -- the endStep time update and loop condition.  A bar stops
// at the first element error, with the failing increment counted.
__global__ __launch_bounds__(64) void k_bstep_dt(BatchArgs B, int nb, int preamble)
{
  const int bar = blockIdx.x * 64 + threadIdx.x;
  if (bar >= nb) return;
  StepState *st = B.states + bar;
  const double crit = __longlong_as_double((long long)B.crit[bar]);
  B.crit[bar] = 0x7FF0000000000000ull;  // +inf
  if (preamble)
  {
    st->active = 0;
    if (st->running && !st->status)
    {
      B.t0[bar] = wall_clock64();
      compute_time_step(st, 1, crit);
    }
    return;
  }
  // a bar whose status was raised before this increment's element pass stays
  // where it stopped
  if (!st->running || (st->status && !st->active))
  {
    st->active = 0;
    return;
  }
  const int rec = (st->increment % st->freq) == 0;
  compute_time_step(st, rec, crit);
  init_step(st);
  if (st->status)
  {
    // an element error in this increment's pass: the increment counts and
    // nothing after it runs, as the reference stops inside
    // computeInternalForces (no nodal update, no endStep)
    st->active = 0;
    B.t1[bar] = wall_clock64();
    return;
  }
  st->active = 1;
  if (!end_step(st)) B.t1[bar] = wall_clock64();
}

// [nn][3] rows of the block's nodes through LDS, so that global loads and
// stores cover whole contiguous spans (the block's nodes are contiguous and
// end at the bar's last node at most)
// (the loads of all four fields go out first, at indices clamped into the
// bar's rows, and the LDS writes follow: see RowLoad)
template <int NT>
struct BRowLoad
{
  double v[3];
  __device__ __forceinline__ void load(const double *g, long long i0, long long iend)
  {
    const long long base = 3 * i0, lim = 3 * iend;
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      const long long j = base + threadIdx.x + k * NT;
      v[k] = g[j < lim ? j : lim - 1];
    }
  }
  __device__ __forceinline__ void store(double *l, long long i0, long long iend) const
  {
    const long long base = 3 * i0, lim = 3 * iend;
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      const int j = threadIdx.x + k * NT;
      l[j] = base + j < lim ? v[k] : 0.0;
    }
  }
};
template <int NT>
__device__ __forceinline__ void bstage_out(double *g, const double *l, long long i0, long long iend)
{
  const long long base = 3 * i0, lim = 3 * iend;
#pragma unroll
  for (int j = threadIdx.x; j < 3 * NT; j += NT)
    if (base + j < lim) g[base + j] = l[j];
}

// nodal phase of one block of a bar's nodes: the arithmetic of bar_node,
// with x, v, a, d staged through LDS (Delta u leaves through the x rows)
template <bool FE32, int NT>
__global__ __launch_bounds__(NT) void k_bstep_nodal(BatchArgs B, const int2 *blk)
{
#pragma clang fp contract(off)
  __shared__ double sx[3 * NT], sv[3 * NT], sa[3 * NT], sd[3 * NT];
  const int2 bb = blk[blockIdx.x];
  const StepState *st = B.states + bb.x;
  if (!st->active) return;
  const BarDesc &d = B.bars[bb.x];
  const NodalArgs &n = B.na;
  const long long i0 = bb.y, iend = d.node0 + d.nn, i = i0 + threadIdx.x;
  const int t3 = 3 * threadIdx.x;
  {
    BRowLoad<NT> rx, rv, ra, rd;
    rx.load(n.x, i0, iend);
    rv.load(n.v, i0, iend);
    ra.load(n.a, i0, iend);
    rd.load(n.d, i0, iend);
    rx.store(sx, i0, iend);
    rv.store(sv, i0, iend);
    ra.store(sa, i0, iend);
    rd.store(sd, i0, iend);
  }
  __syncthreads();
  double dun[3] = {0.0, 0.0, 0.0};
  if (i < iend)
  {
    double ac[3], vc[3], an[3], vn[3], m;
#pragma unroll
    for (int c = 0; c < 3; c++) { ac[c] = sa[t3 + c]; vc[c] = sv[t3 + c]; }
    nodal_core<2, false, FE32>(n, i, st, ac, vc, an, vn, dun, m);
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
      sa[t3 + c] = an[c];
      sv[t3 + c] = vn[c];
      sd[t3 + c] = sd[t3 + c] + dun[c];
      sx[t3 + c] = sx[t3 + c] + dun[c];
    }
  }
  __syncthreads();
  bstage_out<NT>(n.x, sx, i0, iend);
  bstage_out<NT>(n.v, sv, i0, iend);
  bstage_out<NT>(n.a, sa, i0, iend);
  bstage_out<NT>(n.d, sd, i0, iend);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; c++) sx[t3 + c] = dun[c];
  __syncthreads();
  bstage_out<NT>(n.du, sx, i0, iend);
}

// FP64 element pass: NT / 4 elements (NT points) per block; 5 waves per SIMD
// (96 VGPRs, no spills; the LDS of either tile allows as many)
// The block table, the bar descriptors and the per-bar arguments come in as
// restrict-qualified read-only pointers of their own, so that the compiler
// may read them with scalar loads at the start; through BatchArgs every
// argument field is a vector load waited on where it is used.
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_bstep_elem(
    BatchArgs B, const int2 *__restrict__ blk, const BarDesc *__restrict__ bars, const ElemArgs *__restrict__ bea,
    int preamble)
{
  __shared__ QuadLds<NT> Ls;
  const int2 bb = blk[blockIdx.x];
  StepState *st = B.states + bb.x;
  if (!st->running || st->status) return;
  const BarDesc &d = bars[bb.x];
  const ElemArgs &a = bea[bb.x];
  quad_fill_table(Ls);
  __syncthreads();
  const long long e_end = d.elem0 + d.ne, e_raw = (long long)bb.y + (threadIdx.x >> 2);
  elem_quad4ax_body<false, false, false, NT>(a, e_raw, e_end, threadIdx.x, Ls);
  if (preamble || (st->increment % st->freq) == 0)
  {
    double l = __builtin_inf();
    if ((threadIdx.x & 3) == 0 && e_raw < e_end) l = quad_length(gmem(a.x), gmem(a.conn) + e_raw * 4) / st->cs;
    courant_contrib(l, B.crit + bb.x);
  }
}


void launch_batch_f64(const BatchArgs &B, int nbars, hipStream_t s)
{
  hipLaunchKernelGGL((k_batch_quad4<ElemPassF64, false>), dim3(nbars), dim3(256), 0, s, B, B.order, B.bars, B.bar_ea,
                     B.bar_ea32);
}

// tile 0: 64-lane blocks; tile 1: 256-lane blocks (the block tables match)
void launch_bstep_elem_f64(const BatchArgs &B, const int2 *blk, int nblk, int preamble, int tile, hipStream_t s)
{
  if (tile) hipLaunchKernelGGL(k_bstep_elem<256>, dim3(nblk), dim3(256), 0, s, B, blk, B.bars, B.bar_ea, preamble);
  else hipLaunchKernelGGL(k_bstep_elem<64>, dim3(nblk), dim3(64), 0, s, B, blk, B.bars, B.bar_ea, preamble);
}

void launch_bstep_dt(const BatchArgs &B, int nb, int preamble, hipStream_t s)
{
  hipLaunchKernelGGL(k_bstep_dt, dim3((nb + 63) / 64), dim3(64), 0, s, B, nb, preamble);
}

void launch_bstep_nodal(const BatchArgs &B, const int2 *blk, int nblk, bool fe32, int tile, hipStream_t s)
{
  if (tile)
  {
    if (fe32) hipLaunchKernelGGL((k_bstep_nodal<true, 256>), dim3(nblk), dim3(256), 0, s, B, blk);
    else hipLaunchKernelGGL((k_bstep_nodal<false, 256>), dim3(nblk), dim3(256), 0, s, B, blk);
  }
  else
  {
    if (fe32) hipLaunchKernelGGL((k_bstep_nodal<true, 64>), dim3(nblk), dim3(64), 0, s, B, blk);
    else hipLaunchKernelGGL((k_bstep_nodal<false, 64>), dim3(nblk), dim3(64), 0, s, B, blk);
  }
}

}  // namespace dnl
