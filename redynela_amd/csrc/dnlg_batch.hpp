// Persistent batched solver: one workgroup carries one axisymmetric bar
// through its whole ExplicitSolverCH::solve (FE/ExplicitSolverCH.C:115-278),
// its own time step and increment count included; there is no grid-wide
// synchronisation, so bars of uneven size finish independently.  The
// arithmetic is the single-domain FP64 path's (shared device functions of
// dnlg_elem_f64.hpp), so every bar is bit-identical to a run of that bar
// alone.  The element pass is a policy: ElemPassF64 here, the FP32 pass in
// dnlg_batch_f32.hip.
#pragma once
#include "dnlg_elem_f64.hpp"

namespace dnl {

// Domain::computeCourantTimeStep over one bar (FE/Domain.C:1613-1646)
__device__ __forceinline__ double bar_courant(const ElemArgs &a, long long e0, long long e1, double *red)
{
  double t = __builtin_inf();
  for (long long e = e0 + threadIdx.x; e < e1; e += 256) t = fmin(t, quad_length(gmem(a.x), gmem(a.conn) + e * 4) / (a.st->cs));
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  const double c = red[0];
  __syncthreads();
  return c;
}

// the time loop's nodal phase for node i: predictor, element-order force
// gather, corrector, BCs and the grid update (lagMotion::updateNode)
template <bool FE32>
__device__ __forceinline__ void bar_node(const NodalArgs &n, long long i, const StepState *st)
{
#pragma clang fp contract(off)
  double ac[3], vc[3], an[3], vn[3], dun[3], m;
#pragma unroll
  for (int c = 0; c < 3; c++) { ac[c] = n.a[3 * i + c]; vc[c] = n.v[3 * i + c]; }
  nodal_core<2, false, FE32>(n, i, st, ac, vc, an, vn, dun, m);
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
    n.a[3 * i + c] = an[c];
    n.v[3 * i + c] = vn[c];
    n.du[3 * i + c] = dun[c];
    n.d[3 * i + c] = n.d[3 * i + c] + dun[c];
    n.x[3 * i + c] = n.x[3 * i + c] + dun[c];
  }
}

template <bool FE32>
__device__ __forceinline__ void bar_nodal(const NodalArgs &n, long long i0, long long i1, const StepState *st)
{
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) bar_node<FE32>(n, i, st);
}

// element a's contribution to its bar's Courant minimum (one lane per element,
// lanes of other points pass +inf): a wave minimum, then one atomic per wave.
// The minimum of positive doubles is exact in any order, and positive doubles
// order like their bit patterns.
__device__ __forceinline__ void courant_contrib(double l, unsigned long long *crit)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l = fmin(l, __shfl_xor(l, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMin(crit, (unsigned long long)__double_as_longlong(l));
}

// FP64 element pass over one bar: 64 elements (4 points each) per sweep of 256 lanes
struct ElemPassF64
{
  static __device__ __forceinline__ void run(const ElemArgs &a, const ElemArgs32 &, long long e0, long long e1,
                                             QuadLds<256> &Ls)
  {
    for (long long base = e0; base < e1; base += 64)
    {
      elem_quad4ax_body<false, false, false, 256>(a, base + (threadIdx.x >> 2), e1, threadIdx.x, Ls);
      __syncthreads();
    }
  }
};

// The bar order, descriptors and per-bar arguments come in as restrict-qualified
// read-only pointers of their own so that the compiler reads them with scalar
// loads (see k_bstep_elem).
template <class PASS, bool FE32>
__global__ __launch_bounds__(256) void k_batch_quad4(BatchArgs B, const int *__restrict__ order,
                                                     const BarDesc *__restrict__ bars,
                                                     const ElemArgs *__restrict__ bea,
                                                     const ElemArgs32 *__restrict__ bea32)
{
  __shared__ QuadLds<256> Ls;
  __shared__ double red[256];
  __shared__ int s_go, s_rec;
  const int bar = order[blockIdx.x];
  const BarDesc d = bars[bar];
  StepState *st = B.states + bar;
  const ElemArgs &a = bea[bar];      // in global memory: no per-lane copy
  const ElemArgs32 &a32 = bea32[bar];
  const long long e0 = d.elem0, e1 = d.elem0 + d.ne, n0 = d.node0, n1 = d.node0 + d.nn;
  quad_fill_table(Ls);
  __syncthreads();
  if (!st->running)
  {
    // betweenTimeBounds false on entry: solve() does nothing (FE/ExplicitSolverCH.C:121-128)
    if (threadIdx.x == 0) B.t0[bar] = B.t1[bar] = wall_clock64();
    return;
  }
  if (threadIdx.x == 0) B.t0[bar] = wall_clock64();
  // solve() preamble: computeInternalForces, computeTimeStep(True)
  PASS::run(a, a32, e0, e1, Ls);
  // a bar stops at its first element error (status is set by atomics at L2):
  // the failing increment counts and nothing after it runs, as the reference
  // stops inside computeInternalForces
  if (threadIdx.x == 0) s_go = __hip_atomic_load(&st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  __syncthreads();
  if (!s_go)
  {
    if (threadIdx.x == 0) B.t1[bar] = wall_clock64();
    return;
  }
  double crit = bar_courant(a, e0, e1, red);
  if (threadIdx.x == 0) compute_time_step(st, 1, crit);
  __syncthreads();
  // the time loop; every branch below depends only on shared or uniform values
  for (long long it = 0;; it++)
  {
    if (it > 0)
    {
      // endStep of the previous increment and the loop condition; the Courant
      // decision uses the increment count as it stands here, before initStep
      if (threadIdx.x == 0)
      {
        s_go = end_step(st);
        s_rec = (st->increment % st->freq) == 0;
      }
      __syncthreads();
      if (!s_go) break;
      if (it >= B.max_steps)
      {
        if (threadIdx.x == 0) atomicCAS(&st->status, 0, (int)ST_BUDGET);
        break;
      }
      if (s_rec) crit = bar_courant(a, e0, e1, red);
      if (threadIdx.x == 0) compute_time_step(st, s_rec, crit);
    }
    if (threadIdx.x == 0) init_step(st);
    __syncthreads();
    PASS::run(a, a32, e0, e1, Ls);
    if (threadIdx.x == 0) s_go = __hip_atomic_load(&st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    __syncthreads();
    if (!s_go) break;
    bar_nodal<FE32>(B.na, n0, n1, st);
    __syncthreads();
  }
  if (threadIdx.x == 0) B.t1[bar] = wall_clock64();
}

}  // namespace dnl
