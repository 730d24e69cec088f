// FP64 kernels of the explicit step (compiled with -ffp-contract=off).
//
//   k_elem<Hex8|Quad4Ax>   Element::computeInternalForces + computeStress
//                          (FE/Element.C:164-403) with the Jacobian pass of
//                          Element3D/Element2D::computeInternalMatrices fused in
//                          (FE/Element3D.C:53-113, FE/Element2D.C:53-126).
//                          One integration point per lane: 8 lanes per hex8,
//                          4 lanes per axisymmetric quad, one wave per block.
//   k_gather_nodal         ExplicitSolverCH::computePredictions + the Fint
//                          assembly of Domain::computeInternalForces +
//                          explicitSolve + updateGrid, fused per node
//                          (FE/ExplicitSolverCH.C:387-567, FE/Domain.C:160-193).
//                          Deterministic node-centric gather over a
//                          node -> (element, local node) CSR in element order.
//   k_courant              Domain::computeCourantTimeStep (FE/Domain.C:1613-1646)
//   k_dt                   Solver::computeTimeStep + ExplicitSolverCH::initStep/endStep
//                          (FE/Solver.C:220-267, FE/ExplicitSolverCH.C:434-468)
//   k_mass_elem / k_gather_scalar   Domain/Element::computeMassMatrix
//                          (FE/Domain.C:948-1005, FE/Element.C:411-454)
#include "dnlg_elem_f64.hpp"

namespace dnl {

// 256 lanes (32 elements) per workgroup: the four waves share one table of
// shape-function derivatives, so the LDS of a workgroup (40 KB) lets four
// workgroups, 4 waves per SIMD, sit on a CU -- the limit the registers set
// too; with one wave per workgroup, LDS held it to 3.25
constexpr int kHexNT = 256;
template <bool SPEED_BC, bool TRACK, bool JC>
__global__ __launch_bounds__(kHexNT) void k_elem_hex8(ElemArgs a)
{
  // Workgroups are dispatched to the 8 XCDs round-robin, so neighbouring
  // elements, which share nodes, would each pull the nodes' lines into a
  // different XCD's L2.  Runs of 8 workgroups (256 elements) go to one XCD
  // instead, the runs rotating over the XCDs: reads fall from 930 to 790 B
  // per element on the 1M bar, and the plastic zone at the impact face still
  // spreads over every XCD (one contiguous range per XCD would put it all on
  // the first).  A bijection on [0, gridDim.x); the tail past the last full
  // round keeps its place.
  constexpr unsigned kRun = 8;
  const unsigned nb = gridDim.x, b = blockIdx.x, full = (nb / (8 * kRun)) * (8 * kRun);
  unsigned lb = b;
  if (b < full)
  {
    const unsigned x = b & 7, k = b >> 3;
    lb = ((k / kRun) * 8 + x) * kRun + k % kRun;
  }
  const long long blk0 = (long long)lb * kHexNT;
  // the connectivity and table loads go out before the step-state check and
  // the barrier, so that the three latencies overlap instead of adding up
  const unsigned lo = hex_lane_offset<kHexNT>(a, blk0, threadIdx.x);
  const int mynode = (a.conn + blk0)[lo];
  constexpr int kTab = 8 * 8 * 3 + 4;
  const double tab = threadIdx.x < kTab ? kHexTabDev.v[threadIdx.x] : 0.0;
  const int running = a.st->running, halted = a.st->halted;  // one wait for both
  if ((a.check_running && !running) || halted) return;
  __shared__ HexLds<kHexNT> Ls;
  if (threadIdx.x < kTab) Ls.tab[threadIdx.x] = tab;
  __syncthreads();
  elem_hex8_body<SPEED_BC, TRACK, JC, kHexNT>(a, blk0, threadIdx.x, lo, mynode, Ls);
}

template <bool SPEED_BC, bool TRACK, bool JC>
__global__ __launch_bounds__(64) void k_elem_quad4ax(ElemArgs a)
{
  if ((a.check_running && !a.st->running) || a.st->halted) return;
  __shared__ QuadLds<64> Ls;
  quad_fill_table(Ls);
  __syncthreads();
  const long long e_raw = ((long long)blockIdx.x * 64 + threadIdx.x) >> 2;
  elem_quad4ax_body<SPEED_BC, TRACK, JC, 64>(a, e_raw, a.ne, threadIdx.x, Ls);
}

// ---------------------------------------------------------------------------
// lumped mass per element, FE/Element.C:411-454 (one element per thread)
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_mass_elem(ElemArgs a, double rho, double hn0, double hn1, double hn2,
                                                    double hn3, double *me)
{
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= a.ne) return;
  constexpr int NEN = DIM == 3 ? 8 : 4;
  const int *cn = a.conn + e * NEN;
  double M[NEN];
#pragma unroll
  for (int i = 0; i < NEN; i++) M[i] = 0.;
  for (int pt = 0; pt < NEN; pt++)
  {
    double detJ, WxdJ;
    if (DIM == 3)
    {
      T9 J = {{0, 0, 0, 0, 0, 0, 0, 0, 0}};
#pragma unroll
      for (int k = 0; k < 8; k++)
      {
        const double *xk = a.x + 3 * (long long)cn[k];
        double g[3];
        hex_dn0(a, pt, k, g);
        J.v[0] += g[0] * xk[0]; J.v[1] += g[1] * xk[0]; J.v[2] += g[2] * xk[0];
        J.v[3] += g[0] * xk[1]; J.v[4] += g[1] * xk[1]; J.v[5] += g[2] * xk[1];
        J.v[6] += g[0] * xk[2]; J.v[7] += g[1] * xk[2]; J.v[8] += g[2] * xk[2];
      }
      detJ = (J.v[0] * J.v[4] * J.v[8]) + (J.v[3] * J.v[7] * J.v[2]) + (J.v[6] * J.v[1] * J.v[5]) -
             (J.v[6] * J.v[4] * J.v[2]) - (J.v[0] * J.v[7] * J.v[5]) - (J.v[3] * J.v[1] * J.v[8]);
      WxdJ = 1.0 * detJ;
#pragma unroll
      for (int i = 0; i < NEN; i++)
      {
        const int ag = (sg0(pt) == sg0(i)) + (sg1(pt) == sg1(i)) + (sg2(pt) == sg2(i));
        const double N = ag == 3 ? hn3 : ag == 2 ? hn2 : ag == 1 ? hn1 : hn0;
        M[i] += rho * N * WxdJ;
      }
    }
    else
    {
      double J0 = 0, J1 = 0, J2 = 0, J3 = 0, R = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++)
      {
        const double *xk = a.x + 3 * (long long)cn[k];
        double g[2];
        quad_dn0(a, pt, k, g);
        J0 += g[0] * xk[0]; J1 += g[1] * xk[0];
        J2 += g[0] * xk[1]; J3 += g[1] * xk[1];
      }
#pragma unroll
      for (int k = 0; k < 4; k++) R += quad_n(a, pt, k) * a.x[3 * (long long)cn[k]];
      detJ = (J0 * J3) - (J1 * J2);
      WxdJ = 1.0 * detJ;
      WxdJ *= 2. * kPI * R;
#pragma unroll
      for (int i = 0; i < NEN; i++) M[i] += rho * quad_n(a, pt, i) * WxdJ;
    }
  }
#pragma unroll
  for (int i = 0; i < NEN; i++) me[a.fpos[e * NEN + i]] = M[i];
}

// per-node sum of element contributions in element order (mass assembly,
// FE/Domain.C:966-990): slot s of node i holds its s-th element in element order
__global__ __launch_bounds__(256) void k_gather_scalar(const unsigned char *deg, const double *src, double *dst,
                                                      long long nn)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  double s = 0.;
  const int d = deg[i];
  for (int k = 0; k < d; k++) s += src[(long long)k * nn + i];
  dst[i] = s;
}

// Domain::Fint only (the preamble's computeInternalForces): gather in element order
template <int DIM, typename FT>
__global__ __launch_bounds__(256) void k_gather_fint(const unsigned char *deg, const FT *fe, double *fint, long long nn)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  double f[DIM];
#pragma unroll
  for (int j = 0; j < DIM; j++) f[j] = 0.0;
  const int d = deg[i];
  for (int k = 0; k < d; k++)
  {
    const FT *s = fe + ((long long)k * nn + i) * DIM;
#pragma unroll
    for (int j = 0; j < DIM; j++) f[j] += (double)s[j];
  }
#pragma unroll
  for (int j = 0; j < DIM; j++) fint[i * DIM + j] = f[j];
}

// ---------------------------------------------------------------------------
// fused nodal step: predictor, BCs, Fint gather, corrector, BCs, grid update
// ---------------------------------------------------------------------------

// [nn][3] AoS arrays are moved through LDS so that every global load and
// store of a block covers whole contiguous lines (3 x 256 doubles per field).
// Each lane loads its three values of every field first, at indices clamped
// into the array so that the loads carry no branch and go out back to back,
// and writes them to LDS afterwards (zeros past the end): a conditional load
// per value would wait on each load before issuing the next.
template <int NT>
struct RowLoad
{
  double v[3];
  __device__ __forceinline__ void load(const double *g, long long i0, long long nn)
  {
    const long long base = 3 * i0, lim = 3 * nn;
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      const long long j = base + threadIdx.x + k * NT;
      v[k] = g[j < lim ? j : lim - 1];
    }
  }
  __device__ __forceinline__ void store(double *l, long long i0, long long nn) const
  {
    const long long base = 3 * i0, lim = 3 * nn;
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      const int j = threadIdx.x + k * NT;
      l[j] = base + j < lim ? v[k] : 0.0;
    }
  }
};
__device__ __forceinline__ void stage_out(double *g, const double *l, long long i0, long long nn)
{
  const long long base = 3 * i0, lim = 3 * nn;
#pragma unroll
  for (int j = threadIdx.x; j < 768; j += 256)
    if (base + j < lim) g[base + j] = l[j];
}

// LDS: four staging rows (x, v, a, d) of 6 KB; the new increment Delta u goes
// out through the x row once x has been written back, and the kinetic-energy
// partials reduce in the v row: 24 KB per block, 6 blocks per CU
template <int DIM, bool TRACK_D, bool HAS_BCS, bool FE32>
__global__ __launch_bounds__(256) void k_gather_nodal(NodalArgs n)
{
  StepState *st = n.st;
  __shared__ double sx[768], sv[768], sa[768], sd[TRACK_D ? 768 : 1];
  if (!st->running || st->status) return;  // no nodal update after an element error
  const long long i0 = (long long)blockIdx.x * 256;
  const long long i = i0 + threadIdx.x;
  const int t3 = 3 * threadIdx.x;
  {
    RowLoad<256> rx, rv, ra, rd;
    rx.load(n.x, i0, n.nn);
    rv.load(n.v, i0, n.nn);
    ra.load(n.a, i0, n.nn);
    if (TRACK_D) rd.load(n.d, i0, n.nn);
    rx.store(sx, i0, n.nn);
    rv.store(sv, i0, n.nn);
    ra.store(sa, i0, n.nn);
    if (TRACK_D) rd.store(sd, i0, n.nn);
  }
  __syncthreads();
  double ke = 0.0, dun[3] = {0.0, 0.0, 0.0};
  if (i < n.nn)
  {
    double ac[3], vc[3], vn[3], an[3], m;
#pragma unroll
    for (int c = 0; c < 3; c++) { ac[c] = sa[t3 + c]; vc[c] = sv[t3 + c]; }
    nodal_core<DIM, HAS_BCS, FE32>(n, i, st, ac, vc, an, vn, dun, m);
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
      sa[t3 + c] = an[c];
      sv[t3 + c] = vn[c];
      if (TRACK_D) sd[t3 + c] = sd[t3 + c] + dun[c];
      sx[t3 + c] = sx[t3 + c] + dun[c];  // lagMotion::updateNode, FE/lagMotion.C:52-57
    }
    if (n.ke_part) ke = .5 * m * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
  }
  __syncthreads();
  stage_out(n.x, sx, i0, n.nn);
  stage_out(n.v, sv, i0, n.nn);
  stage_out(n.a, sa, i0, n.nn);
  if (TRACK_D) stage_out(n.d, sd, i0, n.nn);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; c++) sx[t3 + c] = dun[c];
  if (n.ke_part) sv[threadIdx.x] = ke;
  __syncthreads();
  stage_out(n.du, sx, i0, n.nn);
  if (n.ke_part)
  {
    for (int w = 128; w > 0; w >>= 1)
    {
      if (threadIdx.x < w) sv[threadIdx.x] += sv[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) n.ke_part[blockIdx.x] = sv[0];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->pending = 1;  // endStep of this increment is due
}

// ---------------------------------------------------------------------------
// Courant critical time per element, min-reduced per block
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_courant(ElemArgs a, double cs, double *part)
{
  __shared__ double red[256];
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  double t = __builtin_inf();
  if (e < a.ne)
  {
    if (DIM == 3)
    {
      int nd[8];
#pragma unroll
      for (int k = 0; k < 8; k++) nd[k] = a.conn[e * 8 + k];
      t = hex_length(a, nd) / (cs);
    }
    else
    {
      const double len = quad_length(a.x, a.conn + e * 4);
      t = len / (cs);
    }
  }
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// the minimum of the first npart block minima (one block): the crit that
// Domain::computeCourantTimeStep returns (FE/Domain.C:1613-1646); min is
// exact, so the order of the reduction does not change the result
__global__ __launch_bounds__(256) void k_min_reduce(const double *part, int npart, double *out)
{
  __shared__ double red[256];
  double c = __builtin_inf();
  for (int k = threadIdx.x; k < npart; k += 256) c = fmin(c, part[k]);
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// scalar step control (one block).
//   mode 0 (start of a step): the pending end of the previous increment
//           (ExplicitSolverCH::endStep then, while running, Solver::computeTimeStep),
//           then ExplicitSolverCH::initStep of the new increment;
//   mode 1 (sync point): only the pending end of the last increment;
//   mode 2 (solve() preamble): forced Solver::computeTimeStep.
// For the Pulse method a recomputing step splits in two around the power
// iteration (its frequency is then in st->pulse_f):
//   mode 3: the pending endStep only;
//   mode 4: forced computeTimeStep then initStep, while running;
//   mode 5: forced computeTimeStep only, while running.
// FE/ExplicitSolverCH.C:172, 244-270, 434-468; FE/Solver.C:169-176, 220-267.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dt(StepState *st, const double *part, int npart, int recompute,
                                           const double *ke_part, int nke, int mode)
{
  __shared__ double red[256];
  __shared__ int go, do_ts, run0, pend0;
  const bool forced = mode == 2 || mode == 4 || mode == 5;
  if (threadIdx.x == 0)
  {
    // after an element error nothing more happens: the failing increment
    // counts, and its endStep never runs (the reference stops inside
    // computeInternalForces)
    run0 = st->running && !st->status;
    pend0 = !forced && st->pending;
    if (st->status) st->halted = 1;
  }
  __syncthreads();
  if (!run0) return;
  const int pending = pend0;
  if (pending && ke_part)
  {
    // Domain::getTotalKineticEnergy of the finished increment, fixed reduction order
    double s = 0.0;
    for (int k = threadIdx.x; k < nke; k += 256) s += ke_part[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) st->ke = red[0];
  }
  if (threadIdx.x == 0)
  {
    do_ts = 0;
    go = 1;
    if (pending)
    {
      st->pending = 0;
      go = end_step(st);
      do_ts = go;
      if (st->plot)
      {
        // the history line of endStep, FE/ExplicitSolverCH.C:459-461: time,
        // the increment's time step, the last computed one, kinetic energy
        double *row = st->plot + 4 * (st->plot_n % st->plot_cap);
        row[0] = st->t;
        row[1] = st->dt;
        row[2] = st->computedTimeStep;
        row[3] = st->ke;
        st->plot_n++;
      }
    }
    if (forced) do_ts = 1;
  }
  __syncthreads();
  if (!go || mode == 3) return;
  double crit = __builtin_inf();
  const int rec = do_ts && (recompute || forced);
  if (rec)
  {
    for (int k = threadIdx.x; k < npart; k += 256) crit = fmin(crit, part[k]);
    red[threadIdx.x] = crit;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
      if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
      __syncthreads();
    }
    crit = red[0];
  }
  if (threadIdx.x != 0) return;
  if (do_ts) compute_time_step(st, rec, crit);
  if (mode == 0 || mode == 4) init_step(st);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void launch_elem_f64(int dim, const ElemArgs &a, bool speed_bc, hipStream_t s)
{
  if (a.ne <= 0) return;
  const bool tr = a.eps != nullptr;
  const long long lanes = a.ne * (dim == 3 ? 8 : 4);
  const int nt = dim == 3 ? kHexNT : 64;
  const unsigned grid = (unsigned)((lanes + nt - 1) / nt);
#define DNL_ELEM(K, B, T)                                                                      \
  do                                                                                           \
  {                                                                                            \
    if (a.law == 2) hipLaunchKernelGGL((K<B, T, true>), dim3(grid), dim3(nt), 0, s, a);       \
    else hipLaunchKernelGGL((K<B, T, false>), dim3(grid), dim3(nt), 0, s, a);                 \
  } while (0)
  if (dim == 3)
  {
    if (speed_bc) { if (tr) DNL_ELEM(k_elem_hex8, true, true); else DNL_ELEM(k_elem_hex8, true, false); }
    else { if (tr) DNL_ELEM(k_elem_hex8, false, true); else DNL_ELEM(k_elem_hex8, false, false); }
  }
  else
  {
    if (speed_bc) { if (tr) DNL_ELEM(k_elem_quad4ax, true, true); else DNL_ELEM(k_elem_quad4ax, true, false); }
    else { if (tr) DNL_ELEM(k_elem_quad4ax, false, true); else DNL_ELEM(k_elem_quad4ax, false, false); }
  }
#undef DNL_ELEM
}

void launch_mass_f64(int dim, const ElemArgs &a, double rho, const double *hn, double *me, hipStream_t s)
{
  const unsigned grid = (unsigned)((a.ne + 255) / 256);
  if (dim == 3) hipLaunchKernelGGL(k_mass_elem<3>, dim3(grid), dim3(256), 0, s, a, rho, hn[0], hn[1], hn[2], hn[3], me);
  else hipLaunchKernelGGL(k_mass_elem<2>, dim3(grid), dim3(256), 0, s, a, rho, hn[0], hn[1], hn[2], hn[3], me);
}

void launch_gather_scalar(const unsigned char *deg, const double *src, double *dst, long long nn, hipStream_t s)
{
  const unsigned grid = (unsigned)((nn + 255) / 256);
  hipLaunchKernelGGL(k_gather_scalar, dim3(grid), dim3(256), 0, s, deg, src, dst, nn);
}

void launch_gather_fint(int dim, const unsigned char *deg, const double *fe, const float *fe32, double *fint,
                        long long nn, hipStream_t s)
{
  const unsigned grid = (unsigned)((nn + 255) / 256);
  if (fe32)
  {
    if (dim == 3) hipLaunchKernelGGL((k_gather_fint<3, float>), dim3(grid), dim3(256), 0, s, deg, fe32, fint, nn);
    else hipLaunchKernelGGL((k_gather_fint<2, float>), dim3(grid), dim3(256), 0, s, deg, fe32, fint, nn);
  }
  else
  {
    if (dim == 3) hipLaunchKernelGGL((k_gather_fint<3, double>), dim3(grid), dim3(256), 0, s, deg, fe, fint, nn);
    else hipLaunchKernelGGL((k_gather_fint<2, double>), dim3(grid), dim3(256), 0, s, deg, fe, fint, nn);
  }
}

unsigned nodal_blocks(long long nn) { return (unsigned)((nn + 255) / 256); }
unsigned courant_blocks(long long ne) { return (unsigned)((ne + 255) / 256); }

void launch_nodal_f64(int dim, const NodalArgs &n, bool track_d, bool has_bcs, hipStream_t s)
{
  const unsigned grid = nodal_blocks(n.nn);
  const bool f32 = n.fe32 != nullptr;
#define DNL_NODAL(D, T, B, F) hipLaunchKernelGGL((k_gather_nodal<D, T, B, F>), dim3(grid), dim3(256), 0, s, n)
#define DNL_NODAL_F(D, T, B) do { if (f32) DNL_NODAL(D, T, B, true); else DNL_NODAL(D, T, B, false); } while (0)
  if (dim == 3)
  {
    if (track_d) { if (has_bcs) DNL_NODAL_F(3, true, true); else DNL_NODAL_F(3, true, false); }
    else { if (has_bcs) DNL_NODAL_F(3, false, true); else DNL_NODAL_F(3, false, false); }
  }
  else
  {
    if (track_d) { if (has_bcs) DNL_NODAL_F(2, true, true); else DNL_NODAL_F(2, true, false); }
    else { if (has_bcs) DNL_NODAL_F(2, false, true); else DNL_NODAL_F(2, false, false); }
  }
#undef DNL_NODAL_F
#undef DNL_NODAL
}

void launch_courant_f64(int dim, const ElemArgs &a, double cs, double *part, hipStream_t s)
{
  const unsigned grid = courant_blocks(a.ne);
  if (dim == 3) hipLaunchKernelGGL(k_courant<3>, dim3(grid), dim3(256), 0, s, a, cs, part);
  else hipLaunchKernelGGL(k_courant<2>, dim3(grid), dim3(256), 0, s, a, cs, part);
}

void launch_min_reduce(const double *part, int npart, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(256), 0, s, part, npart, out);
}

void launch_dt(StepState *st, const double *part, int npart, int recompute, const double *ke_part, int nke, int mode,
               hipStream_t s)
{
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(256), 0, s, st, part, npart, recompute, ke_part, nke, mode);
}

}  // namespace dnl
