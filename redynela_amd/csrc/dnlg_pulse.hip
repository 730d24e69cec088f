// Pulse time step on the GPU: Domain::getMaximumFrequency (FE/Domain.C:816-946).
//
// The reference forms every element's elastic stiffness matrix
// (FE/Element3D.C:242-333 for hex8, FE/Element2D.C:129-222 for the
// axisymmetric quad) from the point dN and detJ of the last
// computeInternalMatrices, then runs a power iteration on M^-1 K: each
// iteration scatters the iterate to the elements, multiplies by Ke with
// LAPACK's reference dgemv, assembles element by element into a zeroed
// vector, divides by the lumped mass, takes the largest magnitude fmax and
// scales the iterate by 1/fmax, until |fmax - fmax0| / fmax < 1e-4.  The
// iterate and fmax0 carry over from one call to the next.
//
// Kernels (one call = pulse_ke with the first iteration's product, then
// pulse_mv, pulse_node and pulse_ctl per iteration; all of them return at
// once when the loop has stopped or converged, so the host may issue
// iterations ahead):
//   k_pulse_begin  reset the iteration counters; start vector when fmax0 == 0
//   k_pulse_ke     Ke of 4 hex8 (16 quad) elements per 64-lane block, two
//                  lanes (one for the quad) per (element, node), each holding
//                  half of that node's rows; per point the lanes share the
//                  columns of C B through LDS.  Ke is stored [8 elements]
//                  [column pair J/2][8 x ND rows][2] so that a row's 16-byte
//                  pairs are adjacent across threads
//   k_pulse_mv     y = Ke x for one row per thread (dgemv order: y = 0, then
//                  y += x_J Ke(I, J) for J in order), y into the element's
//                  ELL slot of the node
//   k_pulse_node   per node: slot sum in element order (the reference's
//                  assembly order), divide by the mass, block max of |val|
//   k_pulse_ctl    max over blocks, 1/fmax, convergence, iteration limit
// The stored iterate is val (before the 1/fmax scaling) and invf = 1/fmax;
// the matvec reads val * invf, which is the reference's ev(i) *= 1/fmax.
#include "dnlg_elem_f64.hpp"
#include "dnlg_pulse.hpp"

namespace dnl {

// Material::getHookeMatrix (FE/Material.C:270-331), row-major: the
// threedimensional 6x6 (DIM 3) or the axisymetric 4x4 (DIM 2)
template <int DIM>
__device__ void hooke(double E, double Nu, double *C)
{
#pragma clang fp contract(off)
  constexpr int NC = DIM == 3 ? 6 : 4;
  for (int k = 0; k < NC * NC; k++) C[k] = 0.0;
  if constexpr (DIM == 3)
  {
    C[0] = 1; C[7] = 1; C[14] = 1;
    C[21] = (1.0 - 2.0 * Nu) / (2. * (1.0 - Nu));
    C[28] = (1.0 - 2.0 * Nu) / (2. * (1.0 - Nu));
    C[35] = (1.0 - 2.0 * Nu) / (2. * (1.0 - Nu));
    C[1] = C[6] = Nu / (1. - Nu);
    C[2] = C[12] = Nu / (1. - Nu);
    C[8] = C[13] = Nu / (1. - Nu);
  }
  else
  {
    C[0] = 1; C[5] = 1; C[15] = 1;
    C[4] = C[1] = Nu / (1. - Nu);
    C[12] = C[13] = Nu / (1. - Nu);
    C[3] = C[7] = Nu / (1. - Nu);
    C[10] = (1.0 - 2.0 * Nu) / (2. * (1.0 - Nu));
  }
  const double f = (E * (1.0 - Nu) / ((1.0 + Nu) * (1.0 - 2.0 * Nu)));
  for (int k = 0; k < NC * NC; k++) C[k] *= f;
}

__global__ __launch_bounds__(256) void k_pulse_begin(PulseArgs p)
{
  if (!p.st->running) return;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const bool fresh = p.ctl->fmax0 == 0.0;  // FE/Domain.C:849-856
  if (fresh && i < p.ndof) p.val[i] = p.ev_init[i];
  if (i == 0)
  {
    if (fresh) p.ctl->invf = 1.0;
    p.ctl->iter = 0;
    p.ctl->done = 0;
  }
}

// Ke of EPB elements per 64-lane block, SPLIT lanes per (element, node):
// lanes (i, h) of an element own the DIM rows of node i, h the h-th of SPLIT
// equal column ranges.  Per point, the lanes of node i compute dN of the
// node and its DIM columns of C B (the C B rows split among them); the block
// shares them through LDS and every lane adds its part of the node's rows of
// B^T C B.  With two lanes per hex8 node a lane holds 36 sums instead of 72,
// which frees registers for more waves per SIMD.  The matrices go to HBM for
// the iterations, and the first iteration's product Ke x is formed here from
// the rows in registers: each row's sum runs over the column ranges in order,
// handed from lane to lane.
template <int NEN, int DIM, int SPLIT>
struct PulseLds
{
  static constexpr int ND = NEN * DIM, NIP = NEN, NC = DIM == 3 ? 6 : 4, EPB = 64 / (NEN * SPLIT);
  double tab[NIP * NEN * 3];               // dN0 (and N for the quad) of the reference element
  double C[NC * NC];
  double geo[EPB][NIP][DIM * DIM + 2];     // inv(J), W (with 2 pi R for the quad), R
  double CB[EPB][NC][ND];
  double loc[EPB][ND];                     // the iterate at the element's dofs
};

template <int NEN, int DIM, int SPLIT>
__global__ __launch_bounds__(64) void k_pulse_ke(PulseArgs p, ElemArgs a)
{
#pragma clang fp contract(off)
  using L = PulseLds<NEN, DIM, SPLIT>;
  constexpr int ND = L::ND, NIP = L::NIP, NC = L::NC, EPB = L::EPB, NT8 = 8 * ND, NJ = ND / SPLIT;
  static_assert(NJ % 2 == 0, "column ranges hold whole pairs");
  __shared__ L s;
  if (!p.st->running) return;
  const int t = threadIdx.x, el = t / (NEN * SPLIT), i = (t / SPLIT) % NEN, h = t % SPLIT;
  const long long e_raw = (long long)blockIdx.x * EPB + el;
  const bool valid = e_raw < a.ne;
  const long long e = valid ? e_raw : a.ne - 1;
  for (int k = t; k < NIP * NEN * 3; k += 64)
  {
    const int pt = k / (NEN * 3), n = (k / 3) % NEN, c = k % 3;
    if constexpr (DIM == 3)
    {
      double g[3];
      hex_dn0(a, pt, n, g);
      s.tab[k] = g[c];
    }
    else
    {
      double g[2];
      quad_dn0(a, pt, n, g);
      s.tab[k] = c < 2 ? g[c] : quad_n(a, pt, n);
    }
  }
  if (t == 0) hooke<DIM>(p.E, p.nu, s.C);
  const int *cn = a.conn + e * NEN;
  // the iterate (ev *= 1/fmax, LA/Vector.C:652-659) for the fused first product
  if (h == 0)
  {
    const double invf = p.ctl->invf;
    const long long node = cn[i];
#pragma unroll
    for (int c = 0; c < DIM; c++) s.loc[el][DIM * i + c] = p.val[node * DIM + c] * invf;
  }
  __syncthreads();
  // point geometry: FE/Element3D.C:75-105, FE/Element2D.C:84-115; lane (i, 0) takes point i
  if (h == 0)
  {
    const int pt = i;
    double *gq = s.geo[el][pt];
    if constexpr (DIM == 3)
    {
      T9 J = {{0, 0, 0, 0, 0, 0, 0, 0, 0}};
      for (int k = 0; k < 8; k++)
      {
        const double *xk = a.x + 3 * (long long)cn[k];
        const double x0 = xk[0], x1 = xk[1], x2 = xk[2];
        const double *g = s.tab + (pt * NEN + k) * 3;
        J.v[0] += g[0] * x0; J.v[1] += g[1] * x0; J.v[2] += g[2] * x0;
        J.v[3] += g[0] * x1; J.v[4] += g[1] * x1; J.v[5] += g[2] * x1;
        J.v[6] += g[0] * x2; J.v[7] += g[1] * x2; J.v[8] += g[2] * x2;
      }
      const double detJ = (J.v[0] * J.v[4] * J.v[8]) + (J.v[3] * J.v[7] * J.v[2]) + (J.v[6] * J.v[1] * J.v[5]) -
                          (J.v[6] * J.v[4] * J.v[2]) - (J.v[0] * J.v[7] * J.v[5]) - (J.v[3] * J.v[1] * J.v[8]);
      T9 inv;
      inv3(J, inv);  // a singular J was reported by the element pass that precedes every call
      for (int c = 0; c < 9; c++) gq[c] = inv.v[c];
      gq[9] = 1.0 * detJ;
      gq[10] = 0.0;
    }
    else
    {
      double J0 = 0, J1 = 0, J2 = 0, J3 = 0, R = 0.;
      for (int k = 0; k < 4; k++)
      {
        const double *xk = a.x + 3 * (long long)cn[k];
        const double x0 = xk[0], x1 = xk[1];
        const double *g = s.tab + (pt * NEN + k) * 3;
        J0 += g[0] * x0; J1 += g[1] * x0;
        J2 += g[0] * x1; J3 += g[1] * x1;
      }
      for (int k = 0; k < 4; k++) R += s.tab[(pt * NEN + k) * 3 + 2] * a.x[3 * (long long)cn[k]];
      const double detJ = (J0 * J3) - (J1 * J2);
      gq[0] = J3 / detJ; gq[1] = -J1 / detJ; gq[2] = -J2 / detJ; gq[3] = J0 / detJ;
      double W = 1.0 * detJ;
      W *= kDeuxPI * R;  // 2 * PI * R: doubling is exact
      gq[4] = W;
      gq[5] = R;
    }
  }
  double k[DIM][NJ];
#pragma unroll
  for (int c = 0; c < DIM; c++)
#pragma unroll
    for (int J = 0; J < NJ; J++) k[c][J] = 0.0;
  const double *C = s.C;
  const int j0 = h * NJ;  // first column of this lane's range
  for (int pt = 0; pt < NIP; pt++)
  {
    __syncthreads();  // geo written / previous point's CB consumed
    const double *gq = s.geo[el][pt];
    const double W = gq[DIM * DIM];
    double d[DIM];
    dgemm_row<DIM>(s.tab + (pt * NEN + i) * 3, gq, d);
    // the DIM columns of C B of node i, rows split among the node's lanes
    if constexpr (DIM == 3)
    {
      const double d0 = d[0], d1 = d[1], d2 = d[2];
#pragma unroll
      for (int r = 0; r < 6; r++)
        if (r % SPLIT == h)
        {
          s.CB[el][r][3 * i] = (C[r * 6 + 0] * d0 + C[r * 6 + 3] * d1 + C[r * 6 + 5] * d2);
          s.CB[el][r][3 * i + 1] = (C[r * 6 + 1] * d1 + C[r * 6 + 3] * d0 + C[r * 6 + 4] * d2);
          s.CB[el][r][3 * i + 2] = (C[r * 6 + 2] * d2 + C[r * 6 + 4] * d1 + C[r * 6 + 5] * d0);
        }
    }
    else
    {
      const double d0 = d[0], d1 = d[1], R = gq[5], Ni = s.tab[(pt * NEN + i) * 3 + 2];
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (r % SPLIT == h)
        {
          double v = (C[r * 4 + 0] * d0 + C[r * 4 + 2] * d1);
          v += C[r * 4 + 3] * Ni / R;
          s.CB[el][r][2 * i] = v;
          s.CB[el][r][2 * i + 1] = (C[r * 4 + 1] * d1 + C[r * 4 + 2] * d0);
        }
    }
    __syncthreads();
    // this lane's columns of the DIM rows of node i of B^T C B
    if constexpr (DIM == 3)
    {
      const double a0 = d[0], a1 = d[1], a2 = d[2];
#pragma unroll
      for (int J = 0; J < NJ; J++)
      {
        const int jc = j0 + J;
        const double b0 = s.CB[el][0][jc], b1 = s.CB[el][1][jc], b2 = s.CB[el][2][jc];
        const double b3 = s.CB[el][3][jc], b4 = s.CB[el][4][jc], b5 = s.CB[el][5][jc];
        k[0][J] += (a0 * b0 + a1 * b3 + a2 * b5) * W;
        k[1][J] += (a1 * b1 + a0 * b3 + a2 * b4) * W;
        k[2][J] += (a2 * b2 + a1 * b4 + a0 * b5) * W;
      }
    }
    else
    {
      const double a0 = d[0], a1 = d[1], R = gq[5], Ni = s.tab[(pt * NEN + i) * 3 + 2];
#pragma unroll
      for (int J = 0; J < NJ; J++)
      {
        const int jc = j0 + J;
        const double b0 = s.CB[el][0][jc], b1 = s.CB[el][1][jc], b2 = s.CB[el][2][jc], b3 = s.CB[el][3][jc];
        k[0][J] += (a0 * b0 + a1 * b2) * W;
        k[0][J] += b3 * Ni / R * W;
        k[1][J] += (a1 * b1 + a0 * b2) * W;
      }
    }
  }
  // Ke to HBM: [e / 8][column pair][8 elements x ND rows][2]
  if (valid)
  {
    double2 *ko = reinterpret_cast<double2 *>(p.ke + (e / 8) * ND * NT8);
    const int row0 = (int)(e % 8) * ND + DIM * i;
#pragma unroll
    for (int c = 0; c < DIM; c++)
#pragma unroll
      for (int J = 0; J < NJ; J += 2)
        ko[(long long)((j0 + J) / 2) * NT8 + row0 + c] = make_double2(k[c][J], k[c][J + 1]);
  }
  // the first iteration's product, LAPACK dgemv order (y = 0, y += x_J Ke(I, J)
  // for J in order): lane (i, h) continues the sums lane (i, h - 1) left
#pragma unroll
  for (int c = 0; c < DIM; c++)
  {
    double y = 0.0;
#pragma unroll
    for (int r = 0; r < SPLIT; r++)
    {
      if (r > 0) y = __shfl(y, t - 1, 64);  // lane (i, r) takes the sum of lane (i, r - 1)
      if (h == r)
#pragma unroll
        for (int J = 0; J < NJ; J++)
        {
          const double tj = 1.0 * s.loc[el][j0 + J];
          y = y + tj * k[c][J];
        }
    }
    if (valid && h == SPLIT - 1) p.y[(long long)p.fpos[e * NEN + i] * DIM + c] = y;
  }
}

template <int NEN, int DIM>
__global__ __launch_bounds__(8 * NEN * DIM) void k_pulse_mv(PulseArgs p)
{
#pragma clang fp contract(off)
  constexpr int ND = NEN * DIM, NT = 8 * ND;
  __shared__ double loc[8][ND];
  if (!p.st->running || p.ctl->done) return;
  const int t = threadIdx.x, el = t / ND, I = t % ND;
  const long long e_raw = (long long)blockIdx.x * 8 + el;
  const bool valid = e_raw < p.ne;
  const long long e = valid ? e_raw : p.ne - 1;
  // the row of Ke first, so that its loads overlap the gather of the iterate
  const double2 *kr = reinterpret_cast<const double2 *>(p.ke + (long long)blockIdx.x * ND * NT) + t;
  double2 kv[ND / 2];
#pragma unroll
  for (int j = 0; j < ND / 2; j++) kv[j] = kr[(long long)j * NT];
  const double invf = p.ctl->invf;
  const int node = p.conn[e * NEN + I / DIM];
  loc[el][I] = p.val[(long long)node * DIM + I % DIM] * invf;  // ev(i) *= 1/fmax, LA/Vector.C:652-659
  __syncthreads();
  double y = 0.0;
#pragma unroll
  for (int j = 0; j < ND / 2; j++)
  {
    const double t0 = 1.0 * loc[el][2 * j], t1 = 1.0 * loc[el][2 * j + 1];
    y = y + t0 * kv[j].x;
    y = y + t1 * kv[j].y;
  }
  if (valid) p.y[(long long)p.fpos[e * NEN + I / DIM] * DIM + I % DIM] = y;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_pulse_node(PulseArgs p)
{
#pragma clang fp contract(off)
  __shared__ double red[256];
  if (!p.st->running || p.ctl->done) return;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  double mx = 0.0;
  if (i < p.nn)
  {
    const int d = p.deg[i];
    const double m = p.mass[i];
#pragma unroll
    for (int j = 0; j < DIM; j++)
    {
      double acc = 0.0;
      for (int k = 0; k < d; k++) acc += p.y[((long long)k * p.nn + i) * DIM + j];
      double val = 0.0;
      val += acc;
      val /= m;
      if (fabs(val) > mx) mx = fabs(val);
      p.val[i * DIM + j] = val;
    }
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_pulse_ctl(PulseArgs p, int nblk)
{
#pragma clang fp contract(off)
  __shared__ double red[256];
  if (!p.st->running || p.ctl->done) return;
  double mx = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256)
    if (p.part[k] > mx) mx = p.part[k];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  PulseCtl *c = p.ctl;
  const double fmax = red[0];
  const int it = ++c->iter;
  c->invf = 1.0 / fmax;
  const double conv = fabs(fmax - c->fmax0) / fmax;
  c->conv = conv;
  c->fmax = fmax;
  c->fmax0 = fmax;
  if (it > PULSE_MAX_ITER)  // FE/Domain.C:936-939, checked after the convergence test
  {
    c->done = 1;
    c->failed = 1;
    raise_status(p.st, ST_PULSE, -1);
  }
  else if (conv < PULSE_CONVERGENCE)
  {
    c->done = 1;
    p.st->pulse_f = sqrt(fmax);
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void launch_pulse_begin(const PulseArgs &p, hipStream_t s)
{
  const unsigned grid = (unsigned)((p.ndof + 255) / 256);
  hipLaunchKernelGGL(k_pulse_begin, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, p);
}

void launch_pulse_ke(int dim, const PulseArgs &p, const ElemArgs &a, hipStream_t s)
{
  // 4 hex8 (two lanes per node) or 16 quad elements per 64-lane block; the first iteration's
  // product is formed here, its node and control kernels follow
  const unsigned grid_n = (unsigned)((p.nn + 255) / 256);
  if (dim == 3)
  {
    hipLaunchKernelGGL((k_pulse_ke<8, 3, 2>), dim3((unsigned)((a.ne + 3) / 4)), dim3(64), 0, s, p, a);
    hipLaunchKernelGGL((k_pulse_node<3>), dim3(grid_n), dim3(256), 0, s, p);
  }
  else
  {
    hipLaunchKernelGGL((k_pulse_ke<4, 2, 1>), dim3((unsigned)((a.ne + 15) / 16)), dim3(64), 0, s, p, a);
    hipLaunchKernelGGL((k_pulse_node<2>), dim3(grid_n), dim3(256), 0, s, p);
  }
  hipLaunchKernelGGL(k_pulse_ctl, dim3(1), dim3(256), 0, s, p, (int)grid_n);
}

void launch_pulse_iteration(int dim, const PulseArgs &p, hipStream_t s)
{
  const unsigned grid_e = (unsigned)((p.ne + 7) / 8), grid_n = (unsigned)((p.nn + 255) / 256);
  if (dim == 3)
  {
    hipLaunchKernelGGL((k_pulse_mv<8, 3>), dim3(grid_e), dim3(192), 0, s, p);
    hipLaunchKernelGGL((k_pulse_node<3>), dim3(grid_n), dim3(256), 0, s, p);
  }
  else
  {
    hipLaunchKernelGGL((k_pulse_mv<4, 2>), dim3(grid_e), dim3(64), 0, s, p);
    hipLaunchKernelGGL((k_pulse_node<2>), dim3(grid_n), dim3(256), 0, s, p);
  }
  hipLaunchKernelGGL(k_pulse_ctl, dim3(1), dim3(256), 0, s, p, (int)grid_n);
}

}  // namespace dnl
