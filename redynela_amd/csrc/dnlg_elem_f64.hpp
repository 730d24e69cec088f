// Device functions of the FP64 element pass, the nodal boundary
// conditions and the Courant lengths, shared by the grid kernels
// (dnlg_f64.hip) and the persistent per-bar kernel (dnlg_batch.hip).
// Every translation unit that includes this header must be compiled with
// -ffp-contract=off: the arithmetic is the reference's, operation for operation.
#pragma once
#include "dnlg_math.hpp"
#include "dnlg_internal.hpp"
#include "dnlg_tables.hpp"

namespace dnl {


// signs of the reference element's node / Gauss point number k:
// (-,-,-),(+,-,-),(+,+,-),(-,+,-),(-,-,+),(+,-,+),(+,+,+),(-,+,+)
__device__ __forceinline__ int sg0(int k) { int q = k & 3; return (q == 1 || q == 2) ? 1 : -1; }
__device__ __forceinline__ int sg1(int k) { return ((k & 3) >= 2) ? 1 : -1; }
__device__ __forceinline__ int sg2(int k) { return (k >= 4) ? 1 : -1; }

// hex8 dN0(ip, node, d): sign of node along d times (2+r3)/12, 1/12 or
// (2-r3)/12 by how many of the two other directions the node and the Gauss
// point share a sign (FE/ElHex8n3D.C:81-116 tables)
__device__ __forceinline__ void hex_dn0(const ElemArgs &a, int ip, int k, double *g)
{
  const int a0 = sg0(ip) == sg0(k), a1 = sg1(ip) == sg1(k), a2 = sg2(ip) == sg2(k);
  const int o0 = a1 + a2, o1 = a0 + a2, o2 = a0 + a1;
  const double m0 = o0 == 2 ? a.c2 : (o0 == 1 ? a.c1 : a.c0);
  const double m1 = o1 == 2 ? a.c2 : (o1 == 1 ? a.c1 : a.c0);
  const double m2 = o2 == 2 ? a.c2 : (o2 == 1 ? a.c1 : a.c0);
  g[0] = sg0(k) > 0 ? m0 : -m0;
  g[1] = sg1(k) > 0 ? m1 : -m1;
  g[2] = sg2(k) > 0 ? m2 : -m2;
}
// device copies of the compile-time shape tables (dnlg_tables.hpp)
static __constant__ HexTab kHexTabDev = make_hex_tab();
static __constant__ QuadTab kQuadTabDev = make_quad_tab();

// quad4 dN0(ip, node, d) = sign(node,d) * (3 +- r3)/12 (FE/ElQua4nAx.C:51-73)
__device__ __forceinline__ void quad_dn0(const ElemArgs &a, int ip, int k, double *g)
{
  const double m0 = (sg1(ip) == sg1(k)) ? a.c1 : a.c0;
  const double m1 = (sg0(ip) == sg0(k)) ? a.c1 : a.c0;
  g[0] = sg0(k) > 0 ? m0 : -m0;
  g[1] = sg1(k) > 0 ? m1 : -m1;
}
// quad4 N(ip, node): (2+r3)/6, 1/6, (2-r3)/6 by shared signs
__device__ __forceinline__ double quad_n(const ElemArgs &a, int ip, int k)
{
  const int ag = (sg0(ip) == sg0(k)) + (sg1(ip) == sg1(k));
  return ag == 2 ? a.n2 : (ag == 1 ? a.n1 : a.n0);
}

// a double at byte offset off of a uniform base p, through a raw buffer
// resource (the base in scalar registers, the 32-bit offset in one vector
// register; no range clamp: 0x00020000 is the gfx9 descriptor word 3)
typedef unsigned int dnl_u2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double buf_ld(const double *p, unsigned off)
{
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, -1, 0x00020000);
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_st(double *p, unsigned off, double v)
{
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(p, 0, -1, 0x00020000);
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(dnl_u2, v), r, off, 0, 0);
}

__device__ __forceinline__ void raise_status(StepState *st, int code, long long e)
{
  if (atomicCAS(&st->status, 0, code) == 0) st->err_elem = (int)e;
}

// J2 radial return and elastic predictor, FE/Element.C:248-348
__device__ __forceinline__ bool stress_update(const ElemArgs &a, const T9 &dEps, T9 &S, double &evp, T9 *epl)
{
  const double trm = (dEps.v[0] + dEps.v[4] + dEps.v[8]) / 3.0;
  T9 dev = dEps;
  dev.v[0] = dEps.v[0] - trm;
  dev.v[4] = dEps.v[4] - trm;
  dev.v[8] = dEps.v[8] - trm;
  const double g2 = 2.0 * a.G;
#pragma unroll
  for (int k = 0; k < 9; k++) S.v[k] += g2 * dev.v[k];
  if (!a.yield) return true;
  // pow(evp, n) and pow(evp, n - 1) share one logarithm (DetPowBase: the same doubles)
  DetPowBase b0(evp + kSqrt23 * 0.0);
  const double Sigv0 = a.A + a.B * b0.pow(a.n);
  const double Strial = kSqrt32 * sqrt(sumsq(S));
  if (!(Strial > Sigv0)) return true;
  double h = a.n * a.B * b0.pow(a.nm1);
  double gamma = (sqrt(sumsq(S)) - kSqrt23 * Sigv0) / (2.0 * a.G * (1.0 + (h) / (3.0 * a.G)));
  int iteration = 0;
  while (true)
  {
    const double ev = evp + kSqrt23 * gamma;
    DetPowBase bi(ev);
    const double Sigv = a.A + a.B * bi.pow(a.n);
    h = a.n * a.B * bi.pow(a.nm1);
    const double dfun = kSqrt32 * 2.0 * a.G + kSqrt23 * h;
    const double fun = dfun * gamma - (Strial - Sigv + h * kSqrt23 * gamma);
    bool run = dabs(fun / Sigv) > kRRPrec;
    if (run)
    {
      double dgamma = -fun / dfun;
      if (dgamma < 0 && dabs(dgamma) > gamma) dgamma = -gamma * 0.9;
      gamma += dgamma;
    }
    iteration++;
    if (iteration > kRRMaxIt) return false;
    if (!run) break;
  }
  const double nrm = sqrt(sumsq(S));
  T9 n1;
#pragma unroll
  for (int k = 0; k < 9; k++) n1.v[k] = S.v[k] / nrm;
  evp += kSqrt23 * gamma;
  const double c = 2.0 * a.G * gamma;
#pragma unroll
  for (int k = 0; k < 9; k++) S.v[k] -= c * n1.v[k];
  if (epl)
#pragma unroll
    for (int k = 0; k < 9; k++) epl->v[k] += gamma * n1.v[k];
  return true;
}

// Johnson-Cook yield stress with rate and thermal factors (law 2, an
// extension: FE/IsoHardJohnsonCook.h:51-68 documents the law, the reference
// evaluates only A + B evp^n).  Newton on the consistency condition
// Strial - sqrt(3/2) 2G gamma = sigma_y(evp + sqrt(2/3) gamma, epsdot, T),
// same stop as FE/Element.C:304-331, kept inside a sign bracket (bisection
// when a step leaves it: the rate factor has a kink at epsdot = epsdot0);
// adiabatic heating with the Taylor-Quinney fraction after convergence.
__device__ __forceinline__ double jc_theta(const ElemArgs &a, double T)
{
  if (a.chi_rho_cp == 0.0) return 1.0;
  if (T < a.T0) return 1.0;
  if (T >= a.Tm) return 0.0;
  return 1.0 - det_pow((T - a.T0) / (a.Tm - a.T0), a.m);
}

__device__ __forceinline__ bool stress_update_jc(const ElemArgs &a, const T9 &dEps, T9 &S, double &evp, double &T,
                                                 double dt, T9 *epl)
{
  const double trm = (dEps.v[0] + dEps.v[4] + dEps.v[8]) / 3.0;
  T9 dev = dEps;
  dev.v[0] = dEps.v[0] - trm;
  dev.v[4] = dEps.v[4] - trm;
  dev.v[8] = dEps.v[8] - trm;
  const double g2 = 2.0 * a.G;
#pragma unroll
  for (int k = 0; k < 9; k++) S.v[k] += g2 * dev.v[k];
  const double theta = jc_theta(a, T);
  DetPowBase b0(evp);
  const double sy0 = (a.A + a.B * b0.pow(a.n)) * theta;
  const double nrm = sqrt(sumsq(S));
  const double Strial = kSqrt32 * nrm;
  if (!(Strial > sy0)) return true;
  const double h0 = a.n * a.B * b0.pow(a.nm1) * theta;
  double gamma = (nrm - kSqrt23 * sy0) / (2.0 * a.G * (1.0 + h0 / (3.0 * a.G)));
  double lo = 0.0, hi = Strial / (kSqrt32 * 2.0 * a.G);
  if (!(gamma > lo && gamma < hi)) gamma = 0.5 * (lo + hi);
  double sy = sy0;
  int it = 0;
  while (true)
  {
    const double ev = evp + kSqrt23 * gamma;
    DetPowBase bi(ev);
    const double base = a.A + a.B * bi.pow(a.n);
    const double dbase = a.n * a.B * bi.pow(a.nm1);
    double rf = 1.0, drf = 0.0;
    if (a.C != 0.0 && gamma > 0.0)
    {
      const double epsd = kSqrt23 * gamma / dt;
      if (epsd >= a.epsd0)
      {
        rf = 1.0 + a.C * det_log(epsd / a.epsd0);
        drf = a.C / gamma;
      }
    }
    sy = base * rf * theta;
    const double dsy = (kSqrt23 * dbase * rf + base * drf) * theta;
    const double res = Strial - kSqrt32 * 2.0 * a.G * gamma - sy;
    it++;
    if (dabs(res / sy) <= kRRPrec) break;
    if (it > kRRMaxIt) return false;
    if (res > 0.0) lo = gamma; else hi = gamma;
    const double gn = gamma + res / (kSqrt32 * 2.0 * a.G + dsy);
    gamma = (gn > lo && gn < hi) ? gn : 0.5 * (lo + hi);
  }
  T9 n1;
#pragma unroll
  for (int k = 0; k < 9; k++) n1.v[k] = S.v[k] / nrm;
  const double dep = kSqrt23 * gamma;
  evp += dep;
  const double c = 2.0 * a.G * gamma;
#pragma unroll
  for (int k = 0; k < 9; k++) S.v[k] -= c * n1.v[k];
  if (epl)
#pragma unroll
    for (int k = 0; k < 9; k++) epl->v[k] += gamma * n1.v[k];
  T += a.chi_rho_cp * sy * dep;
  return true;
}

// ---------------------------------------------------------------------------
// LDS of the element kernels: the shape-function table of the reference
// element (filled once per block) and the exchange of the points' data for
// the nodal force sums: hex8 exchanges J^-1 (9 doubles; the node lane forms
// dN = dN0 J^-1 itself) so that 3 waves fit per SIMD, quad4 the shape
// gradients dN themselves.  Rows have a pitch of NT + 1 doubles against bank
// conflicts.
// ---------------------------------------------------------------------------
template <int NT>
struct HexLds
{
  double tab[8 * 8 * 3 + 4];  // dN0(ip, k, d) at hex_tab(ip, k) + d
  double x[19][NT + 1];       // rows 0..8: J^-1 of the point; 9+c: Sig (9..14 first carry the element's
                              // node coordinates and increments); 18: w detJ; columns hex_col(lane)
};
// LDS banks: one double spans two of the 64 four-byte banks, so doubles i and
// i + 32 collide.  A hex8 exchange read has 8 distinct addresses per wave, one
// per element (lanes 8m + q, m = 0..7, same q); element m + 4 would land on
// element m's banks, so the column of lane t has bit 2 flipped when bit 5 of
// t is set (a permutation within each wave: its writes stay conflict-free).
// The table is read with 8 distinct Gauss points pt (phase 1) or 8 distinct
// nodes (phase 2); 4 doubles of padding after point 3 separate pt from pt + 4.
__device__ __forceinline__ int hex_col(int t) { return t ^ ((t >> 3) & 4); }
__device__ __forceinline__ int hex_tab(int ip, int k) { return (ip * 8 + k) * 3 + (ip >= 4 ? 4 : 0); }
template <int NT>
struct QuadLds
{
  double tab[4 * 4 * 3];  // dN0(ip, k, 0..1), N(ip, k)
  double x[15][NT + 1];   // rows 2k+d: dN(k, d); 8..12: Sig xx, xy, yx, yy, zz; 13: w detJ; 14: r
};

// fill the quad4 table from the compile-time copy (all threads of the block;
// a barrier must follow); the hex8 kernel copies kHexTabDev itself
template <int NT>
__device__ __forceinline__ void quad_fill_table(QuadLds<NT> &L)
{
  for (int i = threadIdx.x; i < 4 * 4 * 3; i += NT) L.tab[i] = kQuadTabDev.v[i];
}

// LDS exchange among the lanes of one wave: the writes of every lane before
// the reads of every lane (a wave's LDS operations are performed in order;
// the fences keep the compiler from moving accesses across the barrier)
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// hex8 element kernel: lane = 8*element + integration point
// ---------------------------------------------------------------------------
// The lane offset of point tid of the workgroup whose first point is blk0: tid
// itself, or for a lane past the end of the mesh the same point of the last
// element (in bounds, never written).  The points of the workgroup that exist
// are a multiple of 8, at least 8.
template <int NT>
__device__ __forceinline__ unsigned hex_lane_offset(const ElemArgs &a, long long blk0, int tid)
{
  const long long nvalid = a.ne * 8 - blk0;
  return (tid < nvalid ? (unsigned)tid : (unsigned)(nvalid - 8) + (unsigned)(tid & 7)) & (unsigned)(NT - 1);
}
// ---------------------------------------------------------------------------
// One integration point of one element per lane; lanes 8k..8k+7 of a wave
// hold one element.  The workgroup's NT lanes take the points blk0 ..
// blk0 + NT - 1 (blk0 a multiple of NT); tid indexes them and the exchange
// rows of L (the table must be filled).  JC: the Johnson-Cook law
// (ElemArgs::law == 2) instead of A + B evp^n.
//
// Point and element arrays are addressed from the workgroup's first point,
// a uniform base that stays in scalar registers, plus one 32-bit lane offset
// shared by every access, instead of a 64-bit address per array and
// component in vector registers.
template <bool SPEED_BC, bool TRACK, bool JC, int NT>
__device__ __forceinline__ void elem_hex8_body(const ElemArgs &a, long long blk0, int tid, unsigned lo, int mynode,
                                               HexLds<NT> &Ls)
{
  static_assert((NT & (NT - 1)) == 0 && NT % 64 == 0, "whole waves, a power of two");
  StepState *st = a.st;
  const int wl = tid & 63;
  const int pt = wl & 7, base = wl & ~7, lbase = tid & ~7;
  const long long nipt = a.ne * 8;
  const bool valid = tid < nipt - blk0;
  const long long e = (blk0 + lo) >> 3;
  // component c of S through a raw buffer view of the workgroup's points
  // (base a.S + c * nipt + blk0, in scalar registers), all nine at the one
  // lane offset: plain pointers would make the compiler fold blk0 + c * nipt
  // into a 64-bit address per lane and component, held from load to store
  const unsigned lob = lo * 8u;
  // lane pt of the element has fetched local node pt (mynode, the caller's
  // connectivity load); it loads the node's coordinates and Current increment
  // and the element's 8 lanes share them through LDS rows 9..14, which hold
  // Sig and w detJ only later in the pass
  {
    const double *xk = a.x + 3 * (long long)mynode;
    const double *uk = a.du + 3 * (long long)mynode;
    double u0 = uk[0], u1 = uk[1], u2 = uk[2];
    if (SPEED_BC)
    {
      const unsigned char m = a.bcs[mynode];
      if (m)
      {
        const double *sp = a.bcs_v + 3 * (long long)mynode;
        if (m & 1) u0 = sp[0] * st->dt;
        if (m & 2) u1 = sp[1] * st->dt;
        if (m & 4) u2 = sp[2] * st->dt;
      }
    }
    const int col = hex_col(tid);
    Ls.x[9][col] = xk[0];
    Ls.x[10][col] = xk[1];
    Ls.x[11][col] = xk[2];
    Ls.x[12][col] = u0;
    Ls.x[13][col] = u1;
    Ls.x[14][col] = u2;
  }
  wave_lds_sync();

  // Jacobian at this Gauss point, FE/Element3D.C:75-89.  The reference's
  // sums start at +0; these start at node 0's product, which gives the same
  // double unless all eight products of an entry are -0, i.e. unless all
  // eight nodes share a zero coordinate -- a flat element, whose singular J
  // fails the inverse either way.
  T9 J;
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    const int ck = hex_col(lbase + k);
    const double x0 = Ls.x[9][ck], x1 = Ls.x[10][ck], x2 = Ls.x[11][ck];
    const double *tg = Ls.tab + hex_tab(pt, k);
    const double g[3] = {tg[0], tg[1], tg[2]};
    if (k == 0)
    {
      J.v[0] = g[0] * x0; J.v[1] = g[1] * x0; J.v[2] = g[2] * x0;
      J.v[3] = g[0] * x1; J.v[4] = g[1] * x1; J.v[5] = g[2] * x1;
      J.v[6] = g[0] * x2; J.v[7] = g[1] * x2; J.v[8] = g[2] * x2;
      continue;
    }
    J.v[0] += g[0] * x0; J.v[1] += g[1] * x0; J.v[2] += g[2] * x0;
    J.v[3] += g[0] * x1; J.v[4] += g[1] * x1; J.v[5] += g[2] * x1;
    J.v[6] += g[0] * x2; J.v[7] += g[1] * x2; J.v[8] += g[2] * x2;
  }
  const double detJ = (J.v[0] * J.v[4] * J.v[8]) + (J.v[3] * J.v[7] * J.v[2]) + (J.v[6] * J.v[1] * J.v[5]) -
                      (J.v[6] * J.v[4] * J.v[2]) - (J.v[0] * J.v[7] * J.v[5]) - (J.v[3] * J.v[1] * J.v[8]);
  T9 inv;
  if (!inv3(J, inv) && valid) raise_status(st, ST_DETJ, e);
#pragma unroll
  for (int c = 0; c < 9; c++) Ls.x[c][hex_col(tid)] = inv.v[c];

  // F = I + sum_k dN(k,:) (x) du_k, FE/Element3D.C:182-205
  T9 F = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    const int ck = hex_col(lbase + k);
    const double u0 = Ls.x[12][ck], u1 = Ls.x[13][ck], u2 = Ls.x[14][ck];
    const double *tg = Ls.tab + hex_tab(pt, k);
    const double g[3] = {tg[0], tg[1], tg[2]};
    double d[3];
    dgemm_row<3>(g, inv.v, d);
    F.v[0] += d[0] * u0; F.v[1] += d[1] * u0; F.v[2] += d[2] * u0;
    F.v[3] += d[0] * u1; F.v[4] += d[1] * u1; F.v[5] += d[2] * u1;
    F.v[6] += d[0] * u2; F.v[7] += d[1] * u2; F.v[8] += d[2] * u2;
  }
  wave_lds_sync();  // every lane has read rows 9..14 before any writes Sig there

  T9 dEps, R;
  if (!polar_lnU(F, dEps, R) && valid) raise_status(st, ST_POLAR, e);

  T9 S;
#pragma unroll
  for (int c = 0; c < 9; c++) S.v[c] = buf_ld(a.S + c * nipt + blk0, lob);
  double evp = (a.evp + blk0)[lo];
  // the strain diagnostics go through buffer views like S; the return needs
  // only EpsPlas, and Eps (+ dEps) is read after it
  T9 eps, epl;
  if (TRACK)
#pragma unroll
    for (int c = 0; c < 9; c++) epl.v[c] = buf_ld(a.epsplas + c * nipt + blk0, lob);
  const double tr = dEps.v[0] + dEps.v[4] + dEps.v[8];
  double Tip = 0.0;
  bool rr_ok;
  if (JC)  // law 2, a separate instance: its registers do not weigh on law 1
  {
    Tip = (a.T + blk0)[lo];
    rr_ok = stress_update_jc(a, dEps, S, evp, Tip, st->dt, TRACK ? &epl : nullptr);
  }
  else
    rr_ok = stress_update(a, dEps, S, evp, TRACK ? &epl : nullptr);
  if (!rr_ok && valid) raise_status(st, ST_RR, e);

  // pressure averaged over the element's points in point order, FE/Element.C:353-355
  double pressure = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) pressure += __shfl(tr, base + q, 64);
  pressure /= 8;
  const double p = (a.p + (blk0 >> 3))[lo >> 3] + a.K * pressure;

  // RFI and Sig = S + p I, FE/Element.C:362-401
  S = mul(R, mul_bt(S, R));
  if (TRACK)
  {
#pragma unroll
    for (int c = 0; c < 9; c++) eps.v[c] = buf_ld(a.eps + c * nipt + blk0, lob) + dEps.v[c];
    eps = mul(R, mul_bt(eps, R));
    epl = mul(R, mul_bt(epl, R));
  }
  T9 Sig;
#pragma unroll
  for (int k = 0; k < 9; k++) Sig.v[k] = S.v[k] + p * ((k % 4 == 0) ? 1.0 : 0.0);
  const double WxdJ = 1.0 * detJ;

  if (valid)
  {
#pragma unroll
    for (int c = 0; c < 9; c++) buf_st(a.S + c * nipt + blk0, lob, S.v[c]);
    (a.evp + blk0)[lo] = evp;
    if (JC) (a.T + blk0)[lo] = Tip;
    if (pt == 0) (a.p + (blk0 >> 3))[lo >> 3] = p;
    if (TRACK)
#pragma unroll
      for (int c = 0; c < 9; c++)
      {
        buf_st(a.eps + c * nipt + blk0, lob, eps.v[c]);
        buf_st(a.epsplas + c * nipt + blk0, lob, epl.v[c]);
        buf_st(a.deps + c * nipt + blk0, lob, dEps.v[c]);
      }
  }

  // exchange Sig and w detJ of the 8 points of the element (dN is in place);
  // the element's 8 lanes sit in one wave, whose LDS operations complete in
  // order, so a wave-level barrier orders the writes before the reads
#pragma unroll
  for (int c = 0; c < 9; c++) Ls.x[9 + c][hex_col(tid)] = Sig.v[c];
  Ls.x[18][hex_col(tid)] = WxdJ;
  wave_lds_sync();

  // nodal force of local node L = pt, accumulated point by point as
  // Element::computeInternalForces does, FE/Element.C:178-202
  const int L = pt;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++)
  {
    const int src = hex_col(lbase + q);
    double sq[9];
#pragma unroll
    for (int c = 0; c < 9; c++) sq[c] = Ls.x[9 + c][src];
    const double W = Ls.x[18][src];
    double iq[9];
#pragma unroll
    for (int c = 0; c < 9; c++) iq[c] = Ls.x[c][src];
    const double *tg = Ls.tab + hex_tab(q, L);
    const double g[3] = {tg[0], tg[1], tg[2]};
    double d[3];
    dgemm_row<3>(g, iq, d);
#pragma unroll
    for (int j = 0; j < 3; j++) f0 -= d[j] * sq[3 * j + 0] * W;
#pragma unroll
    for (int j = 0; j < 3; j++) f1 -= d[j] * sq[3 * j + 1] * W;
#pragma unroll
    for (int j = 0; j < 3; j++) f2 -= d[j] * sq[3 * j + 2] * W;
  }
  if (valid)
  {
    double *fo = a.fe + (long long)(a.fpos + blk0)[lo] * 3;
    fo[0] = f0; fo[1] = f1; fo[2] = f2;
  }
}

// ---------------------------------------------------------------------------
// axisymmetric quad4 element kernel: lane = 4*element + integration point
// ---------------------------------------------------------------------------
template <bool SPEED_BC, bool TRACK, bool JC, int NT>
__device__ __forceinline__ void elem_quad4ax_body(const ElemArgs &a, long long e_raw, long long e_end, int tid,
                                                  QuadLds<NT> &Ls)
{
  StepState *st = a.st;
  const int wl = tid & 63;
  const int pt = wl & 3, base = wl & ~3, lbase = tid & ~3;
  const bool valid = e_raw < e_end;
  const long long e = valid ? e_raw : e_end - 1;
  const long long nipt = a.ne * 4, ip = e * 4 + pt;
  const auto *cn = gmem(a.conn) + e * 4;
  int nd[4];
#pragma unroll
  for (int k = 0; k < 4; k++) nd[k] = cn[k];

  // FE/Element2D.C:84-115 and getRadius_atIntPoint, FE/Element2D.C:233-246.
  // The reference's sums start at +0; these start at node 0's product, the
  // same double unless all four products of a sum are -0, which needs all four
  // nodes at a zero coordinate: a flat element (J singular either way) or an
  // element lying on the axis (radius).
  double J0, J1, J2, J3, Rr;
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    const auto *xk = gmem(a.x) + 3 * (long long)nd[k];
    const double x0 = xk[0], x1 = xk[1];
    const double *tg = Ls.tab + (pt * 4 + k) * 3;
    const double g[2] = {tg[0], tg[1]};
    if (k == 0)
    {
      J0 = g[0] * x0; J1 = g[1] * x0;
      J2 = g[0] * x1; J3 = g[1] * x1;
      Rr = tg[2] * x0;
      continue;
    }
    J0 += g[0] * x0; J1 += g[1] * x0;
    J2 += g[0] * x1; J3 += g[1] * x1;
  }
#pragma unroll
  for (int k = 1; k < 4; k++) Rr += Ls.tab[(pt * 4 + k) * 3 + 2] * gmem(a.x)[3 * (long long)nd[k]];
  const double detJ = (J0 * J3) - (J1 * J2);
  if (detJ < 0. && valid) raise_status(st, ST_DETJ, e);
  const double inv[4] = {J3 / detJ, -J1 / detJ, -J2 / detJ, J0 / detJ};

  // ElementAx::getF_atIntPoint, FE/ElementAx.C:110-136
  T9 F = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
  double vr = 0.;
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    const auto *uk = gmem(a.du) + 3 * (long long)nd[k];
    double u0 = uk[0], u1 = uk[1];
    if (SPEED_BC)
    {
      const unsigned char m = gmem(a.bcs)[nd[k]];
      if (m)
      {
        const auto *sp = gmem(a.bcs_v) + 3 * (long long)nd[k];
        if (m & 1) u0 = sp[0] * st->dt;
        if (m & 2) u1 = sp[1] * st->dt;
      }
    }
    const double *tg = Ls.tab + (pt * 4 + k) * 3;
    const double g[2] = {tg[0], tg[1]};
    double d[2];
    dgemm_row<2>(g, inv, d);
    Ls.x[2 * k][tid] = d[0];
    Ls.x[2 * k + 1][tid] = d[1];
    F.v[0] += d[0] * u0; F.v[1] += d[1] * u0;
    F.v[3] += d[0] * u1; F.v[4] += d[1] * u1;
    vr += tg[2] * u0;
  }
  F.v[8] += vr / Rr;

  T9 dEps, R;
  if (!polar_lnU(F, dEps, R) && valid) raise_status(st, ST_POLAR, e);
  T9 S;
#pragma unroll
  for (int c = 0; c < 9; c++) S.v[c] = gmem(a.S)[c * nipt + ip];
  double evp = gmem(a.evp)[ip];
  T9 eps, epl;
  if (TRACK)
  {
#pragma unroll
    for (int c = 0; c < 9; c++) { eps.v[c] = gmem(a.eps)[c * nipt + ip] + dEps.v[c]; epl.v[c] = gmem(a.epsplas)[c * nipt + ip]; }
  }
  const double tr = dEps.v[0] + dEps.v[4] + dEps.v[8];
  double Tip = 0.0;
  bool rr_ok;
  if (JC)  // law 2, a separate instance: its registers do not weigh on law 1
  {
    Tip = gmem(a.T)[ip];
    rr_ok = stress_update_jc(a, dEps, S, evp, Tip, st->dt, TRACK ? &epl : nullptr);
  }
  else
    rr_ok = stress_update(a, dEps, S, evp, TRACK ? &epl : nullptr);
  if (!rr_ok && valid) raise_status(st, ST_RR, e);
  double pressure = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++) pressure += __shfl(tr, base + q, 64);
  pressure /= 4;
  const double p = gmem(a.p)[e] + a.K * pressure;
  S = mul(R, mul_bt(S, R));
  if (TRACK) { eps = mul(R, mul_bt(eps, R)); epl = mul(R, mul_bt(epl, R)); }
  T9 Sig;
#pragma unroll
  for (int k = 0; k < 9; k++) Sig.v[k] = S.v[k] + p * ((k % 4 == 0) ? 1.0 : 0.0);
  double WxdJ = 1.0 * detJ;
  WxdJ *= kDeuxPI * Rr;

  if (valid)
  {
#pragma unroll
    for (int c = 0; c < 9; c++) gmem(a.S)[c * nipt + ip] = S.v[c];
    gmem(a.evp)[ip] = evp;
    if (JC) gmem(a.T)[ip] = Tip;
    if (pt == 0) gmem(a.p)[e] = p;
    if (TRACK)
#pragma unroll
      for (int c = 0; c < 9; c++)
      {
        gmem(a.eps)[c * nipt + ip] = eps.v[c];
        gmem(a.epsplas)[c * nipt + ip] = epl.v[c];
        gmem(a.deps)[c * nipt + ip] = dEps.v[c];
      }
  }

  Ls.x[8][tid] = Sig.v[0]; Ls.x[9][tid] = Sig.v[1]; Ls.x[10][tid] = Sig.v[3]; Ls.x[11][tid] = Sig.v[4];
  Ls.x[12][tid] = Sig.v[8]; Ls.x[13][tid] = WxdJ; Ls.x[14][tid] = Rr;
  __syncthreads();

  // FE/Element.C:178-212: in-plane terms, then the hoop term of this point
  const int L = pt;
  double f0 = 0.0, f1 = 0.0;
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    const int src = lbase + q;
    const double s00 = Ls.x[8][src], s01 = Ls.x[9][src], s10 = Ls.x[10][src], s11 = Ls.x[11][src];
    const double s22 = Ls.x[12][src], W = Ls.x[13][src], Rq = Ls.x[14][src];
    const double d[2] = {Ls.x[2 * L][src], Ls.x[2 * L + 1][src]};
    f0 -= d[0] * s00 * W;
    f0 -= d[1] * s10 * W;
    f1 -= d[0] * s01 * W;
    f1 -= d[1] * s11 * W;
    f0 -= (Ls.tab[(q * 4 + L) * 3 + 2] * s22 / Rq) * W;
  }
  if (valid)
  {
    auto *fo = gmem(a.fe) + (long long)gmem(a.fpos)[e * 4 + L] * 2;
    fo[0] = f0; fo[1] = f1;
  }
}

__device__ __forceinline__ void apply_constant(unsigned char b, unsigned char s, const double *sp, double dt,
                                               double *du, double *v, double *a, double *ac)
{
  // BoundarySpeed::applyConstant (constant speed), FE/Boundary.C:478-500
#pragma unroll
  for (int c = 0; c < 3; c++)
    if (s & (1 << c)) { a[c] = 0.; ac[c] = 0.; v[c] = sp[c]; du[c] = v[c] * dt; }
  // BoundaryRestrain::applyConstant, FE/Boundary.C:636-666
#pragma unroll
  for (int c = 0; c < 3; c++)
    if (b & (1 << c)) { du[c] = 0.; v[c] = 0.; a[c] = 0.; ac[c] = 0.; }
}

__device__ __forceinline__ double hex_length(const ElemArgs &a, const int *nd)
{
#pragma clang fp contract(off)
  // build-defined hex8 length: V / A_max (V = sum of the 8 detJ, weights 1)
  double V = 0.0;
  for (int pt = 0; pt < 8; pt++)
  {
    T9 J = {{0, 0, 0, 0, 0, 0, 0, 0, 0}};
#pragma unroll
    for (int k = 0; k < 8; k++)
    {
      const double *xk = a.x + 3 * (long long)nd[k];
      double g[3];
      hex_dn0(a, pt, k, g);
      J.v[0] += g[0] * xk[0]; J.v[1] += g[1] * xk[0]; J.v[2] += g[2] * xk[0];
      J.v[3] += g[0] * xk[1]; J.v[4] += g[1] * xk[1]; J.v[5] += g[2] * xk[1];
      J.v[6] += g[0] * xk[2]; J.v[7] += g[1] * xk[2]; J.v[8] += g[2] * xk[2];
    }
    V += (J.v[0] * J.v[4] * J.v[8]) + (J.v[3] * J.v[7] * J.v[2]) + (J.v[6] * J.v[1] * J.v[5]) -
         (J.v[6] * J.v[4] * J.v[2]) - (J.v[0] * J.v[7] * J.v[5]) - (J.v[3] * J.v[1] * J.v[8]);
  }
  const int faces[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {7, 4, 5, 6}, {2, 1, 0, 3}};
  double Amax = 0.0;
#pragma unroll
  for (int f = 0; f < 6; f++)
  {
    const double *xa = a.x + 3 * (long long)nd[faces[f][0]], *xb = a.x + 3 * (long long)nd[faces[f][1]],
                 *xc = a.x + 3 * (long long)nd[faces[f][2]], *xd = a.x + 3 * (long long)nd[faces[f][3]];
    const double d1[3] = {xc[0] - xa[0], xc[1] - xa[1], xc[2] - xa[2]};
    const double d2[3] = {xd[0] - xb[0], xd[1] - xb[1], xd[2] - xb[2]};
    const double c0 = d1[1] * d2[2] - d1[2] * d2[1];
    const double c1 = d1[2] * d2[0] - d1[0] * d2[2];
    const double c2 = d1[0] * d2[1] - d1[1] * d2[0];
    const double area = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (area > Amax) Amax = area;
  }
  return V / Amax;
}

// One node of the fused nodal step: ExplicitSolverCH::computePredictions
// (FE/ExplicitSolverCH.C:408-424) with its BCs, the Fint assembly in element
// order and acc = Fint / M (FE/ExplicitSolverCH.C:506-519), the corrector
// (FE/ExplicitSolverCH.C:522-541) and the BCs again.  ac, vc: Current
// acceleration and velocity in; an, vn, dun: New fields out.
template <int DIM, bool HAS_BCS, bool FE32>
__device__ __forceinline__ void nodal_core(const NodalArgs &n, long long i, const StepState *st, double *ac,
                                           const double *vc, double *an, double *vn, double *dun, double &m)
{
#pragma clang fp contract(off)
  const double dt = st->dt, beta = st->beta, gam = st->gamma, am = st->alphaM;
  const double s1 = dt * (0.5 - beta), s2 = dt * (1 - gam), s3 = (dt * gam), s4 = ((dt * dt) * beta);
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
    double u = ac[c];
    u *= s1;
    u += vc[c];
    u *= dt;
    dun[c] = u;
    double w = ac[c];
    w *= s2;
    w += vc[c];
    vn[c] = w;
    an[c] = 0.0;
  }
  const unsigned char b = n.bc[i];
  const unsigned char s = HAS_BCS ? n.bcs[i] : 0;
  const double *sp = HAS_BCS ? n.bcs_v + 3 * i : nullptr;
  if (b | s) apply_constant(b, s, sp, dt, dun, vn, an, ac);
  double f[DIM];
#pragma unroll
  for (int j = 0; j < DIM; j++) f[j] = 0.0;
  const int dg = n.deg[i];
  for (int k = 0; k < dg; k++)
  {
    const long long q = ((long long)k * n.nn + i) * DIM;
    if (FE32)
    {
#pragma unroll
      for (int j = 0; j < DIM; j++) f[j] += (double)n.fe32[q + j];
    }
    else
    {
#pragma unroll
      for (int j = 0; j < DIM; j++) f[j] += n.fe[q + j];
    }
  }
  if (n.fint)
#pragma unroll
    for (int j = 0; j < DIM; j++) n.fint[i * DIM + j] = f[j];
  m = n.mass[i];
#pragma unroll
  for (int j = 0; j < DIM; j++) an[j] = f[j] / m;
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
    double aa = an[c];
    aa -= am * ac[c];
    aa /= (1.0 - am);
    an[c] = aa;
    vn[c] += s3 * aa;
    dun[c] += s4 * aa;
  }
  if (b | s) apply_constant(b, s, sp, dt, dun, vn, an, ac);
}

// ElQua4nAx::getLength, FE/ElQua4nAx.C:131-143
template <class XP, class CP>
__device__ __forceinline__ double quad_length(XP x, CP cn)
{
#pragma clang fp contract(off)
  const auto x0 = x + 3 * (long long)cn[0], x1 = x + 3 * (long long)cn[1], x2 = x + 3 * (long long)cn[2],
             x3 = x + 3 * (long long)cn[3];
  const double x31 = x2[0] - x0[0];
  const double x24 = x1[0] - x3[0];
  const double y42 = x3[1] - x1[1];
  const double y31 = x2[1] - x0[1];
  return ((x31 * y42 + x24 * y31) / sqrt(2. * ((x24 * x24) + (y42 * y42) + (x31 * x31) + (y31 * y31))));
}

// Solver::computeTimeStep (FE/Solver.C:220-267) given the Courant minimum
// (or, for the Pulse method, the frequency left in pulse_f) when it
// recomputes; ExplicitSolverCH::initStep (FE/ExplicitSolverCH.C:434-447)
__device__ __forceinline__ void compute_time_step(StepState *st, int recompute, double crit)
{
#pragma clang fp contract(off)
  if (recompute)
  {
    const double fmax = st->ts_method == 1 ? st->pulse_f : 2.0 / crit;
    st->dt = st->safety * st->omegaS / fmax;
    st->lastTsDate = st->tsDate;
    st->lastTimeStep = st->computedTimeStep;
    st->tsDate = st->t;
    st->computedTimeStep = st->dt;
  }
  else if (st->tsDate != st->lastTsDate)
  {
    const double dx = (st->t - st->lastTsDate) / (st->tsDate - st->lastTsDate);
    st->dt = ((dx) * (st->computedTimeStep) + (1.0 - (dx)) * (st->lastTimeStep));
  }
}
__device__ __forceinline__ void init_step(StepState *st)
{
#pragma clang fp contract(off)
  if (st->t + st->dt > st->upTime) st->dt = st->upTime - st->t;
  st->increment++;
}
// ExplicitSolverCH::endStep time update and Solver::betweenTimeBounds (FE/Solver.C:169-176)
__device__ __forceinline__ int end_step(StepState *st)
{
#pragma clang fp contract(off)
  st->t += st->dt;
  const double lim = st->endTime < st->upTime ? st->endTime : st->upTime;
  st->running = (st->t >= 0.0) && (st->t < lim);
  return st->running;
}

}  // namespace dnl
