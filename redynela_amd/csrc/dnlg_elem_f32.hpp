// Mixed-precision form of the element pass: element kinematics, the
// polar decomposition / Hencky increment, the J2 radial return and B^T sigma
// in FP32; nodal coordinates, velocities, increments, forces after assembly
// and time in FP64 (the nodal kernel of the FP64 form is reused).
//
// This form is a deliberate departure from the reference arithmetic
// (FE/Element.C:164-403, LA/Tensor2.C:1154-1259):
//   * the Hencky increment is taken from C - I = H + H^T + H^T H with
//     H = grad(du), by cyclic Jacobi in FP32 and log1p of the eigenvalues,
//     because C = F^T F ~ I + 2e loses 3-4 digits in FP32 and the reference's
//     1e-9 Jacobi stop is below FP32 resolution;
//   * the Jacobian is formed from node coordinates relative to node 0
//     (rounded from FP64), which keeps FP32 precision for small elements;
//   * the deviatoric stress is stored symmetric (6 components);
//   * B^T sigma is split as B^T S_rot + p_new B^T 1 so that one lane carries a
//     whole element through its 8 points in one pass;
//   * the radial return stops at |residual / sigma_y| <= 1e-6.
// One element per lane.  Translation units including this header are
// compiled with FMA contraction and fast FP32 division / square root.
#pragma once
#include "dnlg_internal.hpp"

#include <hip/hip_runtime.h>

namespace dnl {

__device__ __forceinline__ int s0(int k) { int q = k & 3; return (q == 1 || q == 2) ? 1 : -1; }
__device__ __forceinline__ int s1(int k) { return ((k & 3) >= 2) ? 1 : -1; }
__device__ __forceinline__ int s2(int k) { return (k >= 4) ? 1 : -1; }

// trilinear shape-function derivatives dN0(ip, k, d) of hex8 at its eight
// 2-point Gauss points, a compile-time table in constant memory: the point
// loop indexes it with a wave-uniform point, so the values arrive by scalar
// loads instead of being formed per lane
struct HexTab32
{
  float v[8 * 8 * 3];
};
__host__ __device__ constexpr HexTab32 make_hex_tab32()
{
  HexTab32 t{};
  const float r3 = 0.57735026918962576f;  // 1/sqrt(3)
  for (int ip = 0; ip < 8; ip++)
    for (int k = 0; k < 8; k++)
    {
      const int q = k & 3, qi = ip & 3;
      const float sk0 = (q == 1 || q == 2) ? 1.f : -1.f, sk1 = q >= 2 ? 1.f : -1.f, sk2 = k >= 4 ? 1.f : -1.f;
      const float si0 = (qi == 1 || qi == 2) ? 1.f : -1.f, si1 = qi >= 2 ? 1.f : -1.f, si2 = ip >= 4 ? 1.f : -1.f;
      const float a0 = 1.f + sk0 * si0 * r3, a1 = 1.f + sk1 * si1 * r3, a2 = 1.f + sk2 * si2 * r3;
      t.v[(ip * 8 + k) * 3 + 0] = 0.125f * sk0 * a1 * a2;
      t.v[(ip * 8 + k) * 3 + 1] = 0.125f * sk1 * a0 * a2;
      t.v[(ip * 8 + k) * 3 + 2] = 0.125f * sk2 * a0 * a1;
    }
  return t;
}
static __constant__ HexTab32 kHexTab32 = make_hex_tab32();

// bilinear shape functions and derivatives of quad4 at its 2-point Gauss points
__device__ __forceinline__ void quad_dn0f(int ip, int k, float *g, float *N)
{
  const float r3 = 0.57735026918962576f;
  const float a0 = 1.f + s0(k) * s0(ip) * r3, a1 = 1.f + s1(k) * s1(ip) * r3;
  g[0] = 0.25f * s0(k) * a1;
  g[1] = 0.25f * s1(k) * a0;
  *N = 0.25f * a0 * a1;
}

// symmetric 3x3 eigen-decomposition by cyclic Jacobi (FP32).  a: xx,yy,zz,xy,yz,xz.
__device__ __forceinline__ void jacobi_sym(float *a, float *V)
{
  float A00 = a[0], A11 = a[1], A22 = a[2], A01 = a[3], A12 = a[4], A02 = a[5];
#pragma unroll
  for (int k = 0; k < 9; k++) V[k] = (k % 4 == 0) ? 1.f : 0.f;
  for (int sweep = 0; sweep < 4; sweep++)
  {
    const float off = A01 * A01 + A12 * A12 + A02 * A02;
    const float dg = A00 * A00 + A11 * A11 + A22 * A22;
    if (off <= 1e-14f * dg || off < 1e-36f) break;
#pragma unroll
    for (int pq = 0; pq < 3; pq++)
    {
      // pairs (0,1), (0,2), (1,2); r is the third index
      float app, aqq, apq, arp, arq;
      if (pq == 0) { app = A00; aqq = A11; apq = A01; arp = A02; arq = A12; }
      else if (pq == 1) { app = A00; aqq = A22; apq = A02; arp = A01; arq = A12; }
      else { app = A11; aqq = A22; apq = A12; arp = A01; arq = A02; }
      if (fabsf(apq) > 1e-30f)
      {
        const float th = (aqq - app) / (2.f * apq);
        const float t = copysignf(1.f, th) / (fabsf(th) + sqrtf(th * th + 1.f));
        const float c = rsqrtf(t * t + 1.f), s = t * c;
        app -= t * apq;
        aqq += t * apq;
        const float nrp = c * arp - s * arq, nrq = s * arp + c * arq;
        const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
#pragma unroll
        for (int i = 0; i < 3; i++)
        {
          const float vp = V[3 * i + p], vq = V[3 * i + q];
          V[3 * i + p] = c * vp - s * vq;
          V[3 * i + q] = s * vp + c * vq;
        }
        if (pq == 0) { A00 = app; A11 = aqq; A01 = 0.f; A02 = nrp; A12 = nrq; }
        else if (pq == 1) { A00 = app; A22 = aqq; A02 = 0.f; A01 = nrp; A12 = nrq; }
        else { A11 = app; A22 = aqq; A12 = 0.f; A01 = nrp; A02 = nrq; }
      }
    }
  }
  a[0] = A00; a[1] = A11; a[2] = A22;
}

// symmetric (6: xx, yy, zz, xy, yz, xz) product of two commuting symmetric
// matrices, P A with P a polynomial in A
__device__ __forceinline__ void sym_mul(const float *P, const float *A, float *M)
{
  M[0] = P[0] * A[0] + P[3] * A[3] + P[5] * A[5];
  M[1] = P[3] * A[3] + P[1] * A[1] + P[4] * A[4];
  M[2] = P[5] * A[5] + P[4] * A[4] + P[2] * A[2];
  M[3] = P[0] * A[3] + P[3] * A[1] + P[5] * A[4];
  M[4] = P[3] * A[5] + P[1] * A[4] + P[4] * A[2];
  M[5] = P[0] * A[5] + P[3] * A[4] + P[5] * A[2];
}

// f(A) = sum_k c_k A^k for symmetric A, by Horner: ((c_n A + c_{n-1} I) A + ...) A + c_0 I
template <int N>
__device__ __forceinline__ void sym_poly(const float *A, const float (&c)[N], float *out)
{
  float P[6] = {c[N - 1], c[N - 1], c[N - 1], 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = N - 2; k >= 0; k--)
  {
    float M[6];
    sym_mul(P, A, M);
    P[0] = M[0] + c[k]; P[1] = M[1] + c[k]; P[2] = M[2] + c[k];
    P[3] = M[3]; P[4] = M[4]; P[5] = M[5];
  }
#pragma unroll
  for (int k = 0; k < 6; k++) out[k] = P[k];
}

// Hencky increment and rotation from H = grad(du): dEps = 1/2 ln(I + E2),
// R = F (I + E2)^(-1/2), E2 = H + H^T + H^T H.  dEps symmetric (6).
// Small increments take the Taylor series of both matrix functions: to
// degree 4 when every entry of E2 is below 2^-8 (spectral radius x below
// 3/256: remainders x^5/10 and 0.25 x^5 below 6e-11, about x^4 relative to
// dEps, under FP32 resolution), to degree 8 below 1/32 (x below 3/32,
// remainders below 1.1e-10); larger increments take the Jacobi
// eigen-decomposition.
template <int N>
__device__ __forceinline__ void polar_series_f32(const float *H, const float *E2, float *dE, float *R)
{
  // 1/2 ln(1 + x) = sum_k (-1)^(k+1) x^k / (2k); (1 + x)^(-1/2) = sum_k binom(-1/2, k) x^k
  const float cl9[9] = {0.f, 0.5f, -0.25f, 1.f / 6.f, -0.125f, 0.1f, -1.f / 12.f, 1.f / 14.f, -0.0625f};
  const float cu9[9] = {1.f, -0.5f, 0.375f, -0.3125f, 0.2734375f, -0.24609375f, 0.2255859375f,
                        -0.20947265625f, 0.196380615234375f};
  float cl[N], cu[N];
#pragma unroll
  for (int k = 0; k < N; k++) { cl[k] = cl9[k]; cu[k] = cu9[k]; }
  float Ui6[6];
  sym_poly(E2, cl, dE);
  sym_poly(E2, cu, Ui6);
  const float Ui[9] = {Ui6[0], Ui6[3], Ui6[5], Ui6[3], Ui6[1], Ui6[4], Ui6[5], Ui6[4], Ui6[2]};
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      float r = Ui[3 * i + j];
#pragma unroll
      for (int k = 0; k < 3; k++) r += H[3 * i + k] * Ui[3 * k + j];
      R[3 * i + j] = r;
    }
}

__device__ __forceinline__ void polar_f32(const float *H, float *dE, float *R)
{
  float E2[6];
  // (H^T H)_ij = sum_k H_ki H_kj
  E2[0] = 2.f * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
  E2[1] = 2.f * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
  E2[2] = 2.f * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  E2[3] = H[1] + H[3] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
  E2[4] = H[5] + H[7] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
  E2[5] = H[2] + H[6] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
  float amax = 0.f;
#pragma unroll
  for (int k = 0; k < 6; k++) amax = fmaxf(amax, fabsf(E2[k]));
  if (amax < 0.00390625f)
  {
    polar_series_f32<5>(H, E2, dE, R);
    return;
  }
  if (amax < 0.03125f)
  {
    polar_series_f32<9>(H, E2, dE, R);
    return;
  }
  float V[9];
  jacobi_sym(E2, V);
  float l[3], ui[3];
#pragma unroll
  for (int i = 0; i < 3; i++)
  {
    l[i] = 0.5f * log1pf(E2[i]);
    ui[i] = rsqrtf(1.f + E2[i]);
  }
  // dE = V diag(l) V^T ; Ui = V diag(ui) V^T
  float Ui[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i; j < 3; j++)
    {
      float se = 0.f, su = 0.f;
#pragma unroll
      for (int k = 0; k < 3; k++)
      {
        const float w = V[3 * i + k] * V[3 * j + k];
        se += l[k] * w;
        su += ui[k] * w;
      }
      Ui[3 * i + j] = su;
      Ui[3 * j + i] = su;
      if (i == j) dE[i] = se;
      else if (i == 0 && j == 1) dE[3] = se;
      else if (i == 1 && j == 2) dE[4] = se;
      else dE[5] = se;
    }
  // R = F Ui, F = I + H
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      float r = Ui[3 * i + j];
#pragma unroll
      for (int k = 0; k < 3; k++) r += H[3 * i + k] * Ui[3 * k + j];
      R[3 * i + j] = r;
    }
}

// yield stress and its derivative with respect to gamma (sigma_y(evp + sqrt(2/3) gamma))
__device__ __forceinline__ void yield_f32(const ElemArgs32 &a, float evp, float gamma, float rate_dt_inv, float theta,
                                          float &sy, float &dsy)
{
  const float k23 = 0.81649658092772603f;
  const float ev = evp + k23 * gamma;
  const float pw = (a.n == 1.f) ? ev : __powf(ev, a.n);
  const float base = a.A + a.B * pw;
  const float dbase = (a.n == 1.f) ? a.B : a.n * a.B * pw / ev;
  if (a.law == 2)
  {
    // rate factor 1 + C ln(epsdot / epsdot0) when epsdot >= epsdot0, else 1
    const float epsd = k23 * gamma * rate_dt_inv;
    float rf = 1.f, drf = 0.f;
    if (a.C != 0.f && epsd >= a.epsd0 && gamma > 0.f)
    {
      rf = 1.f + a.C * __logf(epsd / a.epsd0);
      drf = a.C / gamma;  // d rf / d gamma
    }
    sy = base * rf * theta;
    dsy = (k23 * dbase * rf + base * drf) * theta;
  }
  else
  {
    sy = base;
    dsy = k23 * dbase;
  }
}

__device__ __forceinline__ float thermal_factor(const ElemArgs32 &a, float T)
{
  if (a.law != 2 || a.chi_rho_cp == 0.f) return 1.f;
  if (T < a.T0) return 1.f;
  if (T >= a.Tm) return 0.f;
  return 1.f - __powf((T - a.T0) / (a.Tm - a.T0), a.m);
}

// elastic predictor + radial return on the symmetric deviator (6)
__device__ __forceinline__ bool stress_f32(const ElemArgs32 &a, const float *dE, float *S, float &evp, float &T,
                                           float dt_inv)
{
  const float tr3 = (dE[0] + dE[1] + dE[2]) * (1.f / 3.f);
  const float g2 = 2.f * a.G;
  S[0] += g2 * (dE[0] - tr3);
  S[1] += g2 * (dE[1] - tr3);
  S[2] += g2 * (dE[2] - tr3);
  S[3] += g2 * dE[3];
  S[4] += g2 * dE[4];
  S[5] += g2 * dE[5];
  if (a.law == 0) return true;
  const float nrm = sqrtf(S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2.f * (S[3] * S[3] + S[4] * S[4] + S[5] * S[5]));
  const float k32 = 1.2247448713915890f, k23 = 0.81649658092772603f;
  const float Strial = k32 * nrm;
  const float theta = thermal_factor(a, T);
  float sy0, dsy0;
  yield_f32(a, evp, 0.f, dt_inv, theta, sy0, dsy0);
  if (!(Strial > sy0)) return true;
  // Newton on the consistency condition Strial - sqrt(3/2) 2G gamma = sigma_y(gamma)
  const float h0 = (a.n == 1.f) ? a.B : a.n * a.B * __powf(evp, a.n - 1.f);
  float gamma = (nrm - k23 * sy0) / (2.f * a.G * (1.f + h0 / (3.f * a.G)));
  float lo = 0.f, hi = Strial / (k32 * 2.f * a.G);
  if (!(gamma > lo && gamma < hi)) gamma = 0.5f * (lo + hi);
  int it = 0;
  for (; it < 100; it++)
  {
    float sy, dsy;
    yield_f32(a, evp, gamma, dt_inv, theta, sy, dsy);
    const float res = Strial - k32 * 2.f * a.G * gamma - sy;
    if (fabsf(res / sy) <= 1e-6f) break;
    if (res > 0.f) lo = gamma; else hi = gamma;
    const float gn = gamma + res / (k32 * 2.f * a.G + dsy);
    gamma = (gn > lo && gn < hi) ? gn : 0.5f * (lo + hi);
  }
  if (it == 100) return false;
  const float f = 1.f - g2 * gamma / nrm;
#pragma unroll
  for (int k = 0; k < 6; k++) S[k] *= f;
  const float dev = k23 * gamma;
  if (a.law == 2 && a.chi_rho_cp != 0.f)
  {
    float sy, dsy;
    yield_f32(a, evp, gamma, dt_inv, theta, sy, dsy);
    T += a.chi_rho_cp * sy * dev;   // adiabatic heating, Taylor-Quinney fraction chi
  }
  evp += dev;
  return true;
}

// S_rot = R S R^T for symmetric S (6) -> symmetric (6)
__device__ __forceinline__ void rot_sym(const float *R, float *S)
{
  const float M[9] = {S[0], S[3], S[5], S[3], S[1], S[4], S[5], S[4], S[2]};
  float RM[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) RM[3 * i + j] = R[3 * i] * M[j] + R[3 * i + 1] * M[3 + j] + R[3 * i + 2] * M[6 + j];
  auto e = [&](int i, int j) { return RM[3 * i] * R[3 * j] + RM[3 * i + 1] * R[3 * j + 1] + RM[3 * i + 2] * R[3 * j + 2]; };
  S[0] = e(0, 0); S[1] = e(1, 1); S[2] = e(2, 2); S[3] = e(0, 1); S[4] = e(1, 2); S[5] = e(0, 2);
}

__device__ __forceinline__ void raise32(StepState *st, int code, long long e)
{
  if (atomicCAS(&st->status, 0, code) == 0) st->err_elem = (int)e;
}

// ---------------------------------------------------------------------------
// hex8, one element per lane.  The element's node rows (coordinates relative
// to node 0 and the Current increment, 48 floats) sit in the lane's own LDS
// column lx[r * NT] instead of registers and are read again at every Gauss
// point: 199 -> 155 VGPRs, 2 -> 3 waves per SIMD.  No other lane reads the
// column, so no barrier is needed.
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void elem_hex8_f32_body(const ElemArgs32 &a, long long e, float *lx)
{
  StepState *st = a.st;
  const long long ne = a.ne;
  const float dt_inv = (a.law == 2) ? (float)(1.0 / st->dt) : 0.f;
  int nd[8];
#pragma unroll
  for (int k = 0; k < 8; k++) nd[k] = a.conn[e * 8 + k];
  const double x00 = a.x[3 * (long long)nd[0]], x01 = a.x[3 * (long long)nd[0] + 1], x02 = a.x[3 * (long long)nd[0] + 2];
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    const double *xk = a.x + 3 * (long long)nd[k];
    lx[(6 * k + 0) * NT] = (float)(xk[0] - x00);
    lx[(6 * k + 1) * NT] = (float)(xk[1] - x01);
    lx[(6 * k + 2) * NT] = (float)(xk[2] - x02);
    const double *uk = a.du + 3 * (long long)nd[k];
    lx[(6 * k + 3) * NT] = (float)uk[0];
    lx[(6 * k + 4) * NT] = (float)uk[1];
    lx[(6 * k + 5) * NT] = (float)uk[2];
  }
  float Af[8][3], g[8][3];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int c = 0; c < 3; c++) { Af[k][c] = 0.f; g[k][c] = 0.f; }
  float pressure = 0.f;
  bool ok = true;
#pragma unroll 1
  for (int pt = 0; pt < 8; pt++)
  {
    // a compiler barrier: without it the node rows are hoisted out of the
    // loop into registers again
    __asm__ volatile("" ::: "memory");
    // dN0 of the 8 nodes at this point, then (in place) dN = dN0 J^-1 for H and B^T sigma
    float dn[8][3];
#pragma unroll
    for (int k = 0; k < 8; k++)
    {
      dn[k][0] = kHexTab32.v[(pt * 8 + k) * 3 + 0];
      dn[k][1] = kHexTab32.v[(pt * 8 + k) * 3 + 1];
      dn[k][2] = kHexTab32.v[(pt * 8 + k) * 3 + 2];
    }
    float J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 1; k < 8; k++)
    {
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) J[3 * i + j] += dn[k][j] * lx[(6 * k + i) * NT];
    }
    const float c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
    const float detJ = J[0] * c00 + J[1] * c01 + J[2] * c02;
    const float id = 1.f / detJ;
    float inv[9];
    inv[0] = c00 * id; inv[1] = (J[2] * J[7] - J[1] * J[8]) * id; inv[2] = (J[1] * J[5] - J[2] * J[4]) * id;
    inv[3] = c01 * id; inv[4] = (J[0] * J[8] - J[2] * J[6]) * id; inv[5] = (J[2] * J[3] - J[0] * J[5]) * id;
    inv[6] = c02 * id; inv[7] = (J[1] * J[6] - J[0] * J[7]) * id; inv[8] = (J[0] * J[4] - J[1] * J[3]) * id;
    // H = sum_k u_k (x) dN_k
    float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; k++)
    {
      float d[3];
#pragma unroll
      for (int j = 0; j < 3; j++) d[j] = dn[k][0] * inv[j] + dn[k][1] * inv[3 + j] + dn[k][2] * inv[6 + j];
#pragma unroll
      for (int j = 0; j < 3; j++) dn[k][j] = d[j];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) H[3 * i + j] += lx[(6 * k + 3 + i) * NT] * d[j];
    }
    float dE[6], R[9];
    polar_f32(H, dE, R);
    const long long ip = (long long)pt * ne + e;
    const long long nipt = 8 * ne;
    float S[6];
#pragma unroll
    for (int c = 0; c < 6; c++) S[c] = a.S[c * nipt + ip];
    float evp = a.evp[ip];
    float T = a.T ? a.T[ip] : 0.f;
    ok = stress_f32(a, dE, S, evp, T, dt_inv) && ok;
    pressure += dE[0] + dE[1] + dE[2];
    rot_sym(R, S);
#pragma unroll
    for (int c = 0; c < 6; c++) a.S[c * nipt + ip] = S[c];
    a.evp[ip] = evp;
    if (a.T) a.T[ip] = T;
    const float W = detJ;
    const float M[9] = {S[0], S[3], S[5], S[3], S[1], S[4], S[5], S[4], S[2]};
#pragma unroll
    for (int k = 0; k < 8; k++)
    {
      float d[3];
#pragma unroll
      for (int j = 0; j < 3; j++) d[j] = dn[k][j] * W;
#pragma unroll
      for (int i = 0; i < 3; i++)
      {
        Af[k][i] += d[0] * M[i] + d[1] * M[3 + i] + d[2] * M[6 + i];
        g[k][i] += d[i];
      }
    }
  }
  if (!ok) raise32(st, ST_RR, e);
  const float p = a.p[e] + a.K * (pressure * 0.125f);
  a.p[e] = p;
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    float *fo = a.fe + (long long)a.fpos[e * 8 + k] * 3;
#pragma unroll
    for (int i = 0; i < 3; i++) fo[i] = -(Af[k][i] + p * g[k][i]);
  }
}

// ---------------------------------------------------------------------------
// axisymmetric quad4, one element per lane
// ---------------------------------------------------------------------------
// NT = 0: the element's node rows (coordinates relative to node 0, radius,
// Current increment: 20 floats) in registers; NT > 0: in the lane's own LDS
// column lx[i * NT], read again at every Gauss point (as the hex8 body does)
template <int NT = 0>
__device__ __forceinline__ void elem_quad4ax_f32_body(const ElemArgs32 &a, long long e, float *lx = nullptr)
{
  StepState *st = a.st;
  const long long ne = a.ne;
  const float dt_inv = (a.law == 2) ? (float)(1.0 / st->dt) : 0.f;
  int nd[4];
#pragma unroll
  for (int k = 0; k < 4; k++) nd[k] = gmem(a.conn)[e * 4 + k];
  float rreg[NT == 0 ? 20 : 1];
  // row 5k + c: x_c - x_c(node 0); 5k + 2: radius; 5k + 3 + c: increment
  auto row = [&](int i) -> float & {
    if constexpr (NT == 0) return rreg[i];
    else return lx[i * NT];
  };
  const double x00 = gmem(a.x)[3 * (long long)nd[0]], x01 = gmem(a.x)[3 * (long long)nd[0] + 1];
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    const auto *xk = gmem(a.x) + 3 * (long long)nd[k];
    row(5 * k + 0) = (float)(xk[0] - x00);
    row(5 * k + 1) = (float)(xk[1] - x01);
    row(5 * k + 2) = (float)xk[0];
    row(5 * k + 3) = (float)gmem(a.du)[3 * (long long)nd[k]];
    row(5 * k + 4) = (float)gmem(a.du)[3 * (long long)nd[k] + 1];
  }
  float Af[4][2], g[4][2];
#pragma unroll
  for (int k = 0; k < 4; k++) { Af[k][0] = Af[k][1] = g[k][0] = g[k][1] = 0.f; }
  float pressure = 0.f;
  bool ok = true;
  const float twopi = 6.2831853071795865f;
#pragma unroll
  for (int pt = 0; pt < 4; pt++)
  {
    if constexpr (NT > 0) __asm__ volatile("" ::: "memory");  // keep the LDS rows out of registers
    float J0 = 0, J1 = 0, J2 = 0, J3 = 0, Rr = 0.f, ur = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
      float d0[2], N;
      quad_dn0f(pt, k, d0, &N);
      const float x0 = row(5 * k), x1 = row(5 * k + 1);
      J0 += d0[0] * x0; J1 += d0[1] * x0;
      J2 += d0[0] * x1; J3 += d0[1] * x1;
      Rr += N * row(5 * k + 2);
      ur += N * row(5 * k + 3);
    }
    const float detJ = J0 * J3 - J1 * J2;
    const float id = 1.f / detJ;
    const float inv[4] = {J3 * id, -J1 * id, -J2 * id, J0 * id};
    float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
      float d0[2], N;
      quad_dn0f(pt, k, d0, &N);
      const float d[2] = {d0[0] * inv[0] + d0[1] * inv[2], d0[0] * inv[1] + d0[1] * inv[3]};
      const float u0 = row(5 * k + 3), u1 = row(5 * k + 4);
      H[0] += u0 * d[0]; H[1] += u0 * d[1];
      H[3] += u1 * d[0]; H[4] += u1 * d[1];
    }
    H[8] = ur / Rr;
    float dE[6], R[9];
    polar_f32(H, dE, R);
    const long long ip = (long long)pt * ne + e;
    const long long nipt = 4 * ne;
    float S[6];
#pragma unroll
    for (int c = 0; c < 6; c++) S[c] = gmem(a.S)[c * nipt + ip];
    float evp = gmem(a.evp)[ip];
    float T = a.T ? gmem(a.T)[ip] : 0.f;
    ok = stress_f32(a, dE, S, evp, T, dt_inv) && ok;
    pressure += dE[0] + dE[1] + dE[2];
    rot_sym(R, S);
#pragma unroll
    for (int c = 0; c < 6; c++) gmem(a.S)[c * nipt + ip] = S[c];
    gmem(a.evp)[ip] = evp;
    if (a.T) gmem(a.T)[ip] = T;
    const float W = detJ * twopi * Rr;
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
      float d0[2], N;
      quad_dn0f(pt, k, d0, &N);
      const float d[2] = {(d0[0] * inv[0] + d0[1] * inv[2]) * W, (d0[0] * inv[1] + d0[1] * inv[3]) * W};
      const float hoop = N * W / Rr;
      // in-plane B^T S plus the hoop term N S_zz / R
      Af[k][0] += d[0] * S[0] + d[1] * S[3] + hoop * S[2];
      Af[k][1] += d[0] * S[3] + d[1] * S[1];
      g[k][0] += d[0] + hoop;
      g[k][1] += d[1];
    }
  }
  if (!ok) raise32(st, ST_RR, e);
  const float p = gmem(a.p)[e] + a.K * (pressure * 0.25f);
  gmem(a.p)[e] = p;
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    auto *fo = gmem(a.fe) + (long long)gmem(a.fpos)[e * 4 + k] * 2;
    fo[0] = -(Af[k][0] + p * g[k][0]);
    fo[1] = -(Af[k][1] + p * g[k][1]);
  }
}

}  // namespace dnl
