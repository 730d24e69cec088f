// Device-side arithmetic of the FP64 element update.
//
// Every function keeps the evaluation order of the reference routine it
// restates (cited per function), and the translation unit that includes this
// header is compiled with -ffp-contract=off, so each result is the double the
// reference's own code produces for the same inputs.  The only departure is
// log/exp/pow: these are deterministic implementations built from IEEE
// +,-,*,/ (the reference calls glibc), so that the GPU reproduces the same
// bits on every run and on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DNL_HD __host__ __device__ __forceinline__

namespace dnl {

// BT/Math.h:36-46, FE/DefaultValues.h:28-29, LA/Tensor2.C:34
constexpr double kPI = 3.1415926535897932384626433832795;
constexpr double kDeuxPI = 6.283185307179586476925286766559;
constexpr double kSqrt23 = 0.8164965809277260327324280249019;
constexpr double kSqrt32 = 1.2247448713915890490986420373529;
constexpr double kRRPrec = 1.0E-8;
constexpr int kRRMaxIt = 2000;
constexpr double kPolarPrec = 10.0E-10;

template <typename T> DNL_HD T dabs(T x) { return x < 0 ? -x : x; }  // Abs, BT/Math.h:72

// ---------------------------------------------------------------------------
// Deterministic log / exp / pow (no reference counterpart; the reference uses
// glibc).  log: x = 2^k (1+f) with sqrt(1/2) <= 1+f < sqrt(2), s = f/(2+f),
// log(1+f) = f - hfsq + s (hfsq + R(s^2)), R(z) = Lg1 z + ... + Lg7 z^7 the
// minimax polynomial of fdlibm's e_log.c (|R - (2 atanh(s)/s - 2)| <
// 2^-58.45 on |s| <= 0.1716), summed as its even and odd halves.
// exp: Cody-Waite reduction by ln 2, degree-13 Taylor polynomial.
// ---------------------------------------------------------------------------
DNL_HD double det_log(double x)
{
  if (!(x > 0.0)) return (x == 0.0) ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) return x;
  uint64_t b = __builtin_bit_cast(uint64_t, x);
  int k = 0;
  if ((b >> 52) == 0) { x = x * 18014398509481984.0; b = __builtin_bit_cast(uint64_t, x); k = -54; }
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 1.4142135623730951) { m = m * 0.5; k += 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * (3.999999999940941908e-01 + w * (2.222219843214978396e-01 + w * 1.531383769920937332e-01));
  const double t2 = z * (6.666666666666735130e-01 + w * (2.857142874366239149e-01 +
                    w * (1.818357216161805012e-01 + w * 1.479819860511658591e-01)));
  const double R = t2 + t1;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  return dk * 6.93147180369123816490e-01 - ((hfsq - (s * (hfsq + R) + dk * 1.90821492927058770002e-10)) - f);
}

DNL_HD double det_exp(double y)
{
  if (y != y) return y;
  if (y > 709.78) return __builtin_inf();
  if (y < -745.2) return 0.0;
  const double kd = y * 1.44269504088896338700e+00;
  const int k = (int)(kd < 0.0 ? kd - 0.5 : kd + 0.5);
  const double dk = (double)k;
  const double r = (y - dk * 6.93147180369123816490e-01) - dk * 1.90821492927058770002e-10;
  const double p = 1.0 + r * (1.0 + r * (0.5 + r * (1.66666666666666666667e-01 +
                   r * (4.16666666666666666667e-02 + r * (8.33333333333333333333e-03 +
                   r * (1.38888888888888888889e-03 + r * (1.98412698412698412698e-04 +
                   r * (2.48015873015873015873e-05 + r * (2.75573192239858906526e-06 +
                   r * (2.75573192239858906526e-07 + r * (2.50521083854417187751e-08 +
                   r * (2.08767569878680989792e-09 + r * 1.60590438368216145994e-10))))))))))));
  const int k1 = k / 2, k2 = k - k1;
  const double s1 = __builtin_bit_cast(double, (uint64_t)(k1 + 1023) << 52);
  const double s2 = __builtin_bit_cast(double, (uint64_t)(k2 + 1023) << 52);
  return (p * s1) * s2;
}

DNL_HD double det_pow(double x, double y)
{
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  if (x == 1.0) return 1.0;
  if (!(x > 0.0)) return (x == 0.0) ? (y > 0.0 ? 0.0 : __builtin_inf()) : __builtin_nan("");
  return det_exp(y * det_log(x));
}

// det_pow(x, y) of one base for several exponents with one logarithm: each
// result is det_pow(x, y) bit for bit (the special cases take det_pow's own
// branches; otherwise both compute det_exp(y * det_log(x)))
struct DetPowBase
{
  double x, L;
  bool have;
  DNL_HD explicit DetPowBase(double x_) : x(x_), L(0.0), have(false) {}
  DNL_HD double pow(double y)
  {
    if (y == 0.0 || y == 1.0 || x == 1.0 || !(x > 0.0)) return det_pow(x, y);
    if (!have)
    {
      L = det_log(x);
      have = true;
    }
    return det_exp(y * L);
  }
};

// ---------------------------------------------------------------------------
// 3x3 tensors, row-major v[3i+j] (LA/Tensor2.h)
// ---------------------------------------------------------------------------
struct T9 { double v[9]; };

// c = a * b, LA/Tensor2.C:360-380
DNL_HD T9 mul(const T9 &a, const T9 &b)
{
  T9 t;
  t.v[0] = a.v[0] * b.v[0] + a.v[1] * b.v[3] + a.v[2] * b.v[6];
  t.v[1] = a.v[0] * b.v[1] + a.v[1] * b.v[4] + a.v[2] * b.v[7];
  t.v[2] = a.v[0] * b.v[2] + a.v[1] * b.v[5] + a.v[2] * b.v[8];
  t.v[3] = a.v[3] * b.v[0] + a.v[4] * b.v[3] + a.v[5] * b.v[6];
  t.v[4] = a.v[3] * b.v[1] + a.v[4] * b.v[4] + a.v[5] * b.v[7];
  t.v[5] = a.v[3] * b.v[2] + a.v[4] * b.v[5] + a.v[5] * b.v[8];
  t.v[6] = a.v[6] * b.v[0] + a.v[7] * b.v[3] + a.v[8] * b.v[6];
  t.v[7] = a.v[6] * b.v[1] + a.v[7] * b.v[4] + a.v[8] * b.v[7];
  t.v[8] = a.v[6] * b.v[2] + a.v[7] * b.v[5] + a.v[8] * b.v[8];
  return t;
}
// c = a * b^T (the product with Transpose(), LA/Tensor2.C:570-586 then 360-380)
DNL_HD T9 mul_bt(const T9 &a, const T9 &b)
{
  T9 t;
  t.v[0] = a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2];
  t.v[1] = a.v[0] * b.v[3] + a.v[1] * b.v[4] + a.v[2] * b.v[5];
  t.v[2] = a.v[0] * b.v[6] + a.v[1] * b.v[7] + a.v[2] * b.v[8];
  t.v[3] = a.v[3] * b.v[0] + a.v[4] * b.v[1] + a.v[5] * b.v[2];
  t.v[4] = a.v[3] * b.v[3] + a.v[4] * b.v[4] + a.v[5] * b.v[5];
  t.v[5] = a.v[3] * b.v[6] + a.v[4] * b.v[7] + a.v[5] * b.v[8];
  t.v[6] = a.v[6] * b.v[0] + a.v[7] * b.v[1] + a.v[8] * b.v[2];
  t.v[7] = a.v[6] * b.v[3] + a.v[7] * b.v[4] + a.v[8] * b.v[5];
  t.v[8] = a.v[6] * b.v[6] + a.v[7] * b.v[7] + a.v[8] * b.v[8];
  return t;
}
// c = a^T * b
DNL_HD T9 mul_at(const T9 &a, const T9 &b)
{
  T9 t;
  t.v[0] = a.v[0] * b.v[0] + a.v[3] * b.v[3] + a.v[6] * b.v[6];
  t.v[1] = a.v[0] * b.v[1] + a.v[3] * b.v[4] + a.v[6] * b.v[7];
  t.v[2] = a.v[0] * b.v[2] + a.v[3] * b.v[5] + a.v[6] * b.v[8];
  t.v[3] = a.v[1] * b.v[0] + a.v[4] * b.v[3] + a.v[7] * b.v[6];
  t.v[4] = a.v[1] * b.v[1] + a.v[4] * b.v[4] + a.v[7] * b.v[7];
  t.v[5] = a.v[1] * b.v[2] + a.v[4] * b.v[5] + a.v[7] * b.v[8];
  t.v[6] = a.v[2] * b.v[0] + a.v[5] * b.v[3] + a.v[8] * b.v[6];
  t.v[7] = a.v[2] * b.v[1] + a.v[5] * b.v[4] + a.v[8] * b.v[7];
  t.v[8] = a.v[2] * b.v[2] + a.v[5] * b.v[5] + a.v[8] * b.v[8];
  return t;
}
// sum of squares in component order, LA/Tensor2.C:484-511
DNL_HD double sumsq(const T9 &a)
{
  return a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2] + a.v[3] * a.v[3] + a.v[4] * a.v[4] +
         a.v[5] * a.v[5] + a.v[6] * a.v[6] + a.v[7] * a.v[7] + a.v[8] * a.v[8];
}

// Tensor2::polarDecomposeLnU, LA/Tensor2.C:1154-1259.  Returns false when the
// Jacobi sweep does not converge in 29 rotations.
DNL_HD bool polar_lnU(const T9 &F, T9 &U, T9 &R)
{
  const T9 FF0 = mul_at(F, F);
  // the six distinct entries of the symmetric F^T F (the dense formula below
  // keeps it symmetric bit for bit) and the accumulated rotation
  double f0 = FF0.v[0], f1 = FF0.v[1], f2 = FF0.v[2], f4 = FF0.v[4], f5 = FF0.v[5], f8 = FF0.v[8];
  double r0[3] = {1, 0, 0}, r1[3] = {0, 1, 0}, r2[3] = {0, 0, 1};  // columns of Rot
  // the sweep; a lane that has converged leaves the loop and waits, so the
  // closing step below runs once per wave instead of once per exit iteration
  bool conv = false;
  for (int it = 1; it < 30; it++)
  {
    int q = (dabs(f1) > dabs(f2) ? 1 : 2);
    const double fq = (q == 1 ? f1 : f2);
    const int p = (dabs(fq) > dabs(f5) ? 0 : 1);
    if (p == 1) q = 2;
    const bool c01 = p == 0 && q == 1, c02 = p == 0 && q == 2, c12 = p == 1;
    // entries relabelled by (p, q, m), m the third index
    const double Apq = c01 ? f1 : (c02 ? f2 : f5);
    if (dabs(Apq) < kPolarPrec)
    {
      conv = true;
      break;
    }
    const double App = c12 ? f4 : f0, Aqq = c01 ? f4 : f8;
    const double Apm = c01 ? f2 : f1, Aqm = c12 ? f2 : f5;
    const double alpha = (Aqq - App) / (2. * Apq);
    const double t = (alpha != 0. ? 1. / (alpha + sqrt(alpha * alpha + 1.) * (alpha > 0. ? 1. : -1.)) : 1.);
    const double Ct = 1. / (sqrt(1. + t * t));
    const double St = t * Ct, mSt = -St;
    // R^T FF R of LA/Tensor2.C:1216-1239 with the rotation's zero and unit
    // entries left out: each kept product and the order of each sum are the
    // dense formula's (a dropped term adds +-0, a unit factor is exact)
    const double wpp = App * Ct + Apq * mSt, wpq = Apq * Ct + Aqq * mSt;
    const double wqp = App * St + Apq * Ct, wqq = Apq * St + Aqq * Ct;
    const double nApp = Ct * wpp + mSt * wpq;
    const double nAqq = St * wqp + Ct * wqq;
    const double nApq = St * wpp + Ct * wpq;
    const double nApm = Apm * Ct + Aqm * mSt;
    const double nAqm = Apm * St + Aqm * Ct;
    const double Amm = c01 ? f8 : (c02 ? f4 : f0);
    f0 = c12 ? f0 : nApp;
    f4 = c01 ? nAqq : (c02 ? Amm : nApp);
    f8 = c01 ? Amm : nAqq;
    f1 = c01 ? nApq : nApm;
    f2 = c01 ? nApm : (c02 ? nApq : nAqm);
    f5 = c12 ? nApq : nAqm;
    // Rot = Rot r: columns p and q mix, column m is unchanged
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
      const double cp = p == 0 ? r0[i] : r1[i], cq = q == 1 ? r1[i] : r2[i];
      const double np = cp * Ct + cq * mSt, nq = cp * St + cq * Ct;
      r0[i] = p == 0 ? np : r0[i];
      r1[i] = c01 ? nq : (c02 ? r1[i] : np);
      r2[i] = q == 2 ? nq : r2[i];
    }
  }
  if (!conv) return false;
  // U = sum_c ln(lam_c)/2 v_c v_c^T, R = F sum_c lam_c^-1/2 v_c v_c^T with
  // v_c the columns of the accumulated rotation (LA/Tensor2.C:1180-1197)
  T9 u, Ru;
  const double lam[3] = {f0, f4, f8};
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
    const double *col = c == 0 ? r0 : (c == 1 ? r1 : r2);
    const double v0 = col[0], v1 = col[1], v2 = col[2];
    const double ext[9] = {v0 * v0, v0 * v1, v0 * v2, v1 * v0, v1 * v1, v1 * v2, v2 * v0, v2 * v1, v2 * v2};
    const double a = det_log(lam[c]) / 2.;
    const double b = 1. / sqrt(lam[c]);
#pragma unroll
    for (int k = 0; k < 9; k++)
    {
      if (c == 0) { u.v[k] = a * ext[k]; Ru.v[k] = b * ext[k]; }
      else { u.v[k] += a * ext[k]; Ru.v[k] += b * ext[k]; }
    }
  }
  U = u;
  R = mul(F, Ru);
  return true;
}

// Matrix::inverse of a 3x3 (LA/Matrix.C:1653-1723): LAPACK dgetf2 (partial
// pivoting, reciprocal scaling, rank-1 update), dtrti2 (upper, non-unit),
// dgetri unblocked column sweep and column swaps; reference BLAS loop orders
// (LP/Blas/dger.c, dtrmv.c, dgemv.c, dscal.c, idamax.c).  Returns false on an
// exactly singular pivot.
DNL_HD bool inv3(const T9 &J, T9 &out)
{
  double a[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) a[i][j] = J.v[3 * i + j];
  int ipiv[3];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 3; j++)
  {
    int jp = j;
    double dmax = __builtin_fabs(a[j][j]);
#pragma unroll
    for (int i = j + 1; i < 3; i++)
      if (!(__builtin_fabs(a[i][j]) <= dmax)) { jp = i; dmax = __builtin_fabs(a[i][j]); }
    ipiv[j] = jp;
    if (jp != j)
    {
#pragma unroll
      for (int c = 0; c < 3; c++)
      {
        // row swap with a register-only select: rows j and jp
        double tj = a[j][c], tp = (jp == 1 ? a[1][c] : a[2][c]);
        a[j][c] = tp;
        if (jp == 1) a[1][c] = tj; else a[2][c] = tj;
      }
    }
    if (a[j][j] != 0.)
    {
      if (j < 2)
      {
        const double r = 1. / a[j][j];
#pragma unroll
        for (int i = j + 1; i < 3; i++) a[i][j] = r * a[i][j];
      }
    }
    else
      ok = false;
    if (j < 2)
#pragma unroll
      for (int c = j + 1; c < 3; c++)
        if (a[j][c] != 0.)
        {
          const double temp = -1. * a[j][c];
#pragma unroll
          for (int i = j + 1; i < 3; i++) a[i][c] = a[i][c] + a[i][j] * temp;
        }
  }
  if (!ok) return false;
#pragma unroll
  for (int j = 0; j < 3; j++)
  {
    a[j][j] = 1. / a[j][j];
    const double ajj = -a[j][j];
#pragma unroll
    for (int jj = 0; jj < j; jj++)
      if (a[jj][j] != 0.)
      {
        const double temp = a[jj][j];
#pragma unroll
        for (int i = 0; i < jj; i++) a[i][j] += temp * a[i][jj];
        a[jj][j] *= a[jj][jj];
      }
#pragma unroll
    for (int i = 0; i < j; i++) a[i][j] = ajj * a[i][j];
  }
  double work[3];
#pragma unroll
  for (int j = 2; j >= 0; j--)
  {
#pragma unroll
    for (int i = j + 1; i < 3; i++) { work[i] = a[i][j]; a[i][j] = 0.; }
    if (j < 2)
#pragma unroll
      for (int c = j + 1; c < 3; c++)
        if (work[c] != 0.)
        {
          const double temp = -1. * work[c];
#pragma unroll
          for (int i = 0; i < 3; i++) a[i][j] += temp * a[i][c];
        }
  }
#pragma unroll
  for (int j = 1; j >= 0; j--)
  {
    const int jp = ipiv[j];
    if (jp != j)
#pragma unroll
      for (int i = 0; i < 3; i++)
      {
        double tj = a[i][j], tp = (jp == 1 ? a[i][1] : a[i][2]);
        a[i][j] = tp;
        if (jp == 1) a[i][1] = tj; else a[i][2] = tj;
      }
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) out.v[3 * i + j] = a[i][j];
  return true;
}

// one row of dN = dN0 * invJ in the dgemm NN order (LP/Blas/dgemm.c: for
// each column j, C(i,j) = 0 then += (1*B(l,j)) * A(i,l) for l with
// B(l,j) != 0), starting from the first product instead of from +0 and
// without the zero test.  Both are value-neutral.  The dN0 entries are
// finite and non-zero, so a product is +-0 exactly when the dgemm skips its
// term; dgemm's sum is therefore 0 + (its non-zero products in order), and
// the sum here carries the same non-zero products in the same order with
// +-0 terms in between: +-0 + p = p and s + (+-0) = s for s != 0, a
// cancellation s + p = 0 gives +0 in both, and the two sums can only differ
// if every product is +-0, i.e. if column j of invJ is zero, which an
// inverse never has.  1*b is b.
template <int D>
DNL_HD void dgemm_row(const double *g, const double *inv, double *out)
{
#pragma unroll
  for (int j = 0; j < D; j++)
  {
    double c = inv[j] * g[0];
#pragma unroll
    for (int l = 1; l < D; l++) c = c + inv[l * D + j] * g[l];
    out[j] = c;
  }
}

}  // namespace dnl
