/*
 * ORACLE -- CPU restatement of the reference's explicit FEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product (redynela_amd, libdnlgpu.so)
 * never links or calls it.
 *
 * What it restates (pantale/ReDynELA, paths relative to
 * Sources/Libraries/finiteElement = FE, linearAlgebra = LA, lapack = LP,
 * basicTools = BT):
 *   - Element::computeInternalForces / computeStress       FE/Element.C:164-403
 *   - Element3D / Element2D::computeInternalMatrices        FE/Element3D.C:53-113,
 *                                                           FE/Element2D.C:53-126
 *   - getF_atIntPoint (3D, axisymmetric), getRadius        FE/Element3D.C:182-205,
 *                                                           FE/ElementAx.C:110-136,
 *                                                           FE/Element2D.C:233-246
 *   - Tensor2::polarDecomposeLnU and helpers                LA/Tensor2.C:1154-1259, 484-586
 *   - Matrix::inverse (dgetrf/dgetf2 + dgetri/dtrti2)       LA/Matrix.C:1653-1723,
 *                                                           LP/Lapack/dget{rf,f2,ri}.c, dtrt{ri,i2}.c
 *   - Matrix::operator* (dgemm NN loop order)               LA/Matrix.C:644-671, LP/Blas/dgemm.c
 *   - hardening laws A + B evp^n                             FE/IsoHardElastoplastic.C:171-184,
 *                                                           FE/IsoHardJohnsonCook.C:185-198
 *   - Domain::computeInternalForces, assembly case 0 with one thread
 *     (element index order)                                 FE/Domain.C:130-193
 *   - Domain::computeMassMatrix, Element::computeMassMatrix FE/Domain.C:948-1005,
 *                                                           FE/Element.C:411-454
 *   - Domain::computeCourantTimeStep, Element::getSoundSpeed,
 *     ElQua4nAx::getLength, Solver::computeTimeStep         FE/Domain.C:1613-1646,
 *                                                           FE/Element.C:607-621,
 *                                                           FE/ElQua4nAx.C:131-143,
 *                                                           FE/Solver.C:220-267
 *   - ExplicitSolverCH::solve, computePredictions, explicitSolve, initStep,
 *     endStep, updateGrid, computeIntegrationParameters     FE/ExplicitSolverCH.C:115-567
 *   - BoundaryRestrain::applyConstant, BoundarySpeed        FE/Boundary.C:378-384, 478-524, 636-666
 *   - Domain::getTotalKineticEnergy                         FE/Domain.C:1330-1344
 *   - IntegrationPoint::flush (evp = 1e-10)                 FE/IntegrationPoint.C:103-119
 *
 * Every floating-point expression keeps the reference's evaluation order
 * (compiled with -ffp-contract=off, x86-64 baseline ISA: no FMA), so that this
 * restatement reproduces the reference's doubles.  Two build-defined pieces
 * have no reference counterpart and are stated where they are defined: the
 * hex8 characteristic length (the reference's ElHex8n3D::getLength exits,
 * FE/ElHex8n3D.C:188-193) and the deterministic log/exp/pow used in
 * "det" math mode (the reference calls glibc).
 *
 * How each part is pinned (the tests named are under tests/):
 *   - linear algebra (polar decomposition, 3x3 inverse, dgemm, RFI product,
 *     deviator, J2, norm): bit for bit against the reference's own LA/LP
 *     sources compiled under oracle/_ref (test_oracle_linalg.py);
 *   - element layer: closed forms and invariants -- shape tables, lumped
 *     mass, uniaxial Hencky response, radial return, rigid rotation, patch
 *     test, axisymmetric hoop strain, Courant lengths, sum of forces
 *     (test_oracle_element.py); the axisymmetric nodal forces of a uniform
 *     stress in a ring element, pressure averaging under a quadratic
 *     increment, the first time step (test_oracle_pins.py);
 *   - Chung-Hulbert update at any rho_b: its one-increment amplification
 *     matrix equals the published explicit generalized-alpha recurrence
 *     (test_oracle_pins.py); central-difference identity at rho_b = 1 and the
 *     LINEARdbl time-step extrapolation (test_oracle_run.py);
 *   - Pulse power iteration against numpy eigenvalues (test_oracle_pulse.py);
 *   - the OpenMP assembly strategies 2, 3, 6, 7: case 6's Fint reassembled in
 *     numpy from the per-element forces and the partition the oracle reports,
 *     bit for bit; the partition against numpy's centroid order
 *     (test_oracle_strategies.py);
 *   - HistoryFile points-to-nodes extrapolation: exact on trilinear (hex8) and
 *     bilinear (quad4) fields, node averages (test_oracle_history.py).
 *   Constant BoundarySpeed (o_apply_constant): the driven component's
 *   speed, displacement increment (speed x that increment's time step) and
 *   coordinate update are checked bit for bit at every increment, and the
 *   undriven components must move (test_deck.py).
 *   Parity unpinned: the Newton iterate sequence for n < 1 (only its
 *   converged consistency is checked) and glibc's last bits of log / pow
 *   (within 1 ulp, test_oracle_math.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants, BT/Math.h:36-46, FE/DefaultValues.h:24-36, LA/Tensor2.C:34 ---- */
#define O_PI 3.1415926535897932384626433832795
#define O_DEUXPI 6.283185307179586476925286766559
#define O_SQRT23 0.8164965809277260327324280249019
#define O_SQRT32 1.2247448713915890490986420373529
#define O_INVSQRT3 0.5773502691896257645091487805019
#define O_RR_PREC 1.0E-8
#define O_RR_MAXIT 2000
#define O_POLAR_PREC 10.0E-10
#define O_ABS(x) ((x) < 0 ? -(x) : (x))
#define O_SQ(a) ((a) * (a))

/* status codes (mirrors of the reference's fatalError sites) */
#define O_OK 0
#define O_E_POLAR 2  /* LA/Tensor2.C:1257-1258 */
#define O_E_RR 3     /* FE/Element.C:327-330 */
#define O_E_DETJ 4   /* FE/Element2D.C:101-109 (2D aborts; 3D only warns) */
#define O_E_ARG 5
#define O_E_PULSE 7  /* FE/Domain.C:936-939: power iteration not converged in 1000 iterations */

/* ========================================================================
 * Deterministic math ("det" mode).  Build-defined, no reference counterpart:
 * the reference calls glibc log/pow.  Only IEEE +,-,*,/ are used, so any
 * IEEE-754 binary64 machine evaluating the same sequence produces the same
 * bits.  log: argument reduction x = 2^k (1+f), sqrt(1/2) <= 1+f < sqrt(2),
 * s = f/(2+f), log(1+f) = f - hfsq + s (hfsq + R(s^2)), R(z) = Lg1 z + ...
 * + Lg7 z^7 the minimax polynomial and arrangement of fdlibm's e_log.c
 * (|R - (2 atanh(s)/s - 2)| < 2^-58.45 for |s| <= 0.1716; its even and odd
 * halves summed as there), the one final formula for every k.  exp: Cody-Waite
 * reduction by ln 2 and a degree-13 Taylor polynomial.  pow(x,y) for x>0:
 * exp(y log x), with y == 1 and y == 0 returned exactly.
 * ======================================================================== */
static const double o_ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
static const double o_ln2_lo = 1.90821492927058770002e-10;
static const double o_inv_ln2 = 1.44269504088896338700e+00;

static double o_u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint64_t o_d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

double dnlo_det_log(double x)
{
  if (!(x > 0.0)) return (x == 0.0) ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  uint64_t b = o_d2u(x);
  int k = 0;
  if ((b >> 52) == 0) { x = x * 18014398509481984.0; b = o_d2u(x); k = -54; } /* subnormal: *2^54 */
  k += (int)((b >> 52) & 0x7ff) - 1023;
  double m = o_u2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
  if (m > 1.4142135623730951) { m = m * 0.5; k += 1; }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  double t1 = w * (3.999999999940941908e-01 + w * (2.222219843214978396e-01 + w * 1.531383769920937332e-01));
  double t2 = z * (6.666666666666735130e-01 + w * (2.857142874366239149e-01 +
              w * (1.818357216161805012e-01 + w * 1.479819860511658591e-01)));
  double R = t2 + t1;
  double hfsq = 0.5 * f * f;
  double dk = (double)k;
  return dk * o_ln2_hi - ((hfsq - (s * (hfsq + R) + dk * o_ln2_lo)) - f);
}

double dnlo_det_exp(double y)
{
  if (y != y) return y;
  if (y > 709.78) return INFINITY;
  if (y < -745.2) return 0.0;
  double kd = y * o_inv_ln2;
  int k = (int)(kd < 0.0 ? kd - 0.5 : kd + 0.5);
  double dk = (double)k;
  double r = (y - dk * o_ln2_hi) - dk * o_ln2_lo;
  double p = 1.0 + r * (1.0 + r * (0.5 + r * (1.66666666666666666667e-01 +
             r * (4.16666666666666666667e-02 + r * (8.33333333333333333333e-03 +
             r * (1.38888888888888888889e-03 + r * (1.98412698412698412698e-04 +
             r * (2.48015873015873015873e-05 + r * (2.75573192239858906526e-06 +
             r * (2.75573192239858906526e-07 + r * (2.50521083854417187751e-08 +
             r * (2.08767569878680989792e-09 + r * 1.60590438368216145994e-10))))))))))));
  /* scale by 2^k in two steps so that k down to -1074 stays representable */
  int k1 = k / 2, k2 = k - k1;
  double s1 = o_u2d((uint64_t)(k1 + 1023) << 52);
  double s2 = o_u2d((uint64_t)(k2 + 1023) << 52);
  return (p * s1) * s2;
}

double dnlo_det_pow(double x, double y)
{
  if (y == 0.0) return 1.0;
  if (y == 1.0) return x;
  if (x == 1.0) return 1.0;
  if (!(x > 0.0)) return (x == 0.0) ? (y > 0.0 ? 0.0 : INFINITY) : NAN;
  return dnlo_det_exp(y * dnlo_det_log(x));
}

static double o_log(int mode, double x) { return mode ? dnlo_det_log(x) : log(x); }
static double o_pow(int mode, double x, double y) { return mode ? dnlo_det_pow(x, y) : pow(x, y); }

/* ========================================================================
 * 3x3 tensors, row-major v[3i+j] as LA/Tensor2.h
 * ======================================================================== */
/* c = a * b, LA/Tensor2.C:360-380 */
static void t2_mul(const double *a, const double *b, double *c)
{
  double t[9];
  t[0] = a[0] * b[0] + a[1] * b[3] + a[2] * b[6];
  t[1] = a[0] * b[1] + a[1] * b[4] + a[2] * b[7];
  t[2] = a[0] * b[2] + a[1] * b[5] + a[2] * b[8];
  t[3] = a[3] * b[0] + a[4] * b[3] + a[5] * b[6];
  t[4] = a[3] * b[1] + a[4] * b[4] + a[5] * b[7];
  t[5] = a[3] * b[2] + a[4] * b[5] + a[5] * b[8];
  t[6] = a[6] * b[0] + a[7] * b[3] + a[8] * b[6];
  t[7] = a[6] * b[1] + a[7] * b[4] + a[8] * b[7];
  t[8] = a[6] * b[2] + a[7] * b[5] + a[8] * b[8];
  memcpy(c, t, sizeof t);
}
/* LA/Tensor2.C:570-586 */
static void t2_transpose(const double *a, double *t)
{
  t[0] = a[0]; t[1] = a[3]; t[2] = a[6];
  t[3] = a[1]; t[4] = a[4]; t[5] = a[7];
  t[6] = a[2]; t[7] = a[5]; t[8] = a[8];
}
/* LA/Tensor2.C:484-490 */
static double t2_norme(const double *v)
{
  return sqrt(O_SQ(v[0]) + O_SQ(v[1]) + O_SQ(v[2]) + O_SQ(v[3]) + O_SQ(v[4]) + O_SQ(v[5]) +
              O_SQ(v[6]) + O_SQ(v[7]) + O_SQ(v[8]));
}
/* LA/Tensor2.C:505-511 */
static double t2_J2(const double *v)
{
  return O_SQRT32 * sqrt(O_SQ(v[0]) + O_SQ(v[1]) + O_SQ(v[2]) + O_SQ(v[3]) + O_SQ(v[4]) +
                         O_SQ(v[5]) + O_SQ(v[6]) + O_SQ(v[7]) + O_SQ(v[8]));
}
/* LA/Tensor2.C:543-560 with moyTrace LA/Tensor2.h:331-337 */
static void t2_deviator(const double *v, double *d)
{
  double tr = (v[0] + v[4] + v[8]) / 3.0;
  memcpy(d, v, 9 * sizeof(double));
  d[0] = v[0] - tr; d[4] = v[4] - tr; d[8] = v[8] - tr;
}
/* R (S R^T): the RFI product, FE/Element.C:383-389 */
static void t2_rfi(const double *R, double *S)
{
  double Rt[9], SRt[9];
  t2_transpose(R, Rt);
  t2_mul(S, Rt, SRt);
  t2_mul(R, SRt, S);
}

/* the Jacobi stop |(F^T F)_pq| < 1e-9 of LA/Tensor2.C:34; test-only override
 * (a smaller stop removes the shear dead zone below it from finite differences) */
static double o_polar_prec = O_POLAR_PREC;
void dnlo_set_polar_precision(double tol) { o_polar_prec = tol; }

/* Tensor2::polarDecomposeLnU, LA/Tensor2.C:1154-1259.
 * F (row-major) -> U = ln(sqrt(F^T F)) and R = F (F^T F)^(-1/2). */
int dnlo_polar_lnU(const double *F, double *U, double *R, int mathmode)
{
  double FF[9], Ft[9], Rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Rr[9], u[9];
  int p, q, it = 1;
  t2_transpose(F, Ft);
  t2_mul(Ft, F, FF);
  while (it < 30)
  {
    q = (O_ABS(FF[1]) > O_ABS(FF[2]) ? 1 : 2);
    p = (O_ABS(FF[q]) > O_ABS(FF[5]) ? 0 : 1);
    if (p == 1) q = 2;
    if (O_ABS(FF[3 * p + q]) < o_polar_prec)
    {
      double lam[3] = {FF[0], FF[4], FF[8]};
      double Ru[9];
      for (int c = 0; c < 3; c++)
      {
        double vec[3] = {Rot[c], Rot[3 + c], Rot[6 + c]}; /* getCol */
        double ext[9];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) ext[3 * i + j] = vec[i] * vec[j];
        double a = o_log(mathmode, lam[c]) / 2.;
        double b = 1. / sqrt(lam[c]);
        for (int k = 0; k < 9; k++)
        {
          double ua = a * ext[k], rb = b * ext[k];
          if (c == 0) { u[k] = ua; Ru[k] = rb; }
          else { u[k] += ua; Ru[k] += rb; }
        }
      }
      memcpy(U, u, sizeof u);
      t2_mul(F, Ru, R);
      return O_OK;
    }
    double alpha = (FF[3 * q + q] - FF[3 * p + p]) / (2. * FF[3 * p + q]);
    double t = (alpha != 0. ? 1. / (alpha + sqrt(alpha * alpha + 1.) * (alpha > 0. ? 1. : -1.)) : 1.);
    double Ct = 1. / (sqrt(1. + t * t));
    double St = t * Ct;
    for (int k = 0; k < 9; k++) Rr[k] = (k % 4 == 0) ? 1.0 : 0.0;
    Rr[3 * p + q] = St;
    Rr[3 * q + p] = -St;
    Rr[3 * p + p] = Ct;
    Rr[3 * q + q] = Ct;
    /* R^T FF R with FF symmetric, dense formula LA/Tensor2.C:1216-1239 */
    u[0] = Rr[0] * (FF[0] * Rr[0] + FF[3] * Rr[3] + FF[6] * Rr[6]) +
           Rr[3] * (FF[1] * Rr[0] + FF[4] * Rr[3] + FF[7] * Rr[6]) +
           Rr[6] * (FF[2] * Rr[0] + FF[5] * Rr[3] + FF[8] * Rr[6]);
    u[1] = Rr[1] * (FF[0] * Rr[0] + FF[3] * Rr[3] + FF[6] * Rr[6]) +
           Rr[4] * (FF[1] * Rr[0] + FF[4] * Rr[3] + FF[7] * Rr[6]) +
           (FF[2] * Rr[0] + FF[5] * Rr[3] + FF[8] * Rr[6]) * Rr[7];
    u[2] = Rr[2] * (FF[0] * Rr[0] + FF[3] * Rr[3] + FF[6] * Rr[6]) +
           Rr[5] * (FF[1] * Rr[0] + FF[4] * Rr[3] + FF[7] * Rr[6]) +
           (FF[2] * Rr[0] + FF[5] * Rr[3] + FF[8] * Rr[6]) * Rr[8];
    u[4] = Rr[1] * (FF[0] * Rr[1] + FF[3] * Rr[4] + FF[6] * Rr[7]) +
           Rr[4] * (FF[1] * Rr[1] + FF[4] * Rr[4] + FF[7] * Rr[7]) +
           Rr[7] * (FF[2] * Rr[1] + FF[5] * Rr[4] + FF[8] * Rr[7]);
    u[5] = Rr[2] * (FF[0] * Rr[1] + FF[3] * Rr[4] + FF[6] * Rr[7]) +
           Rr[5] * (FF[1] * Rr[1] + FF[4] * Rr[4] + FF[7] * Rr[7]) +
           (FF[2] * Rr[1] + FF[5] * Rr[4] + FF[8] * Rr[7]) * Rr[8];
    u[8] = Rr[2] * (FF[0] * Rr[2] + FF[3] * Rr[5] + FF[6] * Rr[8]) +
           Rr[5] * (FF[1] * Rr[2] + FF[4] * Rr[5] + FF[7] * Rr[8]) +
           Rr[8] * (FF[2] * Rr[2] + FF[5] * Rr[5] + FF[8] * Rr[8]);
    FF[0] = u[0]; FF[1] = u[1]; FF[2] = u[2];
    FF[3] = u[1]; FF[4] = u[4]; FF[5] = u[5];
    FF[6] = u[2]; FF[7] = u[5]; FF[8] = u[8];
    t2_mul(Rot, Rr, Rot);
    it++;
  }
  return O_E_POLAR;
}

/* ========================================================================
 * Matrix::inverse for 3x3 (LA/Matrix.C:1653-1723): LP dgetrf -> dgetf2
 * (unblocked, ilaenv NB=64 >= 3), then dgetri (unblocked) with dtrtri ->
 * dtrti2.  a[i][j] is the logical (row i, column j) entry.
 * ======================================================================== */
int dnlo_inv3(const double *J, double *inv)
{
  double a[3][3];
  int ipiv[3], info = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = J[3 * i + j];
  /* dgetf2, LP/Lapack/dgetf2.c */
  for (int j = 0; j < 3; j++)
  {
    int jp = j; /* idamax over a[j..2][j], first strict maximum */
    double dmax = fabs(a[j][j]);
    for (int i = j + 1; i < 3; i++)
      if (!(fabs(a[i][j]) <= dmax)) { jp = i; dmax = fabs(a[i][j]); }
    ipiv[j] = jp;
    if (a[jp][j] != 0.)
    {
      if (jp != j)
        for (int c = 0; c < 3; c++) { double t = a[j][c]; a[j][c] = a[jp][c]; a[jp][c] = t; }
      if (j < 2)
      {
        double r = 1. / a[j][j];
        for (int i = j + 1; i < 3; i++) a[i][j] = r * a[i][j];   /* dscal */
      }
    }
    else if (info == 0)
      info = j + 1;
    if (j < 2) /* dger, alpha = -1 */
      for (int c = j + 1; c < 3; c++)
        if (a[j][c] != 0.)
        {
          double temp = -1. * a[j][c];
          for (int i = j + 1; i < 3; i++) a[i][c] = a[i][c] + a[i][j] * temp;
        }
  }
  if (info) return O_E_DETJ;
  /* dtrtri -> dtrti2 upper, non-unit */
  for (int j = 0; j < 3; j++)
  {
    a[j][j] = 1. / a[j][j];
    double ajj = -a[j][j];
    /* dtrmv upper, no-trans, non-unit: x := A x, x = a[0..j-1][j] */
    for (int jj = 0; jj < j; jj++)
      if (a[jj][j] != 0.)
      {
        double temp = a[jj][j];
        for (int i = 0; i < jj; i++) a[i][j] += temp * a[i][jj];
        a[jj][j] *= a[jj][jj];
      }
    for (int i = 0; i < j; i++) a[i][j] = ajj * a[i][j]; /* dscal */
  }
  /* dgetri unblocked: solve inv(A) L = inv(U) */
  double work[3];
  for (int j = 2; j >= 0; j--)
  {
    for (int i = j + 1; i < 3; i++) { work[i] = a[i][j]; a[i][j] = 0.; }
    if (j < 2) /* dgemv no-trans, alpha = -1, beta = 1 */
      for (int c = j + 1; c < 3; c++)
        if (work[c] != 0.)
        {
          double temp = -1. * work[c];
          for (int i = 0; i < 3; i++) a[i][j] += temp * a[i][c];
        }
  }
  for (int j = 1; j >= 0; j--)
  {
    int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < 3; i++) { double t = a[i][j]; a[i][j] = a[i][jp]; a[i][jp] = t; }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) inv[3 * i + j] = a[i][j];
  return O_OK;
}

/* C(m x n) = A(m x k) * B(k x n), dgemm NN loop order (LP/Blas/dgemm.c,
 * alpha = 1, beta = 0): C(i,j) accumulates temp*A(i,l) over l, skipping
 * B(l,j) == 0.  Row-major storage of all three. */
void dnlo_gemm(int m, int n, int k, const double *A, const double *B, double *C)
{
  for (int j = 0; j < n; j++)
  {
    for (int i = 0; i < m; i++) C[i * n + j] = 0.;
    for (int l = 0; l < k; l++)
      if (B[l * n + j] != 0.)
      {
        double temp = 1.0 * B[l * n + j];
        for (int i = 0; i < m; i++) C[i * n + j] = C[i * n + j] + temp * A[i * k + l];
      }
  }
}

/* exported tensor helpers (for golden-vector checks) */
void dnlo_t2_rfi(const double *R, const double *S, double *out) { memcpy(out, S, 72); t2_rfi(R, out); }
double dnlo_t2_J2(const double *S) { return t2_J2(S); }
double dnlo_t2_norme(const double *S) { return t2_norme(S); }
void dnlo_t2_deviator(const double *E, double *out) { t2_deviator(E, out); }

/* ========================================================================
 * Element shape tables.  The values are the closed forms of the trilinear /
 * bilinear shape functions at the 2-point Gauss points, written with the
 * constant expressions of FE/ElHex8n3D.C:35-38, 81-87 and
 * FE/ElQua4nAx.C:41-43, 53-57 so the doubles are the reference's.
 * Node and integration-point order: (-,-,-),(+,-,-),(+,+,-),(-,+,-),
 * (-,-,+),(+,-,+),(+,+,+),(-,+,+) (hex8); (-,-),(+,-),(+,+),(-,+) (quad4).
 * ======================================================================== */
static const int o_sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
static double H8_N[8][8], H8_dN[8][8][3]; /* [ip][node], [ip][node][dir] */
static double Q4_N[4][4], Q4_dN[4][4][2];
static int o_tables_ready = 0;

static void o_build_tables(void)
{
  if (o_tables_ready) return;
  const double deuxPR3s12 = (2.0 + sqrt(3.0)) / 12.0, uns12 = 1.0 / 12.0,
               deuxMR3s12 = (2.0 - sqrt(3.0)) / 12.0;
  const double neufP5R3s36 = (9.0 + 5.0 * sqrt(3.0)) / 36.0, troisPR3s36 = (3.0 + sqrt(3.0)) / 36.0,
               troisMR3s36 = (3.0 - sqrt(3.0)) / 36.0, neufM5R3s36 = (9.0 - 5.0 * sqrt(3.0)) / 36.0;
  for (int p = 0; p < 8; p++)
    for (int k = 0; k < 8; k++)
    {
      int agree = 0; /* number of directions where node and IP lie on the same side */
      for (int d = 0; d < 3; d++) agree += (o_sg[p][d] == o_sg[k][d]);
      H8_N[p][k] = agree == 3 ? neufP5R3s36 : agree == 2 ? troisPR3s36 : agree == 1 ? troisMR3s36 : neufM5R3s36;
      for (int d = 0; d < 3; d++)
      {
        int others = 0;
        for (int e = 0; e < 3; e++)
          if (e != d) others += (o_sg[p][e] == o_sg[k][e]);
        double mag = others == 2 ? deuxPR3s12 : others == 1 ? uns12 : deuxMR3s12;
        H8_dN[p][k][d] = o_sg[k][d] > 0 ? mag : -mag;
      }
    }
  const double deuxPR3s6 = (2.0 + sqrt(3.0)) / 6.0, deuxMR3s6 = (2.0 - sqrt(3.0)) / 6.0, uns6 = 1.0 / 6.0;
  const double troisPR3s12 = (3.0 + sqrt(3.0)) / 12.0, troisMR3s12 = (3.0 - sqrt(3.0)) / 12.0;
  for (int p = 0; p < 4; p++)
    for (int k = 0; k < 4; k++)
    {
      int agree = (o_sg[p][0] == o_sg[k][0]) + (o_sg[p][1] == o_sg[k][1]);
      Q4_N[p][k] = agree == 2 ? deuxPR3s6 : agree == 1 ? uns6 : deuxMR3s6;
      for (int d = 0; d < 2; d++)
      {
        int e = 1 - d;
        double mag = (o_sg[p][e] == o_sg[k][e]) ? troisPR3s12 : troisMR3s12;
        Q4_dN[p][k][d] = o_sg[k][d] > 0 ? mag : -mag;
      }
    }
  o_tables_ready = 1;
}

void dnlo_shape_tables(int dim, double *N, double *dN)
{
  o_build_tables();
  if (dim == 3) { memcpy(N, H8_N, sizeof H8_N); memcpy(dN, H8_dN, sizeof H8_dN); }
  else { memcpy(N, Q4_N, sizeof Q4_N); memcpy(dN, Q4_dN, sizeof Q4_dN); }
}

/* ========================================================================
 * Model state
 * ======================================================================== */
typedef struct
{
  int dim, nen, nip;
  long nn, ne;
  /* nodes: AoS with 3 components, as Node::coords / NodalField Vec3D */
  double *x, *x0;
  double *v[2], *a[2], *du[2], *d[2]; /* the two NodalFields; cur indexes Current */
  int cur;
  double *mass; /* per node (Mass(Id*dim)) */
  unsigned char *bc;   /* bit i: BoundaryRestrain on component i */
  unsigned char *bcs;  /* bit i: constant BoundarySpeed on component i */
  double *bcs_v;       /* prescribed speed (3 per node) */
  int *conn;
  /* integration points [e*nip + pt] */
  double *S, *Eps, *EpsPlas, *dEps, *R, *Sig; /* 9 each */
  double *p, *evp;
  double *dN, *detJ; /* [e*nip+pt][nen*dim], current configuration */
  double *fint;      /* dim*nn, node-major [node*dim + j] */
  double *Fe;        /* ne*nen*dim scratch */
  int *Fe_status;
  /* material */
  double E, nu, rho, A, B, n;
  int yield;
  /* Johnson-Cook rate / thermal extension (law 2) */
  int law;
  double C, epsd0, mjc, T0, Tm, chi_rho_cp;
  double *T;
  /* solver */
  double rho_b, alphaM, beta, gamma, omegaS, safety;
  int freq;
  double t, dt, upTime, endTime;
  long increment;
  double computedTimeStep, timeStepComputationDate, lastTimeStepComputationDate, lastTimeStep;
  int mathmode, mass_done, ke_on;
  double rr_prec; /* radial-return stop, 1e-8 as the reference (FE/DefaultValues.h:29) */
  double ke;
  int status;
  /* time-step method: 0 Courant, 1 Pulse (Solver::setTimeStepMethod, FE/Solver.C:312-327) */
  int ts_method;
  /* assembly strategy of Domain::computeInternalForces (FE/Domain.C:130-738):
   * 0 the oracle's own order (element-parallel work, sequential assembly in
   * element order = case 0 with one thread); 2, 3, 6, 7 the reference's cases
   * (dnlo_set_strategy) */
  int strategy;
  int nthreads;      /* threads of the strategy (omp_get_max_threads in the reference) */
  int mapped;        /* Jobs::mapInitial done (FE/Jobs.C:129-131: once) */
  int *job_el;       /* elements of job 0, then job 1, ... (the sorted centroid order) */
  long *job_off;     /* [nthreads + 1] offsets into job_el */
  double *part_fint; /* [nthreads][nn * dim] per-thread partial Fint */
  /* Domain::getMaximumFrequency state (FE/Domain.h:65-66): the last iterate and
   * fmax0, kept from one call to the next; Ke per element [ne][nd*nd] */
  double fmax0;
  double *ev, *Ke;
  long pulse_iters; /* iterations of the last call */
  double plot_row[4]; /* the endStep history line of the last increment, FE/ExplicitSolverCH.C:459-461 */
  double pulse_conv; /* its last convergence measure */
  double forces_seconds; /* wall time spent in Domain::computeInternalForces by dnlo_step, the
                          * region the reference times itself (FE/ExplicitSolverCH.C:475-478) */
} dnlo;

static double o_wtime(void)
{
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return (double)clock() / CLOCKS_PER_SEC;
#endif
}

static double *o_alloc(size_t n)
{
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(3); }
  return p;
}

void dnlo_free(dnlo *m);

/* Model creation.  x0: nn*3, conn: ne*nen (0-based local order of the
 * reference element), v0: nn*3 initial velocity (BoundarySpeed initial
 * condition, FE/Boundary.C:378-384), bc: nn restraint bitmask. */
dnlo *dnlo_create(int dim, long nn, long ne, const double *x0, const int *conn, const unsigned char *bc,
                  const double *v0)
{
  if (dim != 2 && dim != 3) return NULL;
  o_build_tables();
  dnlo *m = (dnlo *)calloc(1, sizeof(dnlo));
  m->dim = dim;
  m->nen = dim == 3 ? 8 : 4;
  m->nip = dim == 3 ? 8 : 4;
  m->nn = nn;
  m->ne = ne;
  m->x = o_alloc(nn * 3);
  m->x0 = o_alloc(nn * 3);
  for (int f = 0; f < 2; f++)
  {
    m->v[f] = o_alloc(nn * 3); m->a[f] = o_alloc(nn * 3);
    m->du[f] = o_alloc(nn * 3); m->d[f] = o_alloc(nn * 3);
  }
  m->mass = o_alloc(nn);
  m->bc = (unsigned char *)calloc(nn, 1);
  m->bcs = (unsigned char *)calloc(nn, 1);
  m->bcs_v = o_alloc(nn * 3);
  m->conn = (int *)malloc(sizeof(int) * ne * m->nen);
  memcpy(m->conn, conn, sizeof(int) * ne * m->nen);
  memcpy(m->x, x0, sizeof(double) * nn * 3);
  memcpy(m->x0, x0, sizeof(double) * nn * 3);
  if (bc) memcpy(m->bc, bc, nn);
  long nipt = ne * m->nip;
  m->S = o_alloc(nipt * 9); m->Eps = o_alloc(nipt * 9); m->EpsPlas = o_alloc(nipt * 9);
  m->dEps = o_alloc(nipt * 9); m->R = o_alloc(nipt * 9); m->Sig = o_alloc(nipt * 9);
  m->p = o_alloc(nipt); m->evp = o_alloc(nipt);
  m->dN = o_alloc(nipt * m->nen * dim);
  m->detJ = o_alloc(nipt);
  m->fint = o_alloc(nn * dim);
  m->Fe = o_alloc(ne * m->nen * dim);
  m->Fe_status = (int *)calloc(ne, sizeof(int));
  for (long i = 0; i < nipt; i++)
  {
    m->evp[i] = 1e-10; /* IntegrationPoint::flush, FE/IntegrationPoint.C:117 */
    for (int k = 0; k < 9; k++) m->R[9 * i + k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  /* Domain::initSolve applyInitial_V0 (FE/Domain.C:1142-1146): BoundarySpeed
   * sets Current and New mat_speed; BoundaryRestrain::applyInitial_V0 is empty */
  if (v0)
  {
    memcpy(m->v[0], v0, sizeof(double) * nn * 3);
    memcpy(m->v[1], v0, sizeof(double) * nn * 3);
  }
  m->cur = 0;
  m->rho_b = 0.8182; /* FE/DefaultValues.h:25 */
  m->safety = 0.9;
  m->freq = 10;
  m->rr_prec = O_RR_PREC;
  m->endTime = 1e300;
  m->yield = 1;
  return m;
}

/* constant BoundarySpeed on selected components of one node */
void dnlo_set_speed_bc(dnlo *m, long node, int mask, const double *speed)
{
  m->bcs[node] = (unsigned char)mask;
  for (int i = 0; i < 3; i++) m->bcs_v[3 * node + i] = speed[i];
}

void dnlo_set_material(dnlo *m, double E, double nu, double rho, double A, double B, double n, int yield)
{
  m->E = E; m->nu = nu; m->rho = rho; m->A = A; m->B = B; m->n = n; m->yield = yield;
  m->law = yield ? 1 : 0;
}

/* Johnson-Cook with rate and thermal factors (extension, see o_stress_jc).
 * With C = 0 and taylor_quinney = 0 the reference law (law 1) is used. */
void dnlo_set_johnson_cook(dnlo *m, double C, double epsd0, double mexp, double T0, double Tm, double cp, double tq)
{
  m->C = C; m->epsd0 = epsd0; m->mjc = mexp; m->T0 = T0; m->Tm = Tm;
  m->chi_rho_cp = (cp > 0.0) ? tq / (m->rho * cp) : 0.0;
  m->law = (C == 0.0 && m->chi_rho_cp == 0.0) ? 1 : 2;
  long nipt = m->ne * m->nip;
  if (!m->T) m->T = o_alloc(nipt);
  for (long i = 0; i < nipt; i++) m->T[i] = T0;
}

/* ExplicitSolverCH::computeIntegrationParameters, FE/ExplicitSolverCH.C:336-343 */
void dnlo_set_solver(dnlo *m, double rho_b, double safety, int freq, double endTime)
{
  m->rho_b = rho_b;
  m->safety = safety;
  m->freq = freq;
  m->endTime = endTime;
  m->alphaM = (2.0 * rho_b - 1.0) / (1.0 + rho_b);
  m->beta = (5.0 - 3.0 * rho_b) / ((1.0 + rho_b) * (1.0 + rho_b) * (2.0 - rho_b));
  m->gamma = 1.5 - m->alphaM;
  m->omegaS = sqrt((12.0 * pow((1 + rho_b), 3) * (2.0 - rho_b)) /
                   (10 + 15 * rho_b - rho_b * rho_b + pow(rho_b, 3) - pow(rho_b, 4)));
}

void dnlo_set_mathmode(dnlo *m, int mode) { m->mathmode = mode; }
/* test-only: a tighter radial-return stop than the reference's 1e-8, to take
 * the iteration-count jumps out of finite differences of the solve */
void dnlo_set_rr_precision(dnlo *m, double tol) { m->rr_prec = tol; }
void dnlo_set_ke(dnlo *m, int on) { m->ke_on = on; }
void dnlo_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ========================================================================
 * Element kernels
 * ======================================================================== */
/* Element3D::computeInternalMatrices (FE/Element3D.C:53-113) and
 * Element2D::computeInternalMatrices (FE/Element2D.C:53-126) */
static int o_internal_matrices(dnlo *m, long e)
{
  const int *cn = m->conn + e * m->nen;
  int st = O_OK;
  for (int pt = 0; pt < m->nip; pt++)
  {
    long ip = e * m->nip + pt;
    double *dN = m->dN + ip * m->nen * m->dim;
    if (m->dim == 3)
    {
      double J[9] = {0}, inv[9];
      for (int k = 0; k < 8; k++)
      {
        const double *x = m->x + 3 * (long)cn[k];
        const double *d0 = H8_dN[pt][k];
        J[0] += d0[0] * x[0]; J[1] += d0[1] * x[0]; J[2] += d0[2] * x[0];
        J[3] += d0[0] * x[1]; J[4] += d0[1] * x[1]; J[5] += d0[2] * x[1];
        J[6] += d0[0] * x[2]; J[7] += d0[1] * x[2]; J[8] += d0[2] * x[2];
      }
      m->detJ[ip] = (J[0] * J[4] * J[8]) + (J[3] * J[7] * J[2]) + (J[6] * J[1] * J[5]) -
                    (J[6] * J[4] * J[2]) - (J[0] * J[7] * J[5]) - (J[3] * J[1] * J[8]);
      /* negative detJ only warns in 3D (FE/Element3D.C:98-105) */
      if (dnlo_inv3(J, inv) != O_OK) st = O_E_DETJ;
      dnlo_gemm(8, 3, 3, &H8_dN[pt][0][0], inv, dN);
    }
    else
    {
      double J[4] = {0}, inv[4];
      for (int k = 0; k < 4; k++)
      {
        const double *x = m->x + 3 * (long)cn[k];
        const double *d0 = Q4_dN[pt][k];
        J[0] += d0[0] * x[0]; J[1] += d0[1] * x[0];
        J[2] += d0[0] * x[1]; J[3] += d0[1] * x[1];
      }
      double detJ = (J[0] * J[3]) - (J[1] * J[2]);
      m->detJ[ip] = detJ;
      if (detJ < 0.) st = O_E_DETJ; /* FE/Element2D.C:101-109 aborts */
      inv[0] = J[3] / detJ;
      inv[1] = -J[1] / detJ;
      inv[2] = -J[2] / detJ;
      inv[3] = J[0] / detJ;
      dnlo_gemm(4, 2, 2, &Q4_dN[pt][0][0], inv, dN);
    }
  }
  return st;
}

/* getRadius_atIntPoint, FE/Element2D.C:233-246 */
static double o_radius(const dnlo *m, long e, int pt)
{
  const int *cn = m->conn + e * m->nen;
  double R = 0.;
  for (int i = 0; i < 4; i++) R += Q4_N[pt][i] * m->x[3 * (long)cn[i]];
  return R;
}

/* Johnson-Cook with rate and thermal factors -- build extension, no
 * reference counterpart (FE/IsoHardJohnsonCook.h:51-68 documents this law;
 * FE/IsoHardJohnsonCook.C:185-198 evaluates only A + B evp^n):
 *   sigma_y = (A + B evp^n) [1 + C ln(epsdot/epsdot0)] theta(T),
 *   rate factor 1 when epsdot < epsdot0, theta = 1 - ((T-T0)/(Tm-T0))^m on
 *   [T0, Tm), 1 below, 0 above; epsdot = sqrt(2/3) gamma / dt.
 * Newton on Strial - sqrt(3/2) 2G gamma - sigma_y = 0 from the reference's
 * initial guess and with its stop |res/sigma_y| <= 1e-8 (FE/Element.C:292-331),
 * kept inside a sign bracket [lo, hi] (bisection when a step leaves it): the
 * rate factor has a kink at epsdot = epsdot0 where plain Newton cycles.
 * Adiabatic heating dT = chi/(rho cp) sigma_y depsp after convergence. */
static double o_theta(const dnlo *m, double T)
{
  if (m->chi_rho_cp == 0.0) return 1.0;
  if (T < m->T0) return 1.0;
  if (T >= m->Tm) return 0.0;
  return 1.0 - o_pow(m->mathmode, (T - m->T0) / (m->Tm - m->T0), m->mjc);
}

static int o_stress_jc(dnlo *m, long ip, double *S, double G)
{
  const int mm = m->mathmode;
  double evp = m->evp[ip];
  double theta = o_theta(m, m->T[ip]);
  double sy0 = (m->A + m->B * o_pow(mm, evp, m->n)) * theta;
  double nrm = t2_norme(S);
  double Strial = O_SQRT32 * nrm;
  if (!(Strial > sy0)) return O_OK;
  double h0 = m->n * m->B * o_pow(mm, evp, (m->n - 1.)) * theta;
  double gamma = (nrm - O_SQRT23 * sy0) / (2.0 * G * (1.0 + h0 / (3.0 * G)));
  /* res(gamma) decreases from res(0) > 0 to res(hi) = -sigma_y < 0 */
  double lo = 0.0, hi = Strial / (O_SQRT32 * 2.0 * G);
  if (!(gamma > lo && gamma < hi)) gamma = 0.5 * (lo + hi);
  double sy = sy0;
  int it = 0;
  while (1)
  {
    double ev = evp + O_SQRT23 * gamma;
    double base = m->A + m->B * o_pow(mm, ev, m->n);
    double dbase = m->n * m->B * o_pow(mm, ev, (m->n - 1.));
    double rf = 1.0, drf = 0.0;
    if (m->C != 0.0 && gamma > 0.0)
    {
      double epsd = O_SQRT23 * gamma / m->dt;
      if (epsd >= m->epsd0)
      {
        rf = 1.0 + m->C * o_log(mm, epsd / m->epsd0);
        drf = m->C / gamma;
      }
    }
    sy = base * rf * theta;
    double dsy = (O_SQRT23 * dbase * rf + base * drf) * theta;
    double res = Strial - O_SQRT32 * 2.0 * G * gamma - sy;
    it++;
    if (O_ABS(res / sy) <= m->rr_prec) break;
    if (it > O_RR_MAXIT) return O_E_RR;
    if (res > 0.0) lo = gamma; else hi = gamma;
    double gn = gamma + res / (O_SQRT32 * 2.0 * G + dsy);
    gamma = (gn > lo && gn < hi) ? gn : 0.5 * (lo + hi);  /* Newton inside the bracket, else bisection */
  }
  double n1[9];
  for (int k = 0; k < 9; k++) n1[k] = S[k] / nrm;
  double dep = O_SQRT23 * gamma;
  m->evp[ip] += dep;
  double c = 2.0 * G * gamma;
  for (int k = 0; k < 9; k++) S[k] -= c * n1[k];
  for (int k = 0; k < 9; k++) m->EpsPlas[9 * ip + k] += gamma * n1[k];
  m->T[ip] += m->chi_rho_cp * sy * dep;
  return O_OK;
}

/* Element::computeStress, FE/Element.C:217-403 (G2TR_STUDY branch with
 * RFI = True and underIntegratePressure = True, FE/Solver.C:74-75) */
static int o_compute_stress(dnlo *m, long e)
{
  const int *cn = m->conn + e * m->nen;
  const double *DU = m->du[m->cur]; /* getFieldAtTime(0) = Current, FE/Node.h:178-183 */
  const int mm = m->mathmode;
  double pressure = 0.0;
  double K = m->E / (3. * (1. - 2. * m->nu));
  double G = m->E / (2. * (1. + m->nu)); /* Material::mu, FE/Material.h:167-172 */
  for (int pt = 0; pt < m->nip; pt++)
  {
    long ip = e * m->nip + pt;
    const double *dN = m->dN + ip * m->nen * m->dim;
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (m->dim == 3)
    { /* FE/Element3D.C:182-205 */
      for (int k = 0; k < 8; k++)
      {
        const double *u = DU + 3 * (long)cn[k];
        const double *g = dN + 3 * k;
        F[0] += g[0] * u[0]; F[1] += g[1] * u[0]; F[2] += g[2] * u[0];
        F[3] += g[0] * u[1]; F[4] += g[1] * u[1]; F[5] += g[2] * u[1];
        F[6] += g[0] * u[2]; F[7] += g[1] * u[2]; F[8] += g[2] * u[2];
      }
    }
    else
    { /* FE/ElementAx.C:110-136 */
      double v = 0.;
      for (int k = 0; k < 4; k++)
      {
        const double *u = DU + 3 * (long)cn[k];
        const double *g = dN + 2 * k;
        F[0] += g[0] * u[0]; F[1] += g[1] * u[0];
        F[3] += g[0] * u[1]; F[4] += g[1] * u[1];
        v += Q4_N[pt][k] * u[0];
      }
      double R = o_radius(m, e, pt);
      F[8] += v / R;
    }
    double *dEps = m->dEps + 9 * ip, *Rr = m->R + 9 * ip, *S = m->S + 9 * ip;
    if (dnlo_polar_lnU(F, dEps, Rr, mm) != O_OK) return O_E_POLAR;
    for (int k = 0; k < 9; k++) m->Eps[9 * ip + k] += dEps[k];
    pressure += (dEps[0] + dEps[4] + dEps[8]);
    double dev[9];
    t2_deviator(dEps, dev);
    double g2 = 2.0 * G;
    for (int k = 0; k < 9; k++) S[k] += g2 * dev[k];
    if (m->law == 2)
    {
      int st = o_stress_jc(m, ip, S, G);
      if (st) return st;
    }
    else if (m->yield)
    {
      double evp = m->evp[ip];
      double Sigv0 = m->A + m->B * o_pow(mm, evp + O_SQRT23 * 0.0, m->n);
      double Strial = t2_J2(S);
      if (Strial > Sigv0)
      {
        double h = m->n * m->B * o_pow(mm, evp + O_SQRT23 * 0.0, (m->n - 1.));
        double gamma = (t2_norme(S) - O_SQRT23 * Sigv0) / (2.0 * G * (1.0 + (h) / (3.0 * G)));
        double dgamma, Sigv, dfun, fun;
        int run = 1, iteration = 0;
        while (run)
        {
          double ev = evp + O_SQRT23 * gamma;
          Sigv = m->A + m->B * o_pow(mm, ev, m->n);
          h = m->n * m->B * o_pow(mm, ev, (m->n - 1.));
          dfun = O_SQRT32 * 2.0 * G + O_SQRT23 * h;
          fun = dfun * gamma - (Strial - Sigv + h * O_SQRT23 * gamma);
          if (O_ABS(fun / Sigv) > m->rr_prec)
          {
            dgamma = -fun / dfun;
            if (dgamma < 0 && O_ABS(dgamma) > gamma) dgamma = -gamma * 0.9;
            gamma += dgamma;
          }
          else
            run = 0;
          iteration++;
          if (iteration > O_RR_MAXIT) return O_E_RR;
        }
        double nrm = t2_norme(S), n1[9];
        for (int k = 0; k < 9; k++) n1[k] = S[k] / nrm;
        m->evp[ip] += O_SQRT23 * gamma;
        double c = 2.0 * G * gamma;
        for (int k = 0; k < 9; k++) S[k] -= c * n1[k];
        for (int k = 0; k < 9; k++) m->EpsPlas[9 * ip + k] += gamma * n1[k];
      }
    }
  }
  pressure /= m->nip;
  for (int pt = 0; pt < m->nip; pt++)
  {
    long ip = e * m->nip + pt;
    m->p[ip] += K * pressure;
    const double *Rr = m->R + 9 * ip;
    t2_rfi(Rr, m->S + 9 * ip);
    t2_rfi(Rr, m->Eps + 9 * ip);
    t2_rfi(Rr, m->EpsPlas + 9 * ip);
    static const double I9[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int k = 0; k < 9; k++) m->Sig[9 * ip + k] = m->S[9 * ip + k] + m->p[ip] * I9[k];
  }
  return O_OK;
}

/* Element::computeInternalForces, FE/Element.C:164-214 */
static int o_element_forces(dnlo *m, long e, double *Fe)
{
  int st = o_compute_stress(m, e);
  if (st) return st;
  const int dim = m->dim, nen = m->nen;
  for (int k = 0; k < nen * dim; k++) Fe[k] = 0.0;
  for (int pt = 0; pt < m->nip; pt++)
  {
    long ip = e * m->nip + pt;
    const double *dN = m->dN + ip * nen * dim, *Sig = m->Sig + 9 * ip;
    double WxdJ = 1.0 * m->detJ[ip];
    double R = 0.0;
    if (dim == 2)
    {
      R = o_radius(m, e, pt);
      WxdJ *= O_DEUXPI * R;
    }
    for (int I = 0; I < nen; I++)
      for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) Fe[I * dim + i] -= dN[I * dim + j] * Sig[3 * j + i] * WxdJ;
    if (dim == 2)
      for (int I = 0; I < nen; I++) Fe[I * dim] -= (Q4_N[pt][I] * Sig[8] / R) * WxdJ;
  }
  return O_OK;
}

/* Domain::computeInternalMatrices over all elements (FE/Domain.C:98-127) */
int dnlo_compute_internal_matrices(dnlo *m)
{
  int st = O_OK;
#pragma omp parallel for schedule(static) reduction(| : st)
  for (long e = 0; e < m->ne; e++) st |= o_internal_matrices(m, e);
  return st;
}

/* Domain::computeInternalForces, assembly as case 0 with one thread
 * (FE/Domain.C:160-193): Fint = 0, then element by element in index order,
 * local node by local node.  Element work runs in parallel into a per-element
 * buffer; the assembly stays sequential, so the sums are those of the
 * single-threaded reference. */
static int o_forces_element_order(dnlo *m)
{
  const int dim = m->dim, nen = m->nen;
#pragma omp parallel for schedule(static)
  for (long e = 0; e < m->ne; e++) m->Fe_status[e] = o_element_forces(m, e, m->Fe + e * nen * dim);
  for (long e = 0; e < m->ne; e++)
    if (m->Fe_status[e]) return m->Fe_status[e];
  for (long k = 0; k < m->nn * dim; k++) m->fint[k] = 0.0;
  for (long e = 0; e < m->ne; e++)
  {
    const int *cn = m->conn + e * nen;
    const double *Fe = m->Fe + e * nen * dim;
    for (int i = 0; i < nen; i++)
    {
      long glob = (long)cn[i] * dim;
      for (int j = 0; j < dim; j++) m->fint[glob + j] += Fe[i * dim + j];
    }
  }
  return O_OK;
}

/* ---- the reference's multi-threaded assembly strategies ----
 * Jobs::mapInitial (FE/Jobs.C:129-165): the elements sorted by
 * compareElementJobsSort (FE/Jobs.C:104-125: centroid = sum of the node
 * coordinates / n; an element goes after another when its centroid y is
 * greater, or y is equal and its x is not smaller) with List::sort's Shell
 * sort (BT/List.h:1346-1372: gaps h = 3h + 1, insertion while the comparator
 * says "after"), then cut into equal shares: sorted element i goes to job
 * min(floor((i + 1) / (nb / np)), np - 1).  Done once, on the coordinates of
 * the first call, as the reference's `init` flag does. */
static int o_jobs_after(const double *c1, const double *c2)
{
  if (c1[1] < c2[1]) return 0;
  if (c1[1] > c2[1]) return 1;
  if (c1[0] < c2[0]) return 0;
  return 1;
}

static void o_map_initial(dnlo *m)
{
  const long nb = m->ne;
  const int np = m->nthreads, nen = m->nen;
  double *cen = o_alloc((size_t)nb * 3);
  int *lst = (int *)malloc(sizeof(int) * (nb ? nb : 1));
  for (long e = 0; e < nb; e++)
  {
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nen; i++)
      for (int c = 0; c < 3; c++) v[c] += m->x[3 * (long)m->conn[e * nen + i] + c];
    for (int c = 0; c < 3; c++) cen[3 * e + c] = v[c] / nen;
    lst[e] = (int)e;
  }
  long h = 1;
  while (h <= nb) h = 3 * h + 1;
  while (h != 1)
  {
    h = h / 3;
    for (long i = h + 1; i <= nb; i++)
    {
      const int v = lst[i - 1];
      long j = i;
      while (o_jobs_after(cen + 3 * (long)lst[j - h - 1], cen + 3 * (long)v))
      {
        lst[j - 1] = lst[j - h - 1];
        j -= h;
        if (j <= h) break;
      }
      lst[j - 1] = v;
    }
  }
  const double nppr = ((double)nb) / np;
  long *cnt = (long *)calloc(np, sizeof(long));
  int *job = (int *)malloc(sizeof(int) * (nb ? nb : 1));
  for (long i = 0; i < nb; i++)
  {
    const double q = (i + 1) / nppr;
    job[i] = q <= np - 1 ? (int)(long)q : np - 1;
    cnt[job[i]]++;
  }
  free(m->job_el);
  free(m->job_off);
  m->job_el = (int *)malloc(sizeof(int) * (nb ? nb : 1));
  m->job_off = (long *)calloc(np + 1, sizeof(long));
  for (int p = 0; p < np; p++) m->job_off[p + 1] = m->job_off[p] + cnt[p];
  long *fill = (long *)calloc(np, sizeof(long));
  for (long i = 0; i < nb; i++) m->job_el[m->job_off[job[i]] + fill[job[i]]++] = lst[i];
  free(fill);
  free(cnt);
  free(job);
  free(lst);
  free(cen);
  m->mapped = 1;
}

/* per-thread partial Fint of one element's forces: reverse local-node order
 * as Vector::gatherFrom (LA/Vector.C:828-889, cases 5 and 6) or forward as
 * the explicit loops of cases 2, 3, 4 and 7 */
static void o_add_element(const dnlo *m, long e, const double *Fe, double *f, int reverse)
{
  const int dim = m->dim, nen = m->nen;
  const int *cn = m->conn + e * nen;
  for (int k = 0; k < nen; k++)
  {
    const int i = reverse ? nen - 1 - k : k;
    const long glob = (long)cn[i] * dim;
    for (int j = 0; j < dim; j++) f[glob + j] += Fe[i * dim + j];
  }
}

static int o_forces_strategy(dnlo *m)
{
  const long noddl = m->nn * m->dim;
  const int np = m->nthreads, st = m->strategy, nd = m->nen * m->dim;
  if (st != 2 && !m->mapped) o_map_initial(m);
  if (!m->part_fint) m->part_fint = o_alloc((size_t)np * noddl);
  int status = O_OK;
  long next = 0; /* case 2: the shared position in the element list */
  if (st == 3)
    for (long k = 0; k < noddl; k++) m->fint[k] = 0.0; /* Fint = 0.0, FE/Domain.C:146 */
#pragma omp parallel num_threads(np) reduction(| : status)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double *f = m->part_fint + (size_t)t * noddl;
    for (long k = 0; k < noddl; k++) f[k] = 0.0;
    if (st == 2)
    {
      /* FE/Domain.C:241-316: each thread takes the next element of the list in a critical section */
      for (;;)
      {
        long e;
#pragma omp critical(dnlo_case2)
        e = next < m->ne ? next++ : -1;
        if (e < 0) break;
        double *Fe = m->Fe + e * nd;
        const int s = o_element_forces(m, e, Fe);
        status |= s;
        if (!s) o_add_element(m, e, Fe, f, 0);
      }
    }
    else
    {
      /* cases 3, 6, 7 (FE/Domain.C:416-496, 503-594, 654-729): the thread's job in sorted order */
      for (long k = m->job_off[t]; k < m->job_off[t + 1]; k++)
      {
        const long e = m->job_el[k];
        double *Fe = m->Fe + e * nd;
        const int s = o_element_forces(m, e, Fe);
        status |= s;
        if (!s) o_add_element(m, e, Fe, f, st == 6);
      }
    }
    if (st == 3)
    {
      /* Fint += FintLocal in a critical section: the order of arrival */
#pragma omp critical(dnlo_case3)
      for (long k = 0; k < noddl; k++) m->fint[k] += f[k];
    }
  }
  if (status) return status;
  if (st != 3)
  {
    /* the fixed-order reduction, FE/Domain.C:309-316 */
#pragma omp parallel for schedule(static) num_threads(np)
    for (long i = 0; i < noddl; i++)
    {
      double val = 0;
      for (int j = 0; j < np; j++) val += m->part_fint[(size_t)j * noddl + i];
      m->fint[i] = val;
    }
  }
  return O_OK;
}

/* Domain::computeInternalForces: the oracle's order (strategy 0) or one of
 * the reference's multi-threaded strategies */
int dnlo_compute_internal_forces(dnlo *m)
{
  if (m->strategy == 0) return o_forces_element_order(m);
  return o_forces_strategy(m);
}

/* the strategy of Domain::computeInternalForces (the reference's environment
 * variable computeInternalForces, FE/Domain.C:150-153): 0 = the oracle's
 * order, or the reference's case 2, 3, 6 or 7 with `threads` threads (the
 * deterministic ones, 6 and 7, sum the same way for a given thread count).
 * Returns 0, or -1 for another value. */
int dnlo_set_strategy(dnlo *m, int strategy, int threads)
{
  if (strategy != 0 && strategy != 2 && strategy != 3 && strategy != 6 && strategy != 7) return -1;
  if (threads < 1) threads = 1;
  if (threads != m->nthreads)
  {
    free(m->part_fint);
    m->part_fint = NULL;
    m->mapped = 0;
  }
  m->strategy = strategy;
  m->nthreads = threads;
  return 0;
}

/* the jobs of the last Jobs::mapInitial: element of each position, job offsets */
long dnlo_jobs(const dnlo *m, int *elements, long *offsets)
{
  if (!m->mapped) return 0;
  if (elements) memcpy(elements, m->job_el, sizeof(int) * m->ne);
  if (offsets) memcpy(offsets, m->job_off, sizeof(long) * (m->nthreads + 1));
  return m->nthreads;
}

/* Domain::computeMassMatrix + Element::computeMassMatrix (FE/Domain.C:948-1005,
 * FE/Element.C:411-454) */
static void o_mass(dnlo *m)
{
  if (m->mass_done) return;
  for (long i = 0; i < m->nn; i++) m->mass[i] = 0.;
  for (long e = 0; e < m->ne; e++)
  {
    const int *cn = m->conn + e * m->nen;
    double Me[8] = {0};
    for (int pt = 0; pt < m->nip; pt++)
    {
      long ip = e * m->nip + pt;
      double WxdJ = 1.0 * m->detJ[ip];
      if (m->dim == 2)
      {
        double R = o_radius(m, e, pt);
        WxdJ *= 2. * O_PI * R;
      }
      for (int i = 0; i < m->nen; i++)
        Me[i] += m->rho * (m->dim == 3 ? H8_N[pt][i] : Q4_N[pt][i]) * WxdJ;
    }
    for (int i = 0; i < m->nen; i++) m->mass[cn[i]] += Me[i];
  }
  m->mass_done = 1;
}

/* Element::getSoundSpeed (FE/Element.C:607-621); ro is the mean of the
 * nodal densities, all equal to the material density */
static double o_sound_speed(const dnlo *m)
{
  double ro = 0.;
  for (int i = 0; i < m->nen; i++) ro += m->rho;
  ro /= m->nen;
  return sqrt((m->E * (1 - m->nu)) / (ro * (1 + m->nu) * (1 - 2 * m->nu)));
}

/* Characteristic length.
 * quad4 axisymmetric: ElQua4nAx::getLength, FE/ElQua4nAx.C:131-143.
 * hex8: build-defined (the reference's ElHex8n3D::getLength exits,
 * FE/ElHex8n3D.C:188-193): l = V / A_max, V = sum over the 8 Gauss points of
 * detJ (weights 1, current configuration), A_max = the largest of the six
 * face areas 0.5 |(x_c - x_a) x (x_d - x_b)| over the faces (a,b,c,d) of
 * FE/ElHex8n3D.C faces table {0,1,5,4},{1,2,6,5},{2,3,7,6},{3,0,4,7},
 * {7,4,5,6},{2,1,0,3}. */
static const int o_hex_faces[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                      {3, 0, 4, 7}, {7, 4, 5, 6}, {2, 1, 0, 3}};
double dnlo_element_length(const dnlo *m, long e)
{
  const int *cn = m->conn + e * m->nen;
  if (m->dim == 2)
  {
    const double *x0 = m->x + 3 * (long)cn[0], *x1 = m->x + 3 * (long)cn[1], *x2 = m->x + 3 * (long)cn[2],
                 *x3 = m->x + 3 * (long)cn[3];
    double x31 = x2[0] - x0[0];
    double x24 = x1[0] - x3[0];
    double y42 = x3[1] - x1[1];
    double y31 = x2[1] - x0[1];
    return ((x31 * y42 + x24 * y31) / sqrt(2. * (O_SQ(x24) + O_SQ(y42) + O_SQ(x31) + O_SQ(y31))));
  }
  double V = 0.0;
  for (int pt = 0; pt < 8; pt++) V += m->detJ[e * 8 + pt];
  double Amax = 0.0;
  for (int f = 0; f < 6; f++)
  {
    const double *xa = m->x + 3 * (long)cn[o_hex_faces[f][0]], *xb = m->x + 3 * (long)cn[o_hex_faces[f][1]],
                 *xc = m->x + 3 * (long)cn[o_hex_faces[f][2]], *xd = m->x + 3 * (long)cn[o_hex_faces[f][3]];
    double d1[3] = {xc[0] - xa[0], xc[1] - xa[1], xc[2] - xa[2]};
    double d2[3] = {xd[0] - xb[0], xd[1] - xb[1], xd[2] - xb[2]};
    double c0 = d1[1] * d2[2] - d1[2] * d2[1];
    double c1 = d1[2] * d2[0] - d1[0] * d2[2];
    double c2 = d1[0] * d2[1] - d1[1] * d2[0];
    double area = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (area > Amax) Amax = area;
  }
  return V / Amax;
}

/* Domain::computeCourantTimeStep, FE/Domain.C:1613-1646 */
double dnlo_courant_crit(dnlo *m)
{
  double Cs = o_sound_speed(m);
  double crit = dnlo_element_length(m, 0) / (Cs);
  for (long i = 1; i < m->ne; i++)
  {
    double time = dnlo_element_length(m, i) / (Cs);
    if (time < crit) crit = time;
  }
  return crit;
}

/* Material::getHookeMatrix (FE/Material.C:270-331): the threedimensional
 * 6x6 for hex8, the axisymetric 4x4 for the axisymmetric quad; row-major.
 * Returns the order. */
int dnlo_hooke(const dnlo *m, double *C)
{
  const double Nu = m->nu, E = m->E;
  const int nc = m->dim == 3 ? 6 : 4;
  for (int k = 0; k < nc * nc; k++) C[k] = 0.0;
  if (nc == 6)
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
    C[4] = Nu / (1. - Nu); C[1] = Nu / (1. - Nu);
    C[12] = Nu / (1. - Nu); C[13] = Nu / (1. - Nu);
    C[3] = Nu / (1. - Nu); C[7] = Nu / (1. - Nu);
    C[10] = (1.0 - 2.0 * Nu) / (2. * (1.0 - Nu));
  }
  const double f = (E * (1.0 - Nu) / ((1.0 + Nu) * (1.0 - 2.0 * Nu)));
  for (int k = 0; k < nc * nc; k++) C[k] *= f; /* Matrix::operator*=, LA/Matrix.C:547-554 */
  return nc;
}

/* Element3D::computeElasticStiffnessMatrix (FE/Element3D.C:242-333) and
 * Element2D::computeElasticStiffnessMatrix for the axisymmetric family
 * (FE/Element2D.C:129-222), full integration, from the point dN and detJ of
 * the last computeInternalMatrices.  K is [nd][nd] row-major (the reference
 * stores column-major; the entries and their sums are the same). */
void dnlo_element_stiffness(const dnlo *m, long e, double *K)
{
  double C[36];
  const int nc = dnlo_hooke(m, C);
  const int nen = m->nen, dim = m->dim, nd = nen * dim;
#define Cm(r, c) C[(r) * nc + (c)]
#define Km(r, c) K[(r) * nd + (c)]
  for (int k = 0; k < nd * nd; k++) K[k] = 0.0;
  double CB[6][24];
  for (int pt = 0; pt < m->nip; pt++)
  {
    const long ip = e * m->nip + pt;
    const double *dN = m->dN + ip * nd;
    double W = 1.0 * m->detJ[ip];
    if (dim == 3)
    {
      for (int i = 0; i < 8; i++)
      {
        const int I = 3 * i;
        const double d0 = dN[I], d1 = dN[I + 1], d2 = dN[I + 2];
        for (int r = 0; r < 6; r++)
        {
          CB[r][I] = (Cm(r, 0) * d0 + Cm(r, 3) * d1 + Cm(r, 5) * d2);
          CB[r][I + 1] = (Cm(r, 1) * d1 + Cm(r, 3) * d0 + Cm(r, 4) * d2);
          CB[r][I + 2] = (Cm(r, 2) * d2 + Cm(r, 4) * d1 + Cm(r, 5) * d0);
        }
      }
      for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++)
        {
          const int I = 3 * i, J = 3 * j;
          const double a0 = dN[I], a1 = dN[I + 1], a2 = dN[I + 2];
          for (int c = 0; c < 3; c++)
          {
            Km(I, J + c) += (a0 * CB[0][J + c] + a1 * CB[3][J + c] + a2 * CB[5][J + c]) * W;
          }
          for (int c = 0; c < 3; c++)
          {
            Km(I + 1, J + c) += (a1 * CB[1][J + c] + a0 * CB[3][J + c] + a2 * CB[4][J + c]) * W;
          }
          for (int c = 0; c < 3; c++)
          {
            Km(I + 2, J + c) += (a2 * CB[2][J + c] + a1 * CB[4][J + c] + a0 * CB[5][J + c]) * W;
          }
        }
    }
    else
    {
      const double R = o_radius(m, e, pt);
      W *= 2 * O_PI * R;
      for (int i = 0; i < 4; i++)
      {
        const int I = 2 * i;
        const double d0 = dN[I], d1 = dN[I + 1], Ni = Q4_N[pt][i];
        for (int r = 0; r < 3; r++)
        {
          CB[r][I] = (Cm(r, 0) * d0 + Cm(r, 2) * d1);
          CB[r][I + 1] = (Cm(r, 1) * d1 + Cm(r, 2) * d0);
        }
        CB[3][I] = (Cm(3, 0) * d0 + Cm(3, 2) * d1);
        CB[3][I + 1] = (Cm(3, 1) * d1 + Cm(3, 2) * d0);
        for (int r = 0; r < 4; r++) CB[r][I] += Cm(r, 3) * Ni / R;
      }
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
        {
          const int I = 2 * i, J = 2 * j;
          const double a0 = dN[I], a1 = dN[I + 1], Ni = Q4_N[pt][i];
          Km(I, J) += (a0 * CB[0][J] + a1 * CB[2][J]) * W;
          Km(I, J + 1) += (a0 * CB[0][J + 1] + a1 * CB[2][J + 1]) * W;
          Km(I + 1, J) += (a1 * CB[1][J] + a0 * CB[2][J]) * W;
          Km(I + 1, J + 1) += (a1 * CB[1][J + 1] + a0 * CB[2][J + 1]) * W;
          Km(I, J) += CB[3][J] * Ni / R * W;
          Km(I, J + 1) += CB[3][J + 1] * Ni / R * W;
        }
    }
  }
#undef Cm
#undef Km
}

/* Domain::getMaximumFrequency (FE/Domain.C:816-946), the power iteration on
 * M^-1 K with the assembly order of one thread: element by element in index
 * order into a zeroed vector (the per-thread vectors of 843-893 summed for
 * threads = 1).  The start vector, ev(0) = 1, ev(i) = ev(i-1) - 2/(noddl-1),
 * is built when fmax0 is still 0 (FE/Domain.C:849-856); later calls start
 * from the last iterate and compare the first fmax with the last one.  Ke x
 * is LAPACK's reference dgemv ('N', beta 0: y = 0, then y += x_j K(:,j) for
 * j in order; LA/Matrix.C:875-900), ev /= fmax multiplies by 1/fmax
 * (LA/Vector.C:652-659).  Returns O_OK and sqrt(fmax), or O_E_PULSE when
 * iteration > 1000 (FE/Domain.C:936-939, limits FE/DefaultValues.h:35-36). */
int dnlo_max_frequency(dnlo *m, double *out)
{
  const int nen = m->nen, dim = m->dim, nd = nen * dim;
  const long noddl = (long)dim * m->nn;
  if (!m->Ke) m->Ke = o_alloc((size_t)m->ne * nd * nd);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < m->ne; e++) dnlo_element_stiffness(m, e, m->Ke + e * nd * nd);
  if (m->fmax0 == 0 || !m->ev)
  {
    if (!m->ev) m->ev = o_alloc(noddl);
    m->ev[0] = 1.0;
    for (long i = 1; i < noddl; i++) m->ev[i] = m->ev[i - 1] - 2. / (noddl - 1);
  }
  double *ev0 = o_alloc(noddl), *acc = o_alloc(noddl);
  double fmax = 0.0;
  long iteration = 0;
  int ok = 0, st = O_OK;
  while (!ok)
  {
    iteration++;
    memcpy(ev0, m->ev, sizeof(double) * noddl);
    for (long k = 0; k < noddl; k++) acc[k] = 0.;
    for (long e = 0; e < m->ne; e++)
    {
      const int *cn = m->conn + e * nen;
      const double *K = m->Ke + e * nd * nd;
      double loc[24], y[24];
      for (int J = 0; J < nd; J++) loc[J] = ev0[(long)cn[J / dim] * dim + J % dim];
      for (int I = 0; I < nd; I++) y[I] = 0.0;
      for (int J = 0; J < nd; J++)
      {
        const double t = 1.0 * loc[J];
        for (int I = 0; I < nd; I++) y[I] = y[I] + t * K[I * nd + J];
      }
      for (int I = 0; I < nd; I++) acc[(long)cn[I / dim] * dim + I % dim] += y[I];
    }
    fmax = 0.0;
    for (long i = 0; i < noddl; i++)
    {
      double val = 0.0;
      val += acc[i];
      val /= m->mass[i / dim];
      if (fabs(val) > fmax) fmax = fabs(val);
      m->ev[i] = val;
    }
    const double inv = 1.0 / fmax;
    for (long i = 0; i < noddl; i++) m->ev[i] *= inv;
    const double convergence = fabs(fmax - m->fmax0) / fmax;
    m->pulse_conv = convergence;
    if (convergence < 1e-4) ok = 1;
    m->fmax0 = fmax;
    if (iteration > 1000)
    {
      st = O_E_PULSE;
      break;
    }
  }
  free(ev0);
  free(acc);
  m->pulse_iters = iteration;
  if (out) *out = sqrt(fmax);
  return st;
}

void dnlo_set_ts_method(dnlo *m, int method) { m->ts_method = method; }
void dnlo_compute_mass(dnlo *m) { o_mass(m); }
void dnlo_get_plot_row(const dnlo *m, double *out) { memcpy(out, m->plot_row, sizeof m->plot_row); }
long dnlo_pulse_iterations(const dnlo *m) { return m->pulse_iters; }
double dnlo_pulse_fmax0(const dnlo *m) { return m->fmax0; }

/* Solver::computeTimeStep, FE/Solver.C:220-267 (Courant or Pulse) */
static int o_compute_time_step(dnlo *m, int force)
{
  if (force || (m->increment % m->freq == 0))
  {
    double fmax;
    if (m->ts_method == 1)
    {
      int st = dnlo_max_frequency(m, &fmax);
      if (st) return st;
    }
    else
      fmax = 2.0 / dnlo_courant_crit(m);
    m->dt = m->safety * m->omegaS / fmax;
    m->lastTimeStepComputationDate = m->timeStepComputationDate;
    m->lastTimeStep = m->computedTimeStep;
    m->timeStepComputationDate = m->t;
    m->computedTimeStep = m->dt;
  }
  else if (m->timeStepComputationDate != m->lastTimeStepComputationDate)
  {
    double dx = (m->t - m->lastTimeStepComputationDate) /
                (m->timeStepComputationDate - m->lastTimeStepComputationDate);
    m->dt = ((dx) * (m->computedTimeStep) + (1.0 - (dx)) * (m->lastTimeStep)); /* LINEARdbl, BT/Math.h:91 */
  }
  return O_OK;
}

/* boundary conditions on one node: BoundaryRestrain::applyConstant
 * (FE/Boundary.C:636-666) and constant BoundarySpeed::applyConstant
 * (FE/Boundary.C:478-524) */
static void o_apply_constant(dnlo *m, long i)
{
  double *Cv = m->v[m->cur] + 3 * i, *Ca = m->a[m->cur] + 3 * i, *Cd = m->du[m->cur] + 3 * i;
  double *Nv = m->v[1 - m->cur] + 3 * i, *Na = m->a[1 - m->cur] + 3 * i, *Nd = m->du[1 - m->cur] + 3 * i;
  unsigned char s = m->bcs[i];
  for (int c = 0; c < 3; c++)
    if (s & (1 << c))
    {
      double sp = m->bcs_v[3 * i + c];
      Ca[c] = 0.; Na[c] = 0.;
      Cv[c] = sp; Nv[c] = sp;
      Cd[c] = Cv[c] * m->dt;
      Nd[c] = Nv[c] * m->dt;
    }
  unsigned char b = m->bc[i];
  for (int c = 0; c < 3; c++)
    if (b & (1 << c))
    {
      Cd[c] = 0.; Nd[c] = 0.;
      Cv[c] = 0.; Nv[c] = 0.;
      Ca[c] = 0.; Na[c] = 0.;
    }
}

/* ExplicitSolverCH::computePredictions, FE/ExplicitSolverCH.C:387-431 */
static void o_predictions(dnlo *m)
{
  const int C = m->cur, N = 1 - m->cur;
  const double s1 = m->dt * (0.5 - m->beta), s2 = m->dt * (1 - m->gamma);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < m->nn; i++)
  {
    for (int c = 0; c < 3; c++)
    {
      long k = 3 * i + c;
      double du = m->a[C][k];
      du *= s1;
      du += m->v[C][k];
      du *= m->dt;
      m->du[N][k] = du;
      m->d[N][k] = m->d[C][k] + du;
      double v = m->a[C][k];
      v *= s2;
      v += m->v[C][k];
      m->v[N][k] = v;
      m->a[N][k] = 0.0;
    }
    if (m->bc[i] | m->bcs[i]) o_apply_constant(m, i);
  }
}

/* ExplicitSolverCH::explicitSolve + updateGrid, FE/ExplicitSolverCH.C:490-567 */
static void o_explicit_solve(dnlo *m)
{
  const int C = m->cur, N = 1 - m->cur, dim = m->dim;
  const double am = m->alphaM, s3 = m->dt * m->gamma, s4 = O_SQ(m->dt) * m->beta;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < m->nn; i++)
  {
    for (int j = 0; j < dim; j++) m->a[N][3 * i + j] = m->fint[i * dim + j] / m->mass[i];
    for (int c = 0; c < 3; c++)
    {
      long k = 3 * i + c;
      double a = m->a[N][k];
      a -= am * m->a[C][k];
      a /= (1.0 - am);
      m->a[N][k] = a;
      m->v[N][k] += s3 * a;
      m->du[N][k] += s4 * a;
    }
    if (m->bc[i] | m->bcs[i]) o_apply_constant(m, i);
    for (int c = 0; c < 3; c++)
    {
      long k = 3 * i + c;
      m->d[N][k] = m->d[C][k] + m->du[N][k];
      m->x[k] += m->du[N][k]; /* lagMotion::updateNode, FE/lagMotion.C:52-57 */
    }
  }
}

/* Domain::getTotalKineticEnergy, FE/Domain.C:1330-1344 */
static double o_kinetic_energy(const dnlo *m)
{
  double energy = 0.;
  const double *V = m->v[1 - m->cur];
  for (long i = 0; i < m->nn; i++)
    energy += .5 * m->mass[i] * (O_SQ(V[3 * i]) + O_SQ(V[3 * i + 1]) + O_SQ(V[3 * i + 2]));
  return energy;
}

/* Domain::initSolve tail: computeInternalMatrices(True), FE/Domain.C:1219-1223 */
int dnlo_init(dnlo *m)
{
  if (m->alphaM == 0 && m->beta == 0) dnlo_set_solver(m, m->rho_b, m->safety, m->freq, m->endTime);
  m->status = dnlo_compute_internal_matrices(m);
  return m->status;
}

/* The time loop in parts, so that a caller can supply Domain::Fint from
 * outside (the drop-in test: the reference's host loop with the GPU element
 * pass in place of Domain::computeInternalForces).  Each *_pre runs up to the
 * computeInternalForces call of the reference and each *_post from there on;
 * fint = NULL takes the oracle's own Fint. */

/* ExplicitSolverCH::solve preamble up to computeInternalForces,
 * FE/ExplicitSolverCH.C:115-157.  Returns 1 when the loop runs, 0 when
 * betweenTimeBounds is false, <0 on error. */
int dnlo_begin_pre(dnlo *m, double upTime)
{
  m->upTime = upTime;
  if (!((m->t >= 0.0) && (m->t < (m->endTime < upTime ? m->endTime : upTime)))) return 0;
  int st = dnlo_compute_internal_matrices(m);
  if (st) return -st;
  o_mass(m);
  return 1;
}

/* the preamble's computeTimeStep(True), FE/ExplicitSolverCH.C:172 */
int dnlo_begin_post(dnlo *m, const double *fint)
{
  if (fint) memcpy(m->fint, fint, sizeof(double) * m->nn * m->dim);
  int st = o_compute_time_step(m, 1);
  if (st) return -st;
  return 1;
}

/* ExplicitSolverCH::solve preamble, FE/ExplicitSolverCH.C:115-172 */
int dnlo_begin(dnlo *m, double upTime)
{
  int r = dnlo_begin_pre(m, upTime);
  if (r <= 0) return r;
  int st = dnlo_compute_internal_forces(m);
  if (st) return -st;
  return dnlo_begin_post(m, NULL);
}

/* initStep and computePredictions, FE/ExplicitSolverCH.C:181-212, 387-447 */
void dnlo_step_pre(dnlo *m)
{
  if (m->t + m->dt > m->upTime) m->dt = m->upTime - m->t;
  m->increment++;
  o_predictions(m);
}

/* explicitSolve, updateGrid, endStep and, while the loop runs,
 * computeInternalMatrices + computeTimeStep, FE/ExplicitSolverCH.C:222-274.
 * Returns 1 while the loop continues, 0 when it ends, <0 on error. */
int dnlo_step_post(dnlo *m, const double *fint)
{
  if (fint) memcpy(m->fint, fint, sizeof(double) * m->nn * m->dim);
  o_explicit_solve(m);
  /* endStep, FE/ExplicitSolverCH.C:450-468 */
  m->t += m->dt;
  if (m->ke_on) m->ke = o_kinetic_energy(m);
  m->plot_row[0] = m->t;
  m->plot_row[1] = m->dt;
  m->plot_row[2] = m->computedTimeStep;
  m->plot_row[3] = m->ke;
  m->cur = 1 - m->cur; /* transfertNodalField, FE/Node.C:219-229 */
  int run = (m->t >= 0.0) && (m->t < (m->endTime < m->upTime ? m->endTime : m->upTime));
  if (run)
  {
    int st = dnlo_compute_internal_matrices(m);
    if (st) return -st;
    st = o_compute_time_step(m, 0);
    if (st) return -st;
  }
  return run;
}

/* one pass of the while(runStep) loop, FE/ExplicitSolverCH.C:179-274.
 * Returns 1 while the loop continues, 0 when it ends, <0 on error. */
int dnlo_step(dnlo *m)
{
  dnlo_step_pre(m);
  const double t0 = o_wtime();
  int st = dnlo_compute_internal_forces(m);
  m->forces_seconds += o_wtime() - t0;
  if (st) return -st;
  return dnlo_step_post(m, NULL);
}

/* the accumulated time of the internal-force passes of dnlo_step, and its reset */
double dnlo_forces_seconds(const dnlo *m) { return m->forces_seconds; }
void dnlo_reset_forces_seconds(dnlo *m) { m->forces_seconds = 0.0; }

/* ExplicitSolverCH::solve(upTime) in full; returns the number of steps or <0 */
long dnlo_solve(dnlo *m, double upTime)
{
  int r = dnlo_begin(m, upTime);
  if (r <= 0) return r;
  long n = 0;
  while (1)
  {
    r = dnlo_step(m);
    if (r < 0) return r;
    n++;
    if (!r) break;
  }
  return n;
}

/* ---- Node::getNodalValue of the element fields (FE/Node.C:350-418) ----
 * get_IntegrationValue_atNode (FE/Element.C:1180-1198): over the node's
 * elements in the order Domain::initSolve attached them (element order,
 * FE/Domain.C:1090-1107), VAR += integrationPointsToNode(loc)(pt) * point
 * value, then VAR / (number of elements); Node::getNodalValue adds that tensor
 * once per element and divides by the count again.  Tables:
 * ElHex8n3D (FE/ElHex8n3D.C:33-72), ElQua4nAx (FE/ElQua4nAx.C:38-48).
 * field: 13 stress (Sig), 14 strain (Eps), 15 strainInc (dEps), 16
 * plasticStrain (EpsPlas); comp 0 the equivalent (misesEquivalent for the
 * stress, BT/Math.h:87-88; equivalent() for a strain, LA/Tensor2.h:359-372)
 * or 10 i + j for entry (i, j). */
double dnlo_nodal_value(const dnlo *m, long node, int field, int comp)
{
#define O_C5P ((5.0 + 3.0 * sqrt(3.0)) / 4.0)
#define O_C5M ((5.0 - 3.0 * sqrt(3.0)) / 4.0)
#define O_R3P ((sqrt(3.0) + 1.0) / 4.0)
#define O_R3M ((sqrt(3.0) - 1.0) / 4.0)
  const double hex[8][8] = {
      {O_C5P, -O_R3P, O_R3M, -O_R3P, -O_R3P, O_R3M, O_C5M, O_R3M},
      {-O_R3P, O_C5P, -O_R3P, O_R3M, O_R3M, -O_R3P, O_R3M, O_C5M},
      {O_R3M, -O_R3P, O_C5P, -O_R3P, O_C5M, O_R3M, -O_R3P, O_R3M},
      {-O_R3P, O_R3M, -O_R3P, O_C5P, O_R3M, O_C5M, O_R3M, -O_R3P},
      {-O_R3P, O_R3M, O_C5M, O_R3M, O_C5P, -O_R3P, O_R3M, -O_R3P},
      {O_R3M, -O_R3P, O_R3M, O_C5M, -O_R3P, O_C5P, -O_R3P, O_R3M},
      {O_C5M, O_R3M, -O_R3P, O_R3M, O_R3M, -O_R3P, O_C5P, -O_R3P},
      {O_R3M, O_C5M, O_R3M, -O_R3P, -O_R3P, O_R3M, -O_R3P, O_C5P}};
  const double qp = 1.0 + sqrt(3.0) / 2.0, qm = 1.0 - sqrt(3.0) / 2.0, qh = 1.0 / 2.0;
  const double quad[4][4] = {{qp, -qh, qm, -qh}, {-qh, qp, -qh, qm}, {qm, -qh, qp, -qh}, {-qh, qm, -qh, qp}};
#undef O_C5P
#undef O_C5M
#undef O_R3P
#undef O_R3M
  const double *src = field == 13 ? m->Sig : field == 14 ? m->Eps : field == 15 ? m->dEps : m->EpsPlas;
  double V[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  long nb = 0;
  for (long e = 0; e < m->ne; e++)
  {
    int loc = -1;
    for (int i = 0; i < m->nen; i++)
      if (m->conn[e * m->nen + i] == node) loc = i;
    if (loc < 0) continue;
    for (int pt = 0; pt < m->nip; pt++)
    {
      const double c = m->dim == 3 ? hex[loc][pt] : quad[loc][pt];
      const double *T = src + 9 * (e * m->nip + pt);
      for (int k = 0; k < 9; k++) V[k] += c * T[k];
    }
    nb++;
  }
  if (nb == 0) return 0.0;
  double V1[9];
  for (int k = 0; k < 9; k++)
  {
    const double v = V[k] / nb;
    double sum = 0.0;
    for (long j = 0; j < nb; j++) sum += v;
    V1[k] = sum / nb;
  }
  if (comp == 0)
  {
    if (field == 13)
      return sqrt((O_SQ(V1[0] - V1[4]) + O_SQ(V1[0] - V1[8]) + O_SQ(V1[4] - V1[8]) +
                   6.0 * (O_SQ(V1[1]) + O_SQ(V1[2]) + O_SQ(V1[5]))) / 2.0);
    return sqrt(2. * (V1[0] * V1[0] + V1[1] * V1[1] + V1[2] * V1[2] + V1[3] * V1[3] + V1[4] * V1[4] +
                      V1[5] * V1[5] + V1[6] * V1[6] + V1[7] * V1[7] + V1[8] * V1[8]) / 3.);
  }
  return V1[3 * (comp / 10 - 1) + (comp % 10 - 1)];
}

/* ---- state access ---- */
enum { F_X = 0, F_V, F_A, F_DU, F_D, F_FINT, F_MASS, F_S, F_EVP, F_P, F_EPS, F_EPSPLAS, F_SIG, F_X0, F_FE, F_TEMP, F_DEPS };
long dnlo_get(const dnlo *m, int field, double *out)
{
  long n = 0;
  const double *src = NULL;
  switch (field)
  {
  case F_X: src = m->x; n = m->nn * 3; break;
  case F_X0: src = m->x0; n = m->nn * 3; break;
  case F_V: src = m->v[m->cur]; n = m->nn * 3; break;
  case F_A: src = m->a[m->cur]; n = m->nn * 3; break;
  case F_DU: src = m->du[m->cur]; n = m->nn * 3; break;
  case F_D: src = m->d[m->cur]; n = m->nn * 3; break;
  case F_FINT: src = m->fint; n = m->nn * m->dim; break;
  case F_MASS: src = m->mass; n = m->nn; break;
  case F_S: src = m->S; n = m->ne * m->nip * 9; break;
  case F_EVP: src = m->evp; n = m->ne * m->nip; break;
  case F_P: src = m->p; n = m->ne * m->nip; break;
  case F_EPS: src = m->Eps; n = m->ne * m->nip * 9; break;
  case F_EPSPLAS: src = m->EpsPlas; n = m->ne * m->nip * 9; break;
  case F_SIG: src = m->Sig; n = m->ne * m->nip * 9; break;
  case F_FE: src = m->Fe; n = m->ne * m->nen * m->dim; break;
  case F_TEMP: src = m->T; n = m->T ? m->ne * m->nip : 0; break;
  case F_DEPS: src = m->dEps; n = m->ne * m->nip * 9; break;
  default: return -1;
  }
  if (out && n) memcpy(out, src, sizeof(double) * n);
  return n;
}

/* overwrite a state field (a restart, FE/Domain.C:1378-1404: coordinates,
 * the Current nodal fields, point state); returns the doubles written or -1 */
long dnlo_set(dnlo *m, int field, const double *in)
{
  long n = 0;
  double *dst = NULL;
  switch (field)
  {
  case F_X: dst = m->x; n = m->nn * 3; break;
  case F_V: dst = m->v[m->cur]; n = m->nn * 3; break;
  case F_A: dst = m->a[m->cur]; n = m->nn * 3; break;
  case F_DU: dst = m->du[m->cur]; n = m->nn * 3; break;
  case F_D: dst = m->d[m->cur]; n = m->nn * 3; break;
  case F_S: dst = m->S; n = m->ne * m->nip * 9; break;
  case F_EVP: dst = m->evp; n = m->ne * m->nip; break;
  case F_P: dst = m->p; n = m->ne * m->nip; break;
  case F_EPS: dst = m->Eps; n = m->ne * m->nip * 9; break;
  case F_EPSPLAS: dst = m->EpsPlas; n = m->ne * m->nip * 9; break;
  case F_SIG: dst = m->Sig; n = m->ne * m->nip * 9; break;
  case F_DEPS: dst = m->dEps; n = m->ne * m->nip * 9; break;
  default: return -1;
  }
  memcpy(dst, in, sizeof(double) * n);
  return n;
}

/* Structure::readData's currentTime (FE/Structure.C:915-931) */
void dnlo_set_time(dnlo *m, double t) { m->t = t; }

/* overwrite the Current displacement increment (for single-element tests) */
void dnlo_set_du(dnlo *m, const double *du) { memcpy(m->du[m->cur], du, sizeof(double) * m->nn * 3); }

void dnlo_get_time(const dnlo *m, double *out)
{
  out[0] = m->t;
  out[1] = m->dt;
  out[2] = (double)m->increment;
  out[3] = m->computedTimeStep;
  out[4] = m->ke;
  out[5] = m->alphaM;
  out[6] = m->beta;
  out[7] = m->gamma;
  out[8] = m->omegaS;
}

void dnlo_free(dnlo *m)
{
  if (!m) return;
  free(m->x); free(m->x0);
  for (int f = 0; f < 2; f++) { free(m->v[f]); free(m->a[f]); free(m->du[f]); free(m->d[f]); }
  free(m->mass); free(m->bc); free(m->bcs); free(m->bcs_v); free(m->conn);
  free(m->ev); free(m->Ke);
  free(m->S); free(m->Eps); free(m->EpsPlas); free(m->dEps); free(m->R); free(m->Sig);
  free(m->p); free(m->evp); free(m->T); free(m->dN); free(m->detJ); free(m->fint); free(m->Fe); free(m->Fe_status);
  free(m->job_el); free(m->job_off); free(m->part_fint);
  free(m);
}
