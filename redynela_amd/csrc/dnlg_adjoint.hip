// Hand-written adjoint (reverse mode) of one increment of the axisymmetric
// ExplicitSolverCH time loop, for gradients of a function of the final nodal
// coordinates with respect to the hardening parameters (A, B, n) of
// sigma_y = A + B evp^n.  There is no reference counterpart; the forward it
// reverses is, per increment k -> k+1:
//   computeTimeStep  (FE/Solver.C:220-267): Courant every freq increments,
//                    LINEARdbl extrapolation of dt in between
//   computeInternalForces (FE/Element.C:164-403): kinematics on x_k with the
//                    Current displacement increment du_k, Hencky increment
//                    ln U and rotation R of F = R U, J2 radial return
//                    (FE/Element.C:248-348), pressure, S = R S' R^T, B^T sigma
//   explicitSolve    (FE/ExplicitSolverCH.C:408-541): predictor, a = Fint/M,
//                    Chung-Hulbert corrector, restraints, x += du
// The reverse uses the derivative of the exact maps: the 2x2 polar
// decomposition in closed form (rotation angle, eigen-decomposition of U with
// the Daleckii-Krein divided differences for ln U) and the converged
// consistency condition of the return, sqrt(3/2) 2G gamma + sigma_y(evp +
// sqrt(2/3) gamma) = sqrt(3/2)|S_trial|, differentiated implicitly.  With the
// Johnson-Cook temperature factor (law 2 without its rate term) sigma_y
// carries theta(T) = 1 - ((T - T0)/(Tm - T0))^m of the point's temperature
// before the increment, and the point heats by chi/(rho cp) sigma_y d(evp)
// after the return: the temperature has an adjoint of its own, and the
// heating adds its dependence on gamma, evp and (A, B, n).  One backward
// increment is four launches: nodal, element, node gather, scalar.
#include "dnlg_elem_f64.hpp"

namespace dnl {

// ---------------------------------------------------------------------------
// nodal phase: predictor / corrector / restraint of increment k+1 reversed
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_nodal(AdjArgs g)
{
  __shared__ double red[256];
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  double dtb = 0.0;
  if (i < g.nn)
  {
    const double dt = g.st1->dt, be = g.beta, ga = g.gamma, am = g.alphaM;
    const double s1 = dt * (0.5 - be), s2 = dt * (1.0 - ga), s3 = dt * ga, s4 = dt * dt * be;
    const double m = g.mass[i];
    const unsigned char b = g.bc[i];
#pragma unroll
    for (int c = 0; c < 2; c++)
    {
      const long long q = 2 * i + c;
      if (b & (1 << c))
      {
        // x' = x; du', v', a' = 0: only x carries its adjoint through
        g.vb[q] = 0.0;
        g.ab[q] = 0.0;
        g.fb[q] = 0.0;
        continue;
      }
      const double Xb = g.xb[q], Ub = g.ub[q], Vb = g.vb[q], Ab = g.ab[q];
      const double a = g.a[3 * i + c], v = g.v[3 * i + c], aa = g.a1[3 * i + c];
      const double dunb = Xb + Ub;
      const double aab = Ab + Vb * s3 + dunb * s4;
      const double s3b = Vb * aa, s4b = dunb * aa;
      const double s2b = Vb * a, s1b = dunb * a * dt;
      dtb += dunb * (a * s1 + v) + s1b * (0.5 - be) + s2b * (1.0 - ga) + s3b * ga + s4b * 2.0 * dt * be;
      g.fb[q] = aab / (1.0 - am) / m;
      g.ab[q] = -am / (1.0 - am) * aab + Vb * s2 + dunb * s1 * dt;
      g.vb[q] = Vb + dunb * dt;
    }
  }
  red[threadIdx.x] = dtb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) g.dt_part[blockIdx.x] = red[0];
}

// theta(T) of the Johnson-Cook law (FE/IsoHardJohnsonCook.h:51-68) as the
// forward evaluates it, and its derivative
__device__ __forceinline__ void adj_theta(const ElemArgs &a, double T, double &tf, double &dtf)
{
  tf = 1.0;
  dtf = 0.0;
  if (a.chi_rho_cp == 0.0 || T < a.T0) return;
  if (T >= a.Tm)
  {
    tf = 0.0;
    return;
  }
  const double ts = (T - a.T0) / (a.Tm - a.T0);
  tf = 1.0 - pow(ts, a.m);
  if (ts > 0.0) dtf = -a.m * pow(ts, a.m - 1.0) / (a.Tm - a.T0);
}

__device__ __forceinline__ double sum4(double v)
{
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  return v;
}

// ---------------------------------------------------------------------------
// element pass of increment k+1 recomputed and reversed: lane = 4 element + point
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_adj_elem(AdjArgs g)
{
  const ElemArgs &ea = g.ea;
  const int tid = threadIdx.x, pt = tid & 3;
  const long long e_raw = ((long long)blockIdx.x * 64 + tid) >> 2;
  const bool valid = e_raw < g.ne;
  const long long e = valid ? e_raw : g.ne - 1;
  const long long nipt = g.ne * 4, ip = e * 4 + pt;
  int nd[4];
  double X[4][2], U[4][2], gs[4][2], Nq[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    nd[k] = ea.conn[e * 4 + k];
    X[k][0] = g.x[3 * (long long)nd[k]];
    X[k][1] = g.x[3 * (long long)nd[k] + 1];
    U[k][0] = g.du[3 * (long long)nd[k]];
    U[k][1] = g.du[3 * (long long)nd[k] + 1];
    quad_dn0(ea, pt, k, gs[k]);
    Nq[k] = quad_n(ea, pt, k);
  }
  // ---- forward recompute ----
  double J[2][2] = {{0, 0}, {0, 0}}, Rr = 0.0, vr = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int d = 0; d < 2; d++) J[c][d] += gs[k][d] * X[k][c];
    Rr += Nq[k] * X[k][0];
    vr += Nq[k] * U[k][0];
  }
  const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double inv[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
  double dN[4][2];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 2; j++) dN[k][j] = gs[k][0] * inv[0][j] + gs[k][1] * inv[1][j];
  double F[2][2] = {{1, 0}, {0, 1}};
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    F[0][0] += dN[k][0] * U[k][0];
    F[0][1] += dN[k][1] * U[k][0];
    F[1][0] += dN[k][0] * U[k][1];
    F[1][1] += dN[k][1] * U[k][1];
  }
  const double F22 = 1.0 + vr / Rr;
  // F2 = R2 U2, R2 = [[c, -s], [s, c]]
  const double pa = F[0][0] + F[1][1], pb_ = F[1][0] - F[0][1];
  const double rr2 = pa * pa + pb_ * pb_, rr = sqrt(rr2);
  const double cr = pa / rr, sr = pb_ / rr;
  const double R2[2][2] = {{cr, -sr}, {sr, cr}};
  double U2[2][2];
  U2[0][0] = cr * F[0][0] + sr * F[1][0];
  U2[1][1] = -sr * F[0][1] + cr * F[1][1];
  U2[0][1] = U2[1][0] = 0.5 * ((cr * F[0][1] + sr * F[1][1]) + (-sr * F[0][0] + cr * F[1][0]));
  // eigenpairs of U2: la (vector q1 = (cq, sq)) and lb (vector q2 = (-sq, cq))
  const double mu = 0.5 * (U2[0][0] + U2[1][1]), df = 0.5 * (U2[0][0] - U2[1][1]);
  const double rad = sqrt(df * df + U2[0][1] * U2[0][1]);
  const double la = mu + rad, lb = mu - rad;
  const double psi = 0.5 * atan2(U2[0][1], df);
  const double cq = cos(psi), sq = sin(psi);
  const double lnla = log(la), lnlb = log(lb);
  double L2[2][2];
  L2[0][0] = cq * cq * lnla + sq * sq * lnlb;
  L2[1][1] = sq * sq * lnla + cq * cq * lnlb;
  L2[0][1] = L2[1][0] = cq * sq * (lnla - lnlb);
  const double L22 = log(F22);
  const double tr = L2[0][0] + L2[1][1] + L22;
  // stress update of the point (state S_k rotated to the current frame)
  double St[9];
#pragma unroll
  for (int c = 0; c < 9; c++) St[c] = g.S[c * nipt + ip];
  const double trm = tr / 3.0, g2 = 2.0 * g.G;
  St[0] += g2 * (L2[0][0] - trm);
  St[1] += g2 * L2[0][1];
  St[3] += g2 * L2[1][0];
  St[4] += g2 * (L2[1][1] - trm);
  St[8] += g2 * (L22 - trm);
  double s2sum = 0.0;
#pragma unroll
  for (int c = 0; c < 9; c++) s2sum += St[c] * St[c];
  const double snrm = sqrt(s2sum), Strial = kSqrt32 * snrm;
  const double evp = g.evp[ip];
  double tf = 1.0, dtf = 0.0;
  if (g.jc) adj_theta(ea, g.T[ip], tf, dtf);
  const bool plastic = Strial > (g.A + g.B * pow(evp, g.n)) * tf;
  double gam = 0.0;
  if (plastic)
  {
    double h = g.n * g.B * pow(evp, g.n - 1.0) * tf;
    gam = (snrm - kSqrt23 * (g.A + g.B * pow(evp, g.n)) * tf) / (g2 * (1.0 + h / (3.0 * g.G)));
    for (int it = 0; it < 100; it++)
    {
      const double ev = evp + kSqrt23 * gam;
      const double sy = (g.A + g.B * pow(ev, g.n)) * tf;
      h = g.n * g.B * pow(ev, g.n - 1.0) * tf;
      const double phi = kSqrt32 * g2 * gam + sy - Strial;
      if (fabs(phi) <= 1e-15 * sy) break;
      double dg = -phi / (kSqrt32 * g2 + kSqrt23 * h);
      if (dg < 0 && -dg > gam) dg = -0.9 * gam;
      gam += dg;
    }
  }
  const double cfac = plastic ? 1.0 - g2 * gam / snrm : 1.0;
  double Sp[9];
#pragma unroll
  for (int c = 0; c < 9; c++) Sp[c] = cfac * St[c];
  // pressure: p_new = p + K mean over the element's points of tr(ln U)
  const double ptr = sum4(tr);
  const double pnew = g.p[e] + g.K * (ptr / 4.0);
  // Srot = R Sp R^T with R = diag(R2, 1)
  const double R3[3][3] = {{R2[0][0], R2[0][1], 0}, {R2[1][0], R2[1][1], 0}, {0, 0, 1}};
  double RS[3][3], Sig[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      RS[i][j] = R3[i][0] * Sp[0 * 3 + j] + R3[i][1] * Sp[1 * 3 + j] + R3[i][2] * Sp[2 * 3 + j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Sig[i][j] = RS[i][0] * R3[j][0] + RS[i][1] * R3[j][1] + RS[i][2] * R3[j][2] + (i == j ? pnew : 0.0);
  const double W = det * kDeuxPI * Rr;

  // ---- reverse ----
  double fbn[4][2];
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    fbn[k][0] = g.fb[2 * (long long)nd[k]];
    fbn[k][1] = g.fb[2 * (long long)nd[k] + 1];
  }
  // forces f_L0 = -sum_q (d_qL0 Sig00 + d_qL1 Sig10 + N_qL Sig22 / Rr) W,
  //        f_L1 = -sum_q (d_qL0 Sig01 + d_qL1 Sig11) W
  double Sgb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double dNb[4][2], Wb = 0.0, Rrb = 0.0;
#pragma unroll
  for (int L = 0; L < 4; L++)
  {
    const double f0 = fbn[L][0], f1 = fbn[L][1];
    Sgb[0][0] -= f0 * dN[L][0] * W;
    Sgb[1][0] -= f0 * dN[L][1] * W;
    Sgb[2][2] -= f0 * Nq[L] * W / Rr;
    Sgb[0][1] -= f1 * dN[L][0] * W;
    Sgb[1][1] -= f1 * dN[L][1] * W;
    dNb[L][0] = -(f0 * Sig[0][0] + f1 * Sig[0][1]) * W;
    dNb[L][1] = -(f0 * Sig[1][0] + f1 * Sig[1][1]) * W;
    Wb -= f0 * (dN[L][0] * Sig[0][0] + dN[L][1] * Sig[1][0] + Nq[L] * Sig[2][2] / Rr) +
          f1 * (dN[L][0] * Sig[0][1] + dN[L][1] * Sig[1][1]);
    Rrb += f0 * Nq[L] * Sig[2][2] * W / (Rr * Rr);
  }
  double detb = Wb * kDeuxPI * Rr;
  Rrb += Wb * kDeuxPI * det;
  // Sig = Srot + p_new I
  double Srb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Srb[i][j] = g.Sb[(i * 3 + j) * nipt + ip] + Sgb[i][j];
  const double pbt = g.pb[e] + sum4(Sgb[0][0] + Sgb[1][1] + Sgb[2][2]);
  // Srot = R Sp R^T: Spb = R^T Srb R, Rb = Srb R Sp^T + Srb^T R Sp
  double T1[3][3], Spb[3][3], Rb2[2][2];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) T1[i][j] = Srb[i][0] * R3[0][j] + Srb[i][1] * R3[1][j] + Srb[i][2] * R3[2][j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Spb[i][j] = R3[0][i] * T1[0][j] + R3[1][i] * T1[1][j] + R3[2][i] * T1[2][j];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
    {
      // (Srb R Sp^T)_ij + (Srb^T R Sp)_ij
      double v1 = 0.0, v2 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++)
        {
          v1 += Srb[i][k] * R3[k][l] * Sp[j * 3 + l];
          v2 += Srb[k][i] * R3[k][l] * Sp[l * 3 + j];
        }
      Rb2[i][j] = v1 + v2;
    }
  // return mapping reversed
  double Stb[9];
  double cb = 0.0;
#pragma unroll
  for (int c = 0; c < 9; c++)
  {
    Stb[c] = cfac * Spb[c / 3][c % 3];
    cb += Spb[c / 3][c % 3] * St[c];
  }
  const double evbo = g.evb[ip];
  const double Tbo = g.jc ? g.Tb[ip] : 0.0;
  double evbi = evbo, Tbi = Tbo, thA = 0.0, thB = 0.0, thN = 0.0;
  if (plastic)
  {
    double gb = -g2 / snrm * cb + kSqrt23 * evbo;
    double sb = g2 * gam / (snrm * snrm) * cb;
    const double ev = evp + kSqrt23 * gam;
    const double evn = pow(ev, g.n), lev = log(ev);
    const double base = g.A + g.B * evn;
    const double hp = g.n * g.B * pow(ev, g.n - 1.0);
    if (g.jc)
    {
      // T' = T + chi/(rho cp) base(ev) theta(T) dep, dep = sqrt(2/3) gamma
      const double kq = Tbo * ea.chi_rho_cp, dep = kSqrt23 * gam;
      gb += kq * tf * kSqrt23 * (hp * dep + base);
      evbi += kq * tf * hp * dep;
      Tbi += kq * base * dtf * dep;
      thA += kq * tf * dep;
      thB += kq * tf * dep * evn;
      thN += kq * tf * dep * g.B * evn * lev;
    }
    // consistency sqrt(3/2) 2G gamma + base(evp + sqrt(2/3) gamma) theta(T) = Strial
    const double w = gb / (kSqrt32 * g2 + kSqrt23 * hp * tf);
    evbi -= w * hp * tf;
    Tbi -= w * base * dtf;
    thA -= w * tf;
    thB -= w * tf * evn;
    thN -= w * tf * g.B * evn * lev;
    sb += kSqrt32 * w;
#pragma unroll
    for (int c = 0; c < 9; c++) Stb[c] += sb / snrm * St[c];
  }
  // S_t = S + 2G dev(dE): S-bar in, dE-bar = 2G dev(St-bar) + (K/4) p-bar I
  const double trX = g2 * (Stb[0] + Stb[4] + Stb[8]) / 3.0, trb = g.K / 4.0 * pbt;
  const double Lb00 = g2 * Stb[0] - trX + trb, Lb11 = g2 * Stb[4] - trX + trb, L22b = g2 * Stb[8] - trX + trb;
  const double Lb01 = 0.5 * g2 * (Stb[1] + Stb[3]);
  // ln U2 reversed (Daleckii-Krein in the eigenbasis), ln F22
  double Fb[2][2];
  const double F22b = L22b / F22;
  const double M00 = cq * (cq * Lb00 + sq * Lb01) + sq * (cq * Lb01 + sq * Lb11);
  const double M11 = -sq * (-sq * Lb00 + cq * Lb01) + cq * (-sq * Lb01 + cq * Lb11);
  const double M01 = cq * (-sq * Lb00 + cq * Lb01) + sq * (-sq * Lb01 + cq * Lb11);
  const double fdd = (la - lb > 1e-12 * la) ? (lnla - lnlb) / (la - lb) : 1.0 / sqrt(la * lb);
  const double N00 = M00 / la, N11 = M11 / lb, N01 = M01 * fdd;
  double U2b[2][2];
  U2b[0][0] = cq * cq * N00 - 2.0 * cq * sq * N01 + sq * sq * N11;
  U2b[1][1] = sq * sq * N00 + 2.0 * cq * sq * N01 + cq * cq * N11;
  U2b[0][1] = U2b[1][0] = cq * sq * N00 + (cq * cq - sq * sq) * N01 - cq * sq * N11;
  // U2 = R2^T F2: F2-bar = R2 U2b, R2-bar += F2 U2b^T
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
    {
      Fb[i][j] = R2[i][0] * U2b[0][j] + R2[i][1] * U2b[1][j];
      Rb2[i][j] += F[i][0] * U2b[j][0] + F[i][1] * U2b[j][1];
    }
  const double phib = -sr * Rb2[0][0] - cr * Rb2[0][1] + cr * Rb2[1][0] - sr * Rb2[1][1];
  Fb[0][0] -= phib * pb_ / rr2;
  Fb[1][1] -= phib * pb_ / rr2;
  Fb[1][0] += phib * pa / rr2;
  Fb[0][1] -= phib * pa / rr2;
  // F = I + sum_k dN_k (x) u_k, F22 = 1 + vr / Rr
  double Ub[4][2];
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    dNb[k][0] += Fb[0][0] * U[k][0] + Fb[1][0] * U[k][1];
    dNb[k][1] += Fb[0][1] * U[k][0] + Fb[1][1] * U[k][1];
    Ub[k][0] = Fb[0][0] * dN[k][0] + Fb[0][1] * dN[k][1] + F22b * Nq[k] / Rr;
    Ub[k][1] = Fb[1][0] * dN[k][0] + Fb[1][1] * dN[k][1];
  }
  Rrb -= F22b * vr / (Rr * Rr);
  // dN = g inv; inv = J^-1; det
  double invb[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) invb[i][j] += gs[k][i] * dNb[k][j];
  double Jb[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int l = 0; l < 2; l++)
    {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int k = 0; k < 2; k++) s += inv[j][i] * invb[j][k] * inv[l][k];
      Jb[i][l] = -s;
    }
  Jb[0][0] += detb * J[1][1];
  Jb[1][1] += detb * J[0][0];
  Jb[0][1] -= detb * J[1][0];
  Jb[1][0] -= detb * J[0][1];
  double Xb[4][2];
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    Xb[k][0] = Jb[0][0] * gs[k][0] + Jb[0][1] * gs[k][1] + Rrb * Nq[k];
    Xb[k][1] = Jb[1][0] * gs[k][0] + Jb[1][1] * gs[k][1];
  }
  // ---- outputs ----
  if (valid)
  {
    g.Sb[0 * nipt + ip] = Stb[0];
#pragma unroll
    for (int c = 1; c < 9; c++) g.Sb[c * nipt + ip] = Stb[c];
    g.evb[ip] = evbi;
    if (g.jc) g.Tb[ip] = Tbi;
  }
  // per-node sums over the element's four points; lane pt writes node pt
  double ox0 = 0, ox1 = 0, ou0 = 0, ou1 = 0;
#pragma unroll
  for (int L = 0; L < 4; L++)
  {
    const double a0 = sum4(Xb[L][0]), a1 = sum4(Xb[L][1]), b0 = sum4(Ub[L][0]), b1 = sum4(Ub[L][1]);
    if (L == pt) { ox0 = a0; ox1 = a1; ou0 = b0; ou1 = b1; }
  }
  if (valid)
  {
    const long long slot = ea.fpos[e * 4 + pt];
    g.slot_x[2 * slot] = ox0;
    g.slot_x[2 * slot + 1] = ox1;
    g.slot_u[2 * slot] = ou0;
    g.slot_u[2 * slot + 1] = ou1;
    if (pt == 0) g.pb[e] = pbt;
  }
  // parameter adjoints: one partial per block
  double t3[3] = {valid ? thA : 0.0, valid ? thB : 0.0, valid ? thN : 0.0};
#pragma unroll
  for (int j = 0; j < 3; j++)
  {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t3[j] += __shfl_xor(t3[j], o, 64);
  }
  if (tid == 0)
#pragma unroll
    for (int j = 0; j < 3; j++) g.th_part[3 * blockIdx.x + j] = t3[j];
}

// ---------------------------------------------------------------------------
// node gather: x-bar_k = x-bar_{k+1} + element terms, du-bar_k = element terms
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_gather(AdjArgs g)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nn) return;
  double sx0 = 0, sx1 = 0, su0 = 0, su1 = 0;
  const int dg = g.deg[i];
  for (int s = 0; s < dg; s++)
  {
    const long long q = 2 * ((long long)s * g.nn + i);
    sx0 += g.slot_x[q];
    sx1 += g.slot_x[q + 1];
    su0 += g.slot_u[q];
    su1 += g.slot_u[q + 1];
  }
  g.xb[2 * i] += sx0;
  g.xb[2 * i + 1] += sx1;
  g.ub[2 * i] = su0;
  g.ub[2 * i + 1] = su1;
}

// d(length / cs) / d(node coordinates) of ElQua4nAx::getLength (FE/ElQua4nAx.C:131-143)
__device__ void quad_length_grad(const double *x, const int *cn, double cs, double w, double *xb)
{
  const double *x0 = x + 3 * (long long)cn[0], *x1 = x + 3 * (long long)cn[1], *x2 = x + 3 * (long long)cn[2],
               *x3 = x + 3 * (long long)cn[3];
  const double x31 = x2[0] - x0[0], x24 = x1[0] - x3[0], y42 = x3[1] - x1[1], y31 = x2[1] - x0[1];
  const double num = x31 * y42 + x24 * y31;
  const double q = 2.0 * (x24 * x24 + y42 * y42 + x31 * x31 + y31 * y31), den = sqrt(q);
  const double f = num / (2.0 * q * den);
  const double gx31 = (y42 / den - f * 4.0 * x31) * w / cs, gy42 = (x31 / den - f * 4.0 * y42) * w / cs;
  const double gx24 = (y31 / den - f * 4.0 * x24) * w / cs, gy31 = (x24 / den - f * 4.0 * y31) * w / cs;
  xb[2 * (long long)cn[2]] += gx31;
  xb[2 * (long long)cn[0]] -= gx31;
  xb[2 * (long long)cn[1]] += gx24;
  xb[2 * (long long)cn[3]] -= gx24;
  xb[2 * (long long)cn[3] + 1] += gy42;
  xb[2 * (long long)cn[1] + 1] -= gy42;
  xb[2 * (long long)cn[2] + 1] += gy31;
  xb[2 * (long long)cn[0] + 1] -= gy31;
}

// ---------------------------------------------------------------------------
// scalars: partial sums, the time-step computation of increment k+1 reversed
// (Courant minimum through its arg-min element, or the LINEARdbl
// extrapolation, or an unchanged dt), parameter adjoint accumulation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_scalar(AdjArgs g)
{
  __shared__ double red[256];
  __shared__ long long ridx[256];
  __shared__ double th3[3];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int k = tid; k < g.ndt_part; k += 256) s += g.dt_part[k];
  red[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1)
  {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double dtb_n = red[0];
  __syncthreads();
  for (int j = 0; j < 3; j++)
  {
    double t = 0.0;
    for (int k = tid; k < g.nth_part; k += 256) t += g.th_part[3 * k + j];
    red[tid] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) th3[j] = red[0];
    __syncthreads();
  }
  const int rec = g.k > 0 && (g.k % g.freq) == 0;
  long long amin = -1;
  if (rec)
  {
    // arg-min of the Courant lengths on x_k (ties to the lower element index)
    double best = __builtin_inf();
    long long bi = -1;
    for (long long e = tid; e < g.ne; e += 256)
    {
      const double l = quad_length(g.x, g.ea.conn + e * 4) / g.cs;
      if (l < best) { best = l; bi = e; }
    }
    red[tid] = best;
    ridx[tid] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
      if (tid < w)
      {
        const double o = red[tid + w];
        const long long oi = ridx[tid + w];
        if (o < red[tid] || (o == red[tid] && oi >= 0 && (ridx[tid] < 0 || oi < ridx[tid])))
        {
          red[tid] = o;
          ridx[tid] = oi;
        }
      }
      __syncthreads();
    }
    amin = ridx[0];
  }
  if (tid != 0) return;
  AdjScalars &c = *g.sc;
#pragma unroll
  for (int j = 0; j < 3; j++) c.th[j] += th3[j];
  const double dtb = c.tb + c.dtb_carry + dtb_n;
  c.dtb_carry = 0.0;
  if (g.k == 0) return;  // dt of the first increment comes from the initial mesh
  const StepState &s0 = *g.st0;
  if (rec)
  {
    const double critb = (dtb + c.Cb) * g.safety * g.omegaS / 2.0;
    if (amin >= 0) quad_length_grad(g.x, g.ea.conn + amin * 4, g.cs, critb, g.xb);
    const double tadd = c.tsb;
    c.tsb = c.ltsb;
    c.Cb = c.Lb;
    c.ltsb = 0.0;
    c.Lb = 0.0;
    c.tb += tadd;
  }
  else if (s0.tsDate != s0.lastTsDate)
  {
    const double t = g.st1->t, ts = s0.tsDate, lts = s0.lastTsDate;
    const double dx = (t - lts) / (ts - lts);
    c.Cb += dtb * dx;
    c.Lb += dtb * (1.0 - dx);
    const double dxb = dtb * (s0.computedTimeStep - s0.lastTimeStep);
    c.tb += dxb / (ts - lts);
    c.tsb -= dxb * (t - lts) / ((ts - lts) * (ts - lts));
    c.ltsb += dxb * (t - ts) / ((ts - lts) * (ts - lts));
  }
  else
    c.dtb_carry = dtb;
}

// state snapshot / restore: up to kCopyMax arrays copied in one launch (blockIdx.y = array)
constexpr int kCopyMax = 12;
struct CopyList
{
  const double *src[kCopyMax];
  double *dst[kCopyMax];
  long long n[kCopyMax];
};

__global__ __launch_bounds__(256) void k_copy_list(CopyList L)
{
  const int a = blockIdx.y;
  const long long n = L.n[a];
  const double *src = L.src[a];
  double *dst = L.dst[a];
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

bool launch_copy_list(const double *const *src, double *const *dst, const long long *n, int count, hipStream_t s)
{
  if (count < 1 || count > kCopyMax) return false;
  CopyList L{};
  long long mx = 0;
  for (int k = 0; k < count; k++)
  {
    L.src[k] = src[k];
    L.dst[k] = dst[k];
    L.n[k] = n[k];
    mx = n[k] > mx ? n[k] : mx;
  }
  long long blocks = (mx + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_copy_list, dim3((unsigned)blocks, count), dim3(256), 0, s, L);
  return true;
}

void launch_adjoint_step(const AdjArgs &g, hipStream_t s)
{
  hipLaunchKernelGGL(k_adj_nodal, dim3(g.ndt_part), dim3(256), 0, s, g);
  hipLaunchKernelGGL(k_adj_elem, dim3(g.nth_part), dim3(64), 0, s, g);
  hipLaunchKernelGGL(k_adj_gather, dim3(g.ndt_part), dim3(256), 0, s, g);
  hipLaunchKernelGGL(k_adj_scalar, dim3(1), dim3(256), 0, s, g);
}

}  // namespace dnl
