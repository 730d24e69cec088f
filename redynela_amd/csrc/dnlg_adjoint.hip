// Hand-written adjoint (reverse mode) of one increment of the axisymmetric or
// hex8 ExplicitSolverCH time loop, for gradients of a function of the final
// nodal coordinates with respect to the hardening parameters (A, B, n) of
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
// the Daleckii-Krein divided differences for ln U), in 3D ln U = f(F^T F) and
// U^-1 = g(F^T F) through the eigen-decomposition of F^T F converged to
// round-off (Daleckii-Krein for both, R = F U^-1), and the converged
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
template <int DIM>
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
    for (int c = 0; c < DIM; c++)
    {
      const long long q = DIM * i + c;
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

// the radial return of one point as the exact map (FE/Element.C:248-348,
// Newton converged to round-off): trial deviator St in, the plastic switch,
// gamma and the scaling cfac of S' = cfac St out
struct AdjReturn
{
  double snrm, tf, dtf, gam, cfac;
  bool plastic;
};

__device__ __forceinline__ AdjReturn adj_return_fwd(const AdjArgs &g, const double *St, double evp, double T)
{
  AdjReturn r;
  const double g2 = 2.0 * g.G;
  double s2sum = 0.0;
#pragma unroll
  for (int c = 0; c < 9; c++) s2sum += St[c] * St[c];
  r.snrm = sqrt(s2sum);
  const double Strial = kSqrt32 * r.snrm;
  r.tf = 1.0;
  r.dtf = 0.0;
  if (g.jc) adj_theta(g.ea, T, r.tf, r.dtf);
  const double tf = r.tf;
  r.plastic = Strial > (g.A + g.B * pow(evp, g.n)) * tf;
  double gam = 0.0;
  if (r.plastic)
  {
    double h = g.n * g.B * pow(evp, g.n - 1.0) * tf;
    gam = (r.snrm - kSqrt23 * (g.A + g.B * pow(evp, g.n)) * tf) / (g2 * (1.0 + h / (3.0 * g.G)));
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
  r.gam = gam;
  r.cfac = r.plastic ? 1.0 - g2 * gam / r.snrm : 1.0;
  return r;
}

// its reverse: Spb (adjoint of S'), evbo and Tbo (adjoints of evp' and T')
// in; Stb, evbi, Tbi and the (A, B, n) contributions out
__device__ __forceinline__ void adj_return_rev(const AdjArgs &g, const AdjReturn &r, const double *St, double evp,
                                               const double *Spb, double evbo, double Tbo, double *Stb, double &evbi,
                                               double &Tbi, double *th)
{
  const double g2 = 2.0 * g.G, tf = r.tf, dtf = r.dtf, gam = r.gam, snrm = r.snrm;
  double cb = 0.0;
#pragma unroll
  for (int c = 0; c < 9; c++)
  {
    Stb[c] = r.cfac * Spb[c];
    cb += Spb[c] * St[c];
  }
  evbi = evbo;
  Tbi = Tbo;
  th[0] = th[1] = th[2] = 0.0;
  if (!r.plastic) return;
  double gb = -g2 / snrm * cb + kSqrt23 * evbo;
  double sb = g2 * gam / (snrm * snrm) * cb;
  const double ev = evp + kSqrt23 * gam;
  const double evn = pow(ev, g.n), lev = log(ev);
  const double base = g.A + g.B * evn;
  const double hp = g.n * g.B * pow(ev, g.n - 1.0);
  if (g.jc)
  {
    // T' = T + chi/(rho cp) base(ev) theta(T) dep, dep = sqrt(2/3) gamma
    const double kq = Tbo * g.ea.chi_rho_cp, dep = kSqrt23 * gam;
    gb += kq * tf * kSqrt23 * (hp * dep + base);
    evbi += kq * tf * hp * dep;
    Tbi += kq * base * dtf * dep;
    th[0] += kq * tf * dep;
    th[1] += kq * tf * dep * evn;
    th[2] += kq * tf * dep * g.B * evn * lev;
  }
  // consistency sqrt(3/2) 2G gamma + base(evp + sqrt(2/3) gamma) theta(T) = Strial
  const double w = gb / (kSqrt32 * g2 + kSqrt23 * hp * tf);
  evbi -= w * hp * tf;
  Tbi -= w * base * dtf;
  th[0] -= w * tf;
  th[1] -= w * tf * evn;
  th[2] -= w * tf * g.B * evn * lev;
  sb += kSqrt32 * w;
#pragma unroll
  for (int c = 0; c < 9; c++) Stb[c] += sb / snrm * St[c];
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
  const double evp = g.evp[ip];
  const AdjReturn ret = adj_return_fwd(g, St, evp, g.jc ? g.T[ip] : 0.0);
  const double cfac = ret.cfac;
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
  double Stb[9], Spf[9], evbi, Tbi, th3r[3];
#pragma unroll
  for (int c = 0; c < 9; c++) Spf[c] = Spb[c / 3][c % 3];
  adj_return_rev(g, ret, St, evp, Spf, g.evb[ip], g.jc ? g.Tb[ip] : 0.0, Stb, evbi, Tbi, th3r);
  const double thA = th3r[0], thB = th3r[1], thN = th3r[2];
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
// hex8: the same element pass in 3D, lane = 8 element + point
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sum8(double v)
{
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}

// symmetric 3x3 eigen-decomposition C = Q diag(lam) Q^T by cyclic Jacobi
// rotations, swept until the off-diagonal part is below round-off (the
// forward's polar decomposition stops at 1e-9; its derivative is taken on the
// exact decomposition)
__device__ void sym3_eig(const double C[3][3], double lam[3], double Q[3][3])
{
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      A[i][j] = C[i][j];
      Q[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 50; sweep++)
  {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (!(off > 1e-36 * dia)) break;
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
      const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2, m = 3 - p - q;
      const double apq = A[p][q];
      if (apq == 0.0) continue;
      const double th = (A[q][q] - A[p][p]) / (2.0 * apq);
      const double t = fabs(th) > 1e150 ? 0.5 / th : (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c, tau = sn / (1.0 + c);
      A[p][p] -= t * apq;
      A[q][q] += t * apq;
      A[p][q] = A[q][p] = 0.0;
      const double amp = A[m][p], amq = A[m][q];
      A[m][p] = A[p][m] = amp - sn * (amq + tau * amp);
      A[m][q] = A[q][m] = amq + sn * (amp - tau * amq);
#pragma unroll
      for (int k = 0; k < 3; k++)
      {
        const double gk = Q[k][p], hk = Q[k][q];
        Q[k][p] = gk - sn * (hk + tau * gk);
        Q[k][q] = hk + sn * (gk - tau * hk);
      }
    }
  }
  lam[0] = A[0][0];
  lam[1] = A[1][1];
  lam[2] = A[2][2];
}

// divided differences of f = ln(lam) / 2 and of g = lam^(-1/2), the
// Daleckii-Krein kernels of ln U = f(C) and U^-1 = g(C)
__device__ __forceinline__ double dd_half_log(double a, double b)
{
  const double s = a + b, t = (a - b) / s;
  return t == 0.0 ? 1.0 / s : atanh(t) / (t * s);
}
__device__ __forceinline__ double dd_inv_sqrt(double a, double b)
{
  const double ra = sqrt(a), rb = sqrt(b);
  return -1.0 / (ra * rb * (ra + rb));
}

__global__ __launch_bounds__(64) void k_adj_elem_hex8(AdjArgs g)
{
  const ElemArgs &ea = g.ea;
  const int tid = threadIdx.x, pt = tid & 7;
  const long long e_raw = ((long long)blockIdx.x * 64 + tid) >> 3;
  const bool valid = e_raw < g.ne;
  const long long e = valid ? e_raw : g.ne - 1;
  const long long nipt = g.ne * 8, ip = e * 8 + pt;
  const int *cn = ea.conn + e * 8;
  // ---- forward recompute ----
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    const double *xk = g.x + 3 * (long long)cn[k];
    double gk[3];
    hex_dn0(ea, pt, k, gk);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int d = 0; d < 3; d++) J[c][d] += gk[d] * xk[c];
  }
  double cof[3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3, d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      cof[c][d] = J[c1][d1] * J[c2][d2] - J[c1][d2] * J[c2][d1];
    }
  const double det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
  double inv[3][3];  // J^-1 = cof^T / det
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) inv[i][j] = cof[j][i] / det;
  double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    const double *uk = g.du + 3 * (long long)cn[k];
    double gk[3], dn[3];
    hex_dn0(ea, pt, k, gk);
#pragma unroll
    for (int j = 0; j < 3; j++) dn[j] = gk[0] * inv[0][j] + gk[1] * inv[1][j] + gk[2] * inv[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) F[i][j] += uk[i] * dn[j];
  }
  // F = R U, C = F^T F = Q diag(lam) Q^T, ln U = Q diag(ln(lam)/2) Q^T, U^-1 = Q diag(lam^-1/2) Q^T
  double Cm[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Cm[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
  double lam[3], Q[3][3];
  sym3_eig(Cm, lam, Q);
  const double fl[3] = {0.5 * log(lam[0]), 0.5 * log(lam[1]), 0.5 * log(lam[2])};
  const double gl[3] = {1.0 / sqrt(lam[0]), 1.0 / sqrt(lam[1]), 1.0 / sqrt(lam[2])};
  double dE[3][3], Vi[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      dE[i][j] = Q[i][0] * fl[0] * Q[j][0] + Q[i][1] * fl[1] * Q[j][1] + Q[i][2] * fl[2] * Q[j][2];
      Vi[i][j] = Q[i][0] * gl[0] * Q[j][0] + Q[i][1] * gl[1] * Q[j][1] + Q[i][2] * gl[2] * Q[j][2];
    }
  double R[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) R[i][j] = F[i][0] * Vi[0][j] + F[i][1] * Vi[1][j] + F[i][2] * Vi[2][j];
  const double tr = dE[0][0] + dE[1][1] + dE[2][2];
  const double g2 = 2.0 * g.G;
  double St[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) St[3 * i + j] = g.S[(3 * i + j) * nipt + ip] + g2 * (dE[i][j] - (i == j ? tr / 3.0 : 0.0));
  const double evp = g.evp[ip];
  const AdjReturn ret = adj_return_fwd(g, St, evp, g.jc ? g.T[ip] : 0.0);
  double Sp[9];
#pragma unroll
  for (int c = 0; c < 9; c++) Sp[c] = ret.cfac * St[c];
  const double pnew = g.p[e] + g.K * (sum8(tr) / 8.0);
  double RS[3][3], Sig[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) RS[i][j] = R[i][0] * Sp[j] + R[i][1] * Sp[3 + j] + R[i][2] * Sp[6 + j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Sig[i][j] = RS[i][0] * R[j][0] + RS[i][1] * R[j][1] + RS[i][2] * R[j][2] + (i == j ? pnew : 0.0);
  const double W = det;  // Gauss weight 1

  // ---- reverse ----
  // f_L,i = -sum_q sum_j dN_qL,j Sig_q,ji W_q
  double Sgb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, invb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double Wb = 0.0;
#pragma unroll
  for (int L = 0; L < 8; L++)
  {
    const double *fbL = g.fb + 3 * (long long)cn[L];
    const double f[3] = {fbL[0], fbL[1], fbL[2]};
    double gk[3], dn[3], dnb[3];
    hex_dn0(ea, pt, L, gk);
#pragma unroll
    for (int j = 0; j < 3; j++) dn[j] = gk[0] * inv[0][j] + gk[1] * inv[1][j] + gk[2] * inv[2][j];
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      double sj = 0.0;
#pragma unroll
      for (int i = 0; i < 3; i++)
      {
        Sgb[j][i] -= f[i] * dn[j] * W;
        sj += f[i] * Sig[j][i];
      }
      dnb[j] = -W * sj;
      Wb -= dn[j] * sj;
    }
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int j = 0; j < 3; j++) invb[d][j] += gk[d] * dnb[j];
  }
  // Sig = Srot + p_new I, Srot = R Sp R^T
  double Srb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Srb[i][j] = g.Sb[(3 * i + j) * nipt + ip] + Sgb[i][j];
  const double pbt = g.pb[e] + sum8(Sgb[0][0] + Sgb[1][1] + Sgb[2][2]);
  double T1[3][3], Spb[9], Rb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) T1[i][j] = Srb[i][0] * R[0][j] + Srb[i][1] * R[1][j] + Srb[i][2] * R[2][j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Spb[3 * i + j] = R[0][i] * T1[0][j] + R[1][i] * T1[1][j] + R[2][i] * T1[2][j];
  // R-bar = Srb R Sp^T + Srb^T R Sp
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      double v1 = 0.0, v2 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++)
        {
          v1 += Srb[i][k] * R[k][l] * Sp[3 * j + l];
          v2 += Srb[k][i] * R[k][l] * Sp[3 * l + j];
        }
      Rb[i][j] = v1 + v2;
    }
  double Stb[9], evbi, Tbi, th3[3];
  adj_return_rev(g, ret, St, evp, Spb, g.evb[ip], g.jc ? g.Tb[ip] : 0.0, Stb, evbi, Tbi, th3);
  // St = S + 2G dev(dE), p_new = p + K mean tr(dE): dE-bar, symmetric
  const double trX = g2 * (Stb[0] + Stb[4] + Stb[8]) / 3.0, trb = g.K / 8.0 * pbt;
  double Lb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Lb[i][j] = 0.5 * g2 * (Stb[3 * i + j] + Stb[3 * j + i]) + (i == j ? trb - trX : 0.0);
  // R = F U^-1: F-bar = Rb U^-1, (U^-1)-bar = F^T Rb (symmetrised)
  double Fb[3][3], Vb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      Fb[i][j] = Rb[i][0] * Vi[0][j] + Rb[i][1] * Vi[1][j] + Rb[i][2] * Vi[2][j];
      Vb[i][j] = F[0][i] * Rb[0][j] + F[1][i] * Rb[1][j] + F[2][i] * Rb[2][j];
    }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i + 1; j < 3; j++) Vb[i][j] = Vb[j][i] = 0.5 * (Vb[i][j] + Vb[j][i]);
  // C-bar = Q [(Q^T Lb Q) o Df + (Q^T Vb Q) o Dg] Q^T
  double ML[3][3], MV[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++)
    {
      double sl = 0.0, sv = 0.0;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
        {
          sl += Q[i][a] * Lb[i][j] * Q[j][b];
          sv += Q[i][a] * Vb[i][j] * Q[j][b];
        }
      const double df = a == b ? 0.5 / lam[a] : dd_half_log(lam[a], lam[b]);
      const double dg = a == b ? -0.5 * gl[a] / lam[a] : dd_inv_sqrt(lam[a], lam[b]);
      ML[a][b] = sl * df + sv * dg;
    }
  double Cb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) v += Q[i][a] * ML[a][b] * Q[j][b];
      MV[i][j] = v;
    }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Cb[i][j] = 0.5 * (MV[i][j] + MV[j][i]);
  // C = F^T F: F-bar += 2 F C-bar
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Fb[i][j] += 2.0 * (F[i][0] * Cb[0][j] + F[i][1] * Cb[1][j] + F[i][2] * Cb[2][j]);
  // F = I + sum_k du_k (x) dN_k: du-bar per node, dN-bar into J^-1-bar
  double ou[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    const double *uk = g.du + 3 * (long long)cn[k];
    const double u[3] = {uk[0], uk[1], uk[2]};
    double gk[3], dn[3];
    hex_dn0(ea, pt, k, gk);
#pragma unroll
    for (int j = 0; j < 3; j++) dn[j] = gk[0] * inv[0][j] + gk[1] * inv[1][j] + gk[2] * inv[2][j];
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
      const double ub = sum8(Fb[i][0] * dn[0] + Fb[i][1] * dn[1] + Fb[i][2] * dn[2]);
      if (k == pt) ou[i] = ub;
    }
#pragma unroll
    for (int j = 0; j < 3; j++)
    {
      const double dnb = Fb[0][j] * u[0] + Fb[1][j] * u[1] + Fb[2][j] * u[2];
#pragma unroll
      for (int d = 0; d < 3; d++) invb[d][j] += gk[d] * dnb;
    }
  }
  // J^-1-bar -> J-bar = -J^-T invb J^-T, det-bar = W-bar: J-bar += W-bar cof
  double Jb[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int l = 0; l < 3; l++)
    {
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) v += inv[j][i] * invb[j][k] * inv[l][k];
      Jb[i][l] = -v + Wb * cof[i][l];
    }
  // J[c][d] = sum_k dN0_k,d x_k,c
  double ox[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 8; k++)
  {
    double gk[3];
    hex_dn0(ea, pt, k, gk);
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
      const double xb = sum8(Jb[c][0] * gk[0] + Jb[c][1] * gk[1] + Jb[c][2] * gk[2]);
      if (k == pt) ox[c] = xb;
    }
  }
  // ---- outputs ----
  if (valid)
  {
#pragma unroll
    for (int c = 0; c < 9; c++) g.Sb[c * nipt + ip] = Stb[c];
    g.evb[ip] = evbi;
    if (g.jc) g.Tb[ip] = Tbi;
    const long long slot = ea.fpos[e * 8 + pt];
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
      g.slot_x[3 * slot + c] = ox[c];
      g.slot_u[3 * slot + c] = ou[c];
    }
    if (pt == 0) g.pb[e] = pbt;
  }
  double t3[3] = {valid ? th3[0] : 0.0, valid ? th3[1] : 0.0, valid ? th3[2] : 0.0};
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
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_gather(AdjArgs g)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nn) return;
  double sx[DIM], su[DIM];
#pragma unroll
  for (int c = 0; c < DIM; c++) sx[c] = su[c] = 0.0;
  const int dg = g.deg[i];
  for (int s = 0; s < dg; s++)
  {
    const long long q = DIM * ((long long)s * g.nn + i);
#pragma unroll
    for (int c = 0; c < DIM; c++)
    {
      sx[c] += g.slot_x[q + c];
      su[c] += g.slot_u[q + c];
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; c++)
  {
    g.xb[DIM * i + c] += sx[c];
    g.ub[DIM * i + c] = su[c];
  }
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

// d(length / cs) / d(node coordinates) of the hex8 length V / A_max (V the
// sum of the 8 Gauss-point detJ, A_max the largest face area, first face on
// ties), the length the forward's Courant step uses
__device__ void hex_length_grad(const AdjArgs &g, const int *cn, double w, double *xb)
{
  const ElemArgs &ea = g.ea;
  double X[8][3];
  for (int k = 0; k < 8; k++)
    for (int c = 0; c < 3; c++) X[k][c] = g.x[3 * (long long)cn[k] + c];
  double V = 0.0, dV[8][3];
  for (int k = 0; k < 8; k++)
    for (int c = 0; c < 3; c++) dV[k][c] = 0.0;
  for (int pt = 0; pt < 8; pt++)
  {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 8; k++)
    {
      double gk[3];
      hex_dn0(ea, pt, k, gk);
      for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) J[c][d] += gk[d] * X[k][c];
    }
    double cof[3][3];
    for (int c = 0; c < 3; c++)
      for (int d = 0; d < 3; d++)
      {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3, d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        cof[c][d] = J[c1][d1] * J[c2][d2] - J[c1][d2] * J[c2][d1];
      }
    V += J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
    for (int k = 0; k < 8; k++)
    {
      double gk[3];
      hex_dn0(ea, pt, k, gk);
      for (int c = 0; c < 3; c++) dV[k][c] += cof[c][0] * gk[0] + cof[c][1] * gk[1] + cof[c][2] * gk[2];
    }
  }
  const int faces[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {7, 4, 5, 6}, {2, 1, 0, 3}};
  double Amax = 0.0, nrm[3] = {0, 0, 0}, d1m[3] = {0, 0, 0}, d2m[3] = {0, 0, 0};
  int fm = 0;
  for (int f = 0; f < 6; f++)
  {
    double d1[3], d2[3];
    for (int c = 0; c < 3; c++)
    {
      d1[c] = X[faces[f][2]][c] - X[faces[f][0]][c];
      d2[c] = X[faces[f][3]][c] - X[faces[f][1]][c];
    }
    const double cr[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2], d1[0] * d2[1] - d1[1] * d2[0]};
    const double len = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const double area = 0.5 * len;
    if (area > Amax)
    {
      Amax = area;
      fm = f;
      for (int c = 0; c < 3; c++)
      {
        nrm[c] = cr[c] / len;
        d1m[c] = d1[c];
        d2m[c] = d2[c];
      }
    }
  }
  const double lb = w / g.cs, Vb = lb / Amax, Ab = -lb * V / (Amax * Amax);
  for (int k = 0; k < 8; k++)
    for (int c = 0; c < 3; c++) xb[3 * (long long)cn[k] + c] += Vb * dV[k][c];
  // A = |d1 x d2| / 2: dA/dd1 = (d2 x n) / 2, dA/dd2 = (n x d1) / 2
  const double g1[3] = {0.5 * (d2m[1] * nrm[2] - d2m[2] * nrm[1]), 0.5 * (d2m[2] * nrm[0] - d2m[0] * nrm[2]),
                        0.5 * (d2m[0] * nrm[1] - d2m[1] * nrm[0])};
  const double g2[3] = {0.5 * (nrm[1] * d1m[2] - nrm[2] * d1m[1]), 0.5 * (nrm[2] * d1m[0] - nrm[0] * d1m[2]),
                        0.5 * (nrm[0] * d1m[1] - nrm[1] * d1m[0])};
  for (int c = 0; c < 3; c++)
  {
    xb[3 * (long long)cn[faces[fm][2]] + c] += Ab * g1[c];
    xb[3 * (long long)cn[faces[fm][0]] + c] -= Ab * g1[c];
    xb[3 * (long long)cn[faces[fm][3]] + c] += Ab * g2[c];
    xb[3 * (long long)cn[faces[fm][1]] + c] -= Ab * g2[c];
  }
}

// ---------------------------------------------------------------------------
// arg-min of the Courant lengths on x_k, per block of 256 elements (ties to
// the lower element index), for an increment that recomputes the time step
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_adj_argmin(AdjArgs g)
{
  __shared__ double red[256];
  __shared__ long long ridx[256];
  const int tid = threadIdx.x;
  const long long e = (long long)blockIdx.x * 256 + tid;
  double l = __builtin_inf();
  long long li = -1;
  if (e < g.ne)
  {
    ElemArgs ex = g.ea;
    ex.x = g.x;
    l = (DIM == 3 ? hex_length(ex, g.ea.conn + e * 8) : quad_length(g.x, g.ea.conn + e * 4)) / g.cs;
    li = e;
  }
  red[tid] = l;
  ridx[tid] = li;
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
  if (tid == 0)
  {
    g.am_len[blockIdx.x] = red[0];
    g.am_idx[blockIdx.x] = ridx[0];
  }
}

// ---------------------------------------------------------------------------
// scalars: partial sums, the time-step computation of increment k+1 reversed
// (Courant minimum through its arg-min element, or the LINEARdbl
// extrapolation, or an unchanged dt), parameter adjoint accumulation
// ---------------------------------------------------------------------------
template <int DIM>
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
    // arg-min over the per-block arg-mins of k_adj_argmin (ties to the lower element index)
    double best = __builtin_inf();
    long long bi = -1;
    for (int b = tid; b < g.nam_part; b += 256)
    {
      const double l = g.am_len[b];
      const long long li = g.am_idx[b];
      if (li >= 0 && (l < best || (l == best && (bi < 0 || li < bi))))
      {
        best = l;
        bi = li;
      }
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
    if (amin >= 0)
    {
      if (DIM == 3)
        hex_length_grad(g, g.ea.conn + amin * 8, critb, g.xb);
      else
        quad_length_grad(g.x, g.ea.conn + amin * 4, g.cs, critb, g.xb);
    }
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
  const bool rec = g.k > 0 && (g.k % g.freq) == 0;
  if (g.dim == 3)
  {
    if (rec) hipLaunchKernelGGL(k_adj_argmin<3>, dim3(g.nam_part), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_adj_nodal<3>, dim3(g.ndt_part), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_adj_elem_hex8, dim3(g.nth_part), dim3(64), 0, s, g);
    hipLaunchKernelGGL(k_adj_gather<3>, dim3(g.ndt_part), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_adj_scalar<3>, dim3(1), dim3(256), 0, s, g);
  }
  else
  {
    if (rec) hipLaunchKernelGGL(k_adj_argmin<2>, dim3(g.nam_part), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_adj_nodal<2>, dim3(g.ndt_part), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_adj_elem, dim3(g.nth_part), dim3(64), 0, s, g);
    hipLaunchKernelGGL(k_adj_gather<2>, dim3(g.ndt_part), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_adj_scalar<2>, dim3(1), dim3(256), 0, s, g);
  }
}

}  // namespace dnl
