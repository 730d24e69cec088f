// Golden-vector harness for the reference's linear-algebra layer.
//
// Test infrastructure only.  Links the reference's own Tensor2 / Matrix /
// LAPACK sources (compiled unmodified from /root/reference by
// oracle/refbuild/Makefile) and records, for seeded inputs, the exact doubles
// that the reference produces for the numerical kernels of the element
// stress update:
//   * Tensor2::polarDecomposeLnU      (LA/Tensor2.C:1154-1259)
//   * Matrix::inverse (dgetrf+dgetri) (LA/Matrix.C:1653-1723)
//   * Matrix::operator* (dgemm)       (LA/Matrix.C:644-671)
//   * Tensor2 J2 / norme / deviator / product / Transpose
//                                     (LA/Tensor2.C:484-586, 360-380)
// The outputs are raw little-endian float64 files, one per array, written to
// the directory given as argv[1].  tests/golden/make_la_golden.py turns them
// into the committed fixture.
#include <Tensor2.h>
#include <Matrix.h>
#include <Vec3D.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

static void dump(const std::string &dir, const char *name, const std::vector<double> &v)
{
  std::string p = dir + "/" + name + ".bin";
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) { perror(p.c_str()); exit(2); }
  fwrite(v.data(), sizeof(double), v.size(), f);
  fclose(f);
}

// rotation matrix about unit axis a by angle t (Rodrigues), row-major
static void rotation(double ax, double ay, double az, double t, double *Q)
{
  double n = std::sqrt(ax * ax + ay * ay + az * az);
  ax /= n; ay /= n; az /= n;
  double c = std::cos(t), s = std::sin(t), C = 1 - c;
  Q[0] = c + ax * ax * C;      Q[1] = ax * ay * C - az * s; Q[2] = ax * az * C + ay * s;
  Q[3] = ay * ax * C + az * s; Q[4] = c + ay * ay * C;      Q[5] = ay * az * C - ax * s;
  Q[6] = az * ax * C - ay * s; Q[7] = az * ay * C + ax * s; Q[8] = c + az * az * C;
}

int main(int argc, char **argv)
{
  if (argc < 2) { fprintf(stderr, "usage: %s outdir\n", argv[0]); return 1; }
  std::string out = argv[1];
  std::mt19937_64 rng(20261203ULL);
  std::uniform_real_distribution<double> U(-1.0, 1.0);

  // ---------------- polar decomposition ----------------
  const double scales[] = {0.0, 1e-7, 1e-5, 1e-4, 1e-3, 1e-2, 5e-2, 0.2};
  std::vector<double> Fin, lnU, Rout;
  for (int kind = 0; kind < 3; kind++)       // 0: general 3D, 1: axisymmetric shape, 2: rotated
    for (double sc : scales)
      for (int r = 0; r < 64; r++)
      {
        double H[9];
        for (int k = 0; k < 9; k++) H[k] = sc * U(rng);
        if (kind == 1) { H[2] = H[5] = H[6] = H[7] = 0.0; }
        double F[9];
        for (int k = 0; k < 9; k++) F[k] = H[k] + ((k % 4) == 0 ? 1.0 : 0.0);
        if (kind == 2)
        {
          double Q[9], G[9];
          rotation(U(rng), U(rng), U(rng) + 1.5, 0.3 * U(rng), Q);
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
            {
              G[3 * i + j] = 0;
              for (int k = 0; k < 3; k++) G[3 * i + j] += Q[3 * i + k] * F[3 * k + j];
            }
          for (int k = 0; k < 9; k++) F[k] = G[k];
        }
        Tensor2 T, Lu, R;
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) T(i, j) = F[3 * i + j];
        T.polarDecomposeLnU(Lu, R);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++)
          {
            Fin.push_back(T(i, j));
            lnU.push_back(Lu(i, j));
            Rout.push_back(R(i, j));
          }
      }
  dump(out, "polar_F", Fin);
  dump(out, "polar_lnU", lnU);
  dump(out, "polar_R", Rout);

  // ---------------- 3x3 inverse (dgetrf/dgetri) and dN = dN0 * J^-1 (dgemm) ----------------
  std::vector<double> Jin, Jinv, dN0in, dNout;
  for (int r = 0; r < 512; r++)
  {
    double h[3] = {1e-4 * (1.5 + U(rng)), 1e-4 * (1.5 + U(rng)), 1e-4 * (1.5 + U(rng))};
    Matrix J(3, 3);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
      {
        double base = (i == j ? h[i] : 0.0);
        double pert = (r % 4 == 0 ? 1.0 : 0.3) * 1e-4 * U(rng);  // every 4th: strongly non-diagonal (pivoting)
        J(i, j) = base + pert;
      }
    Matrix inv = J.inverse();
    Matrix dN0(8, 3);
    for (int k = 0; k < 8; k++)
      for (int j = 0; j < 3; j++) dN0(k, j) = 0.5 * U(rng);
    Matrix dN = dN0 * inv;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) { Jin.push_back(J(i, j)); Jinv.push_back(inv(i, j)); }
    for (int k = 0; k < 8; k++)
      for (int j = 0; j < 3; j++) { dN0in.push_back(dN0(k, j)); dNout.push_back(dN(k, j)); }
  }
  dump(out, "inv_J", Jin);
  dump(out, "inv_Jinv", Jinv);
  dump(out, "gemm_dN0", dN0in);
  dump(out, "gemm_dN", dNout);

  // ---------------- 2x2 dgemm (axisymmetric dN = dN0 * InvJ, 4x2 * 2x2) ----------------
  std::vector<double> A2, B2, C2;
  for (int r = 0; r < 256; r++)
  {
    Matrix A(4, 2), B(2, 2);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) A(i, j) = 0.5 * U(rng);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) B(i, j) = 1e4 * U(rng);
    Matrix C = A * B;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 2; j++) { A2.push_back(A(i, j)); C2.push_back(C(i, j)); }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) B2.push_back(B(i, j));
  }
  dump(out, "gemm2_A", A2);
  dump(out, "gemm2_B", B2);
  dump(out, "gemm2_C", C2);

  // ---------------- Tensor2 helpers used by the stress update ----------------
  std::vector<double> Sin, Rin, Ein, rfi, j2, nrm, dev;
  for (int r = 0; r < 512; r++)
  {
    Tensor2 S, R, E;
    for (int k = 0; k < 9; k++)
    {
      S(k / 3, k % 3) = 4e8 * U(rng);
      E(k / 3, k % 3) = 1e-3 * U(rng);
    }
    double Q[9];
    rotation(U(rng), U(rng), U(rng) + 1.5, 0.5 * U(rng), Q);
    for (int k = 0; k < 9; k++) R(k / 3, k % 3) = Q[k];
    Tensor2 P = R * (S * R.Transpose());
    Tensor2 D = E.deviator();
    for (int k = 0; k < 9; k++)
    {
      Sin.push_back(S(k / 3, k % 3));
      Rin.push_back(R(k / 3, k % 3));
      Ein.push_back(E(k / 3, k % 3));
      rfi.push_back(P(k / 3, k % 3));
      dev.push_back(D(k / 3, k % 3));
    }
    j2.push_back(S.J2());
    nrm.push_back(S.norme());
  }
  dump(out, "t2_S", Sin);
  dump(out, "t2_R", Rin);
  dump(out, "t2_E", Ein);
  dump(out, "t2_RSRt", rfi);
  dump(out, "t2_devE", dev);
  dump(out, "t2_J2", j2);
  dump(out, "t2_norme", nrm);
  printf("ok\n");
  return 0;
}
