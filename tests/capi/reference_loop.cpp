// Drop-in test of the C++ facade (include/dnlg_domain.hpp), test
// infrastructure only.
//
// The reference's ExplicitSolverCH::solve loop (FE/ExplicitSolverCH.C:115-278)
// runs on the host in the oracle (oracle/liboracle.so, the checker): initStep
// and computePredictions, then Domain::computeInternalForces() -- served by
// the GPU through dnlg::Domain, from the host's coordinates and Current
// displacement increment -- then explicitSolve, updateGrid, endStep and
// computeTimeStep on the host.  A second oracle model runs the same loop with
// its own forces.  Every increment's Fint, coordinates and time scalars must
// be the same doubles, and so must the final point state.  A second part
// runs dnlg::ExplicitSolverCH::solve against the oracle's solve.
//
// usage: reference_loop axi|axi-rho|hex   (exit status 0 when every check passes)
#include "dnlg_domain.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
typedef struct dnlo dnlo;
dnlo *dnlo_create(int dim, long nn, long ne, const double *x0, const int *conn, const unsigned char *bc,
                  const double *v0);
void dnlo_set_material(dnlo *m, double E, double nu, double rho, double A, double B, double n, int yield);
void dnlo_set_solver(dnlo *m, double rho_b, double safety, int freq, double endTime);
void dnlo_set_mathmode(dnlo *m, int mode);
void dnlo_set_threads(int n);
int dnlo_init(dnlo *m);
int dnlo_begin(dnlo *m, double upTime);
int dnlo_step(dnlo *m);
long dnlo_solve(dnlo *m, double upTime);
int dnlo_begin_pre(dnlo *m, double upTime);
int dnlo_begin_post(dnlo *m, const double *fint);
void dnlo_step_pre(dnlo *m);
int dnlo_step_post(dnlo *m, const double *fint);
long dnlo_get(const dnlo *m, int field, double *out);
void dnlo_get_time(const dnlo *m, double *out);
double dnlo_courant_crit(dnlo *m);
void dnlo_free(dnlo *m);
}

namespace {
enum { O_X = 0, O_DU = 3, O_FINT = 5, O_S = 7, O_EVP = 8, O_P = 9 };

struct Model
{
  int dim = 0;
  long long nn = 0, ne = 0;
  std::vector<double> x, v0;
  std::vector<int> conn;
  std::vector<unsigned char> bc;
};

std::vector<double> oget(const dnlo *m, int f)
{
  const long n = dnlo_get(m, f, nullptr);
  std::vector<double> out(n > 0 ? n : 0);
  if (n > 0) dnlo_get(m, f, out.data());
  return out;
}

int failures = 0;
void fail(const std::string &what)
{
  fprintf(stdout, "MISMATCH: %s\n", what.c_str());
  fflush(stdout);
  failures++;
}

bool same(const std::vector<double> &a, const std::vector<double> &b) { return a.size() == b.size() && !memcmp(a.data(), b.data(), a.size() * sizeof(double)); }

dnlo *oracle(const Model &M, double rho_b)
{
  dnlo *o = dnlo_create(M.dim, (long)M.nn, (long)M.ne, M.x.data(), M.conn.data(), M.bc.data(), M.v0.data());
  dnlo_set_material(o, 115e9, 0.3, 8930.0, 400e6, 100e6, 1.0, 1);
  dnlo_set_solver(o, rho_b, 0.9, 10, 1e300);
  dnlo_set_mathmode(o, 1);  // the deterministic log/pow the GPU uses
  if (dnlo_init(o)) throw std::runtime_error("oracle init failed");
  return o;
}

void throwing_handler(const char *fn, const char *msg) { throw std::runtime_error(std::string(fn) + ": " + msg); }

// the reference's host loop with the GPU element pass, against the oracle's own loop
void run_loop(const Model &M, double rho_b, double up_time, long max_inc, const char *name)
{
  dnlg_mesh mesh{M.dim, M.nn, M.ne, M.x.data(), M.conn.data()};
  dnlg_material mat{};
  mat.young = 115e9; mat.poisson = 0.3; mat.density = 8930.0;
  mat.A = 400e6; mat.B = 100e6; mat.n = 1.0; mat.hardening = 1;
  dnlg_bc bc{M.bc.data(), M.v0.data(), nullptr, nullptr};
  dnlg::ExplicitSolverCH solver;
  solver.setTimeStepMethod("Courant");
  solver.setDissipation(rho_b);
  dnlg::Domain gpu(mesh, mat, bc, solver);

  dnlo *host = oracle(M, rho_b);  // the caller: the reference's loop, forces from the GPU
  dnlo *ref = oracle(M, rho_b);   // the oracle's own run
  // preamble, FE/ExplicitSolverCH.C:115-172
  if (dnlo_begin_pre(host, up_time) != 1 || dnlo_begin(ref, up_time) != 1) throw std::runtime_error("begin");
  std::vector<double> x = oget(host, O_X), du = oget(host, O_DU);
  gpu.bindNodes(x.data(), du.data());
  gpu.computeInternalForces();
  dnlo_begin_post(host, gpu.Fint.data());
  if (!same(gpu.Fint, oget(ref, O_FINT))) fail(std::string(name) + ": preamble Fint");
  long inc = 0;
  int run = 1;
  double th[9] = {0}, tr[9] = {0};
  while (run == 1 && inc < max_inc && !failures)
  {
    dnlo_step_pre(host);  // initStep + computePredictions
    x = oget(host, O_X);
    du = oget(host, O_DU);  // Current->delta_disp after the predictions' BCs
    gpu.bindNodes(x.data(), du.data());
    gpu.computeInternalForces();
    if (inc % 10 == 0)
    {
      // Domain::computeCourantTimeStep on the same coordinates
      const double cg = gpu.computeCourantTimeStep(), co = dnlo_courant_crit(host);
      if (cg != co) fail(std::string(name) + ": Courant crit at increment " + std::to_string(inc + 1));
    }
    run = dnlo_step_post(host, gpu.Fint.data());
    const int rr = dnlo_step(ref);
    inc++;
    if (run != rr) fail(std::string(name) + ": loop flag at increment " + std::to_string(inc));
    if (!same(gpu.Fint, oget(ref, O_FINT))) fail(std::string(name) + ": Fint at increment " + std::to_string(inc));
    if (!same(oget(host, O_X), oget(ref, O_X))) fail(std::string(name) + ": coordinates at increment " + std::to_string(inc));
    dnlo_get_time(host, th);
    dnlo_get_time(ref, tr);
    if (th[0] != tr[0] || th[1] != tr[1]) fail(std::string(name) + ": time at increment " + std::to_string(inc));
  }
  // the final point state: the GPU's S [9][points] and p [elements] against the oracle's [points][9], [points]
  const long long np = M.ne * (M.dim == 3 ? 8 : 4), nip = M.dim == 3 ? 8 : 4;
  std::vector<double> Sg = gpu.getField(DNLG_F_S), So = oget(ref, O_S);
  std::vector<double> St(Sg.size());
  for (long long i = 0; i < np; i++)
    for (int c = 0; c < 9; c++) St[9 * i + c] = Sg[c * np + i];
  if (!same(St, So)) fail(std::string(name) + ": final S");
  if (!same(gpu.getField(DNLG_F_EVP), oget(ref, O_EVP))) fail(std::string(name) + ": final evp");
  std::vector<double> pg = gpu.getField(DNLG_F_P), po = oget(ref, O_P), pe(np);
  for (long long i = 0; i < np; i++) pe[i] = pg[i / nip];
  if (!same(pe, po)) fail(std::string(name) + ": final p");
  printf("%s: host loop with GPU forces, %ld increments, t = %.6e s: Fint, coordinates, time, Courant crit and "
         "final S, evp, p %s\n", name, inc, th[0], failures ? "DIFFER" : "bit-equal to the oracle");
  fflush(stdout);
  dnlo_free(host);
  dnlo_free(ref);
}

// dnlg::ExplicitSolverCH::solve(upTime) against the oracle's solve
void run_solve(const Model &M, double rho_b, double up_time, const char *name)
{
  dnlg_mesh mesh{M.dim, M.nn, M.ne, M.x.data(), M.conn.data()};
  dnlg_material mat{};
  mat.young = 115e9; mat.poisson = 0.3; mat.density = 8930.0;
  mat.A = 400e6; mat.B = 100e6; mat.n = 1.0; mat.hardening = 1;
  dnlg_bc bc{M.bc.data(), M.v0.data(), nullptr, nullptr};
  dnlg::ExplicitSolverCH solver;
  solver.setTimeStepMethod("Courant");
  solver.setDissipation(rho_b);
  dnlg::Domain gpu(mesh, mat, bc, solver);
  dnlo *ref = oracle(M, rho_b);
  solver.solve(up_time);
  const long n = dnlo_solve(ref, up_time);
  if (n != solver.lastIncrements()) fail(std::string(name) + ": increment count of solve()");
  if (!same(gpu.getField(DNLG_F_COORDS), oget(ref, O_X))) fail(std::string(name) + ": coordinates after solve()");
  if (!same(gpu.getField(DNLG_F_EVP), oget(ref, O_EVP))) fail(std::string(name) + ": evp after solve()");
  printf("%s: ExplicitSolverCH::solve(%.3e), %ld increments: coordinates and evp %s\n", name, up_time, n,
         failures ? "DIFFER" : "bit-equal to the oracle");
  fflush(stdout);
  dnlo_free(ref);
}

Model axi_bar(int nr, int nz)
{
  Model M;
  M.dim = DNLG_DIM_AXI;
  M.nn = (long long)(nr + 1) * (nz + 1);
  M.ne = (long long)nr * nz;
  M.x.resize(M.nn * 3);
  M.v0.resize(M.nn * 3);
  M.conn.resize(M.ne * 4);
  M.bc.resize(M.nn);
  if (dnlg_mesh_axi_bar(nr, nz, 3.2e-3, 32.4e-3, 227.0, M.x.data(), M.conn.data(), M.bc.data(), M.v0.data()))
    throw std::runtime_error("mesh");
  return M;
}

Model ogrid_bar(int n, int m, int nz)
{
  Model M;
  M.dim = DNLG_DIM_3D;
  if (dnlg_mesh_ogrid_sizes(n, m, nz, &M.nn, &M.ne)) throw std::runtime_error("mesh sizes");
  M.x.resize(M.nn * 3);
  M.v0.resize(M.nn * 3);
  M.conn.resize(M.ne * 8);
  M.bc.resize(M.nn);
  if (dnlg_mesh_ogrid_bar(n, m, nz, 3.2e-3, 32.4e-3, 0.4, 227.0, M.x.data(), M.conn.data(), M.bc.data(), M.v0.data()))
    throw std::runtime_error("mesh");
  return M;
}
}  // namespace

int main(int argc, char **argv)
{
  const std::string which = argc > 1 ? argv[1] : "axi";
  dnlo_set_threads(1);
  try
  {
    // the facade's errors are the reference's fatalError texts; with a handler they can be caught
    dnlg::set_error_handler(throwing_handler);
    bool raised = false;
    try
    {
      dnlg::ExplicitSolverCH s;
      s.setDissipation(1.5);
    }
    catch (const std::runtime_error &e)
    {
      raised = std::string(e.what()).find("dissipation value out of range") != std::string::npos;
    }
    if (!raised) fail("setDissipation(1.5) did not raise the reference's error");

    if (which == "axi")
    {
      // the C0 bar: 10 x 50 quads, the whole 80 us at rho_b = 1
      run_loop(axi_bar(10, 50), 1.0, 80e-6, 1000000, "axi 10x50 rho_b=1");
      run_solve(axi_bar(10, 50), 1.0, 20e-6, "axi 10x50 rho_b=1");
    }
    else if (which == "axi-rho")
    {
      run_loop(axi_bar(10, 50), 0.8182, 80e-6, 1000000, "axi 10x50 rho_b=0.8182");
    }
    else if (which == "hex")
    {
      run_loop(ogrid_bar(4, 2, 12), 1.0, 1e300, 300, "hex8 O-grid 4+2 x 12 layers");
      run_solve(ogrid_bar(4, 2, 12), 1.0, 3e-6, "hex8 O-grid 4+2 x 12 layers");
    }
    else
    {
      fprintf(stderr, "usage: reference_loop axi|axi-rho|hex\n");
      return 2;
    }
  }
  catch (const std::exception &e)
  {
    fprintf(stdout, "ERROR: %s\n", e.what());
    return 1;
  }
  return failures ? 1 : 0;
}
