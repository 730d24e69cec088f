/*
 * dnlg_domain.hpp -- header-only C++ facade of libdnlgpu with the reference's
 * class and method names, so that a DynELA call site keeps its shape:
 *
 *   domain->computeInternalForces();   FE/Domain.h:103, called by
 *                                      ExplicitSolverCH::initIteration,
 *                                      FE/ExplicitSolverCH.C:471-481
 *   solver->solve(upTime);             FE/Solver.h:126, called by
 *                                      Domain::solve, FE/Domain.C:1288
 *
 * dnlg::Domain owns one handle of include/dnlg.h.  Two ways to use it:
 *
 *  - per call, inside the caller's own time loop: bindNodes() points the
 *    domain at the caller's packed Node::coords and Current->delta_disp
 *    ([n_nodes][3] doubles, node i = Node::Id i), and every
 *    computeInternalForces() sends them to the device, runs the element pass
 *    (the integration-point state stays in device memory, as it stays in the
 *    reference's IntegrationPoints) and fills Fint in the reference's layout
 *    ([Id * dim + j], FE/Domain.h:71);
 *  - whole solves on the device: dnlg::ExplicitSolverCH::solve(upTime) runs
 *    ExplicitSolverCH::solve (FE/ExplicitSolverCH.C:115-278) with the state
 *    resident; getField() reads it back.
 *
 * Errors go where the reference's go: fatalError (BT/Error.C:141-160) prints
 * "FATAL ERROR in function: ..." and exits with -1.  A caller that prefers an
 * exception installs dnlg::set_error_handler().
 */
#ifndef DNLG_DOMAIN_HPP
#define DNLG_DOMAIN_HPP

#include "dnlg.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace dnlg {

using ErrorHandler = void (*)(const char *function, const char *message);

inline ErrorHandler &error_handler_slot()
{
  static ErrorHandler h = nullptr;
  return h;
}
inline void set_error_handler(ErrorHandler h) { error_handler_slot() = h; }

// fatalError, BT/Error.C:141-160
inline void fatalError(const char *function, const char *message)
{
  if (error_handler_slot())
  {
    error_handler_slot()(function, message);
    return;
  }
  fprintf(stderr, "\nFATAL ERROR in function: %s\n", function);
  fprintf(stderr, "%s", message);
  fprintf(stderr, "\n");
  fprintf(stderr, "program aborted\n");
  exit(-1);
}

class Domain;

// The parameters ExplicitSolverCH and Solver keep (FE/Solver.h:66-121,
// FE/ExplicitSolverCH.h), set before the domain is built, as a deck sets them
// before struct.addSolver(solver); solve() runs on the domain it is added to.
class ExplicitSolverCH
{
public:
  ExplicitSolverCH() = default;

  // FE/ExplicitSolverCH.C:353-363
  void setDissipation(double val)
  {
    if (val < 0.0 || val > 1.0)
    {
      fatalError("ExplicitSolverCH::setDissipation", "dissipation value out of range [0.0, 1.0]");
      return;
    }
    rho_b = val;
  }
  // FE/Solver.C:277-284; the reference tests the value held before the
  // assignment, and so does this
  void setTimeStepSafetyFactor(double val)
  {
    if (safety <= 0)
    {
      fatalError("Solver::setTimeStepSafetyFactor", "negative or null safety factor value");
      return;
    }
    safety = val;
  }
  // FE/Solver.C:300-304
  void setComputeTimeStepFrequency(long val) { freq = (int)val; }
  // FE/Solver.C:129-133
  void setTimes(double startT, double stopT)
  {
    startTime = startT;
    endTime = stopT;
  }
  // FE/Solver.C:312-327
  void setTimeStepMethod(const char *meth)
  {
    if (!strcmp(meth, "Courant"))
      method = DNLG_TS_COURANT;
    else if (!strcmp(meth, "Pulse"))
      method = DNLG_TS_PULSE;
    else
    {
      std::string m = std::string("Unknown method ") + meth + " \n";
      fatalError("setTimeStepMethod", m.c_str());
    }
  }
  // the arithmetic form of the element pass (no reference counterpart)
  void setForm(int f) { form = f; }

  // Solver::solve(upTime), FE/Solver.h:126, as ExplicitSolverCH::solve
  inline void solve(double upTime);
  // increments run by the last solve()
  long long lastIncrements() const { return last_steps; }

  dnlg_solver_params params() const
  {
    dnlg_solver_params p;
    p.rho_b = rho_b;
    p.safety = safety;
    p.ts_freq = freq;
    p.end_time = endTime;
    p.form = form;
    return p;
  }

  Domain *domain = nullptr;
  // ExplicitSolverCH's default dissipation is 0.8182 (FE/DefaultValues.h:25) and
  // its time step Pulse (FE/ExplicitSolverCH.C:83); safety 0.9 and frequency 10
  // (FE/Solver.C:45-76)
  double rho_b = 0.8182, safety = 0.9, startTime = 0.0, endTime = 1e300;
  int freq = 10, method = DNLG_TS_PULSE, form = DNLG_FORM_FP64;

private:
  long long last_steps = 0;
};

class Domain
{
public:
  // Structure::addSolver + Domain::initSolve (FE/Structure.C:586,
  // FE/Domain.C:1018-1226): the mesh, material and BCs go to the device
  Domain(const dnlg_mesh &mesh, const dnlg_material &mat, const dnlg_bc &bc, ExplicitSolverCH &solver)
      : numberOfDim(mesh.dim), n_nodes(mesh.n_nodes), n_elems(mesh.n_elems)
  {
    if (solver.startTime != 0.0)
    {
      // Solver::betweenTimeBounds (FE/Solver.C:169-176) with the default start time 0 only
      fatalError("Solver::setTimes", "a start time other than 0 is not supported by libdnlgpu");
      return;
    }
    const dnlg_solver_params sp = solver.params();
    if (dnlg_create(&mesh, &mat, &bc, &sp, &h) != DNLG_OK)
    {
      fatalError("Domain::initSolve", dnlg_last_error(nullptr));
      return;
    }
    if (solver.method != DNLG_TS_COURANT) check(dnlg_set_option(h, DNLG_OPT_TIME_STEP_METHOD, solver.method),
                                                "Solver::setTimeStepMethod");
    solver.domain = this;
    Fint.assign((size_t)(n_nodes * numberOfDim), 0.0);
  }
  ~Domain()
  {
    if (h) dnlg_destroy(h);
  }
  Domain(const Domain &) = delete;
  Domain &operator=(const Domain &) = delete;

  // the caller's packed Node::coords and Current->delta_disp, [n_nodes][3];
  // read by every computeInternalForces() until bound again (NULL: keep the
  // device copy, e.g. while the device runs the loop itself)
  void bindNodes(const double *coords, const double *currentDeltaDisp)
  {
    node_coords = coords;
    node_delta_disp = currentDeltaDisp;
  }

  // Domain::computeInternalForces(), FE/Domain.h:103
  void computeInternalForces()
  {
    if (node_coords || node_delta_disp)
      check(dnlg_compute_internal_forces_from(h, node_coords, node_delta_disp, Fint.data()),
            "Domain::computeInternalForces");
    else
      check(dnlg_compute_internal_forces(h, Fint.data()), "Domain::computeInternalForces");
  }

  // Domain::computeCourantTimeStep(), FE/Domain.h:128
  double computeCourantTimeStep()
  {
    double crit = 0.0;
    check(dnlg_courant_time_step(h, &crit), "Domain::computeCourantTimeStep");
    return crit;
  }

  // a state field (DNLG_F_*) in the layout include/dnlg.h gives
  std::vector<double> getField(int field)
  {
    const long long n = dnlg_field_size(h, field);
    std::vector<double> out(n > 0 ? (size_t)n : 0);
    if (n > 0) check(dnlg_get_field(h, field, out.data(), n), "Domain::getField");
    return out;
  }
  void setField(int field, const std::vector<double> &in)
  {
    check(dnlg_set_field(h, field, in.data(), (long long)in.size()), "Domain::setField");
  }

  // Domain::currentTime and the time-step state: the 9 values of dnlg_get_time
  std::vector<double> getTime()
  {
    std::vector<double> t(9, 0.0);
    check(dnlg_get_time(h, t.data()), "Domain::getTime");
    return t;
  }

  dnlg_handle *handle() { return h; }

  void check(int rc, const char *function)
  {
    if (rc != DNLG_OK) fatalError(function, dnlg_last_error(h));
  }

  std::vector<double> Fint;  // Domain::Fint, FE/Domain.h:71: [Id * dim + j]
  const int numberOfDim;
  const long long n_nodes, n_elems;

private:
  dnlg_handle *h = nullptr;
  const double *node_coords = nullptr, *node_delta_disp = nullptr;
};

inline void ExplicitSolverCH::solve(double upTime)
{
  if (!domain)
  {
    fatalError("ExplicitSolverCH::solve", "no domain: the solver has not been added to one");
    return;
  }
  long long n = 0;
  domain->check(dnlg_solve(domain->handle(), upTime, &n), "ExplicitSolverCH::solve");
  last_steps = n;
}

}  // namespace dnlg

#endif /* DNLG_DOMAIN_HPP */
