// Pulse time step (Domain::getMaximumFrequency) device state and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include "dnlg_internal.hpp"

namespace dnl {

constexpr int PULSE_MAX_ITER = 1000;          // maximumFrequencyMaxIterations, FE/DefaultValues.h:35
constexpr double PULSE_CONVERGENCE = 1e-4;    // maximumFrequencyConvergence, FE/DefaultValues.h:36

// power-iteration state kept in device memory from one call to the next
struct PulseCtl
{
  double fmax0;  // Domain::fmax0 (FE/Domain.h:66), 0 until the first call
  double invf;   // 1 / fmax of the last iteration (1 for the start vector)
  double fmax, conv;
  int iter, done, failed, pad;
};

struct PulseArgs
{
  const int *conn;            // [ne][nen]
  const int *fpos;            // [ne*nen] ELL slot*nn + node
  const unsigned char *deg;   // [nn] slots per node
  const double *mass;         // [nn] lumped mass
  const double *ev_init;      // [nn*dim] start vector ev(0) = 1, ev(i) = ev(i-1) - 2/(noddl-1)
  double *ke;                 // [ceil(ne/8)][nd/2][8*nd][2]: column pairs of the rows of 8 elements
  double *y;                  // [maxdeg][nn][dim] element products Ke x, ELL slots
  double *val;                // [nn*dim] iterate before the 1/fmax scaling
  double *part;               // [ceil(nn/256)] block maxima
  PulseCtl *ctl;
  StepState *st;
  long long ne, nn, ndof;
  double E, nu;
};

void launch_pulse_begin(const PulseArgs &p, hipStream_t s);
void launch_pulse_ke(int dim, const PulseArgs &p, const ElemArgs &a, hipStream_t s);
void launch_pulse_iteration(int dim, const PulseArgs &p, hipStream_t s);

}  // namespace dnl
