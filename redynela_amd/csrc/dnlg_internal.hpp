// Internal shared definitions of libdnlgpu: device-resident step state,
// kernel argument blocks and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dnl {

// p as a pointer into global memory.  Every array the kernels touch is device
// global memory, but a pointer read from memory (the per-bar arguments of the
// batched kernels) reaches the compiler as a generic one, and the flat loads
// and stores it then emits also count against the LDS wait counter.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p)
{
  return (__attribute__((address_space(1))) T *)p;
}

// device status codes (mirror include/dnlg.h)
enum : int { ST_OK = 0, ST_POLAR = 2, ST_RR = 3, ST_DETJ = 4, ST_BUDGET = 6, ST_PULSE = 7 };

// Time-integration scalars kept on the device so that a step is launch-only.
// Mirrors Solver / ExplicitSolverCH / Domain scalar state:
// FE/Solver.C:45-76, FE/ExplicitSolverCH.C:336-343, FE/Domain.h (currentTime).
struct StepState
{
  double t;          // Domain::currentTime
  double dt;         // Solver::timeStep
  double upTime;     // Solver::upTime
  double endTime;    // Solver::endTime
  double computedTimeStep, tsDate, lastTsDate, lastTimeStep;
  double alphaM, beta, gamma, omegaS, safety;
  double cs;         // Element::getSoundSpeed (material constant)
  double ke;         // Domain::getTotalKineticEnergy of the last step
  long long increment;
  int freq;          // computeTimeStepFrequency
  int running;       // betweenTimeBounds at the end of the last step
  int status;        // first device error
  int err_elem;
  int pending;       // the last issued increment still needs its endStep
  int active;        // batched stepped engine: this increment runs its nodal phase
  int ts_method;     // Solver::timeStepMethod: 0 Courant, 1 Pulse
  int plot_cap;      // rows of the .plot ring (0: no ring)
  double pulse_f;    // Domain::getMaximumFrequency of the last Pulse computation
  double *plot;      // [plot_cap][4] ring of endStep history lines (t, dt, computed dt, KE) or null
  long long plot_n;  // lines written so far
  int halted;        // single-domain path: an element error has been seen by k_dt; element passes skip
  int pad_;
};

struct ElemArgs
{
  const int *conn;       // [ne][nen]
  const double *x;       // [nn][3]
  const double *du;      // [nn][3] Current delta_disp
  double *S;             // [9][ne*nip] deviatoric stress (full tensor)
  double *evp;           // [ne*nip]
  double *p;             // [ne] pressure (equal at every IP of an element)
  double *eps;           // [9][ne*nip] or null
  double *epsplas;       // [9][ne*nip] or null
  double *deps;          // [9][ne*nip] or null: IntegrationPoint::dEps of the last pass
  double *fe;            // [maxdeg][nn][dim] element nodal forces, slot-major per node (ELL)
  const int *fpos;       // [ne*nen] slot*nn + node of each (element, local node)
  const unsigned char *bcs;  // speed-BC component mask per node or null
  const double *bcs_v;       // [nn][3]
  StepState *st;
  long long ne;
  double K, G, A, B, n, nm1;
  int yield;
  int check_running;
  // Johnson-Cook rate / thermal extension (law 2); T is the point temperature
  int law;               // 1: A + B evp^n (reference), 2: Johnson-Cook with rate and adiabatic heating
  double C, epsd0, m, T0, Tm, chi_rho_cp;
  double *T;             // [ne*nip] or null
  // shape constants (closed forms of the 2-point Gauss tables)
  double c0, c1, c2;     // hex8 dN0 magnitudes (2-r3)/12, 1/12, (2+r3)/12 ; quad4 (3-r3)/12, (3+r3)/12
  double n0, n1, n2, n3; // quad4 N: (2-r3)/6, 1/6, (2+r3)/6 ; hex8 N not needed here
};

// mixed-form element pass arguments (FP32 point state)
struct ElemArgs32
{
  const int *conn;
  const double *x, *du;
  float *S;      // [6][nip][ne]  (xx, yy, zz, xy, yz, xz), point-major per component
  float *evp;    // [nip][ne]
  float *T;      // [nip][ne] temperature (Johnson-Cook thermal softening) or null
  float *p;      // [ne]
  float *fe;     // [maxdeg][nn][dim] (ELL, see ElemArgs::fe)
  const int *fpos;
  StepState *st;
  long long ne;
  float K, G, A, B, n;
  // Johnson-Cook rate / thermal extension (law == 2)
  float C, epsd0, m, T0, Tm, chi_rho_cp;  // chi / (rho cp)
  int law;       // 0 elastic, 1 A + B evp^n, 2 Johnson-Cook with rate and adiabatic heating
  int check_running;
};

// fused nodal step arguments
struct NodalArgs
{
  double *x, *v, *a, *du, *d;
  const double *mass;
  const unsigned char *bc, *bcs;
  const double *bcs_v;
  const unsigned char *deg; // elements per node; slot s < deg holds the s-th element in element order
  const double *fe;        // FP64 element forces (FP64 form)
  const float *fe32;       // FP32 element forces (mixed form)
  double *fint;     // optional Domain::Fint copy
  double *ke_part;  // optional per-block kinetic energy
  StepState *st;
  long long nn;
};

// one domain of a batch (packed into one device mesh)
struct BarDesc
{
  long long node0, nn, elem0, ne;
  double K, G, A, B, n, nm1;
  int law, pad;
};

struct BatchArgs
{
  const BarDesc *bars;
  const int *order;          // block b runs bar order[b] (largest first)
  StepState *states;         // one per bar
  ElemArgs ea;               // packed arrays, FP64 form
  ElemArgs32 ea32;           // packed arrays, mixed form
  NodalArgs na;              // packed arrays
  unsigned long long *t0, *t1;  // wall clock at the start / end of each bar
  long long max_steps;
  unsigned long long *crit;  // stepped engine: per-bar Courant minimum (bits of a positive double)
  const ElemArgs *bar_ea;    // per bar: ea with the bar's state pointer and material constants
  const ElemArgs32 *bar_ea32;
};

// adjoint of one increment (dnlg_adjoint.hip): scalar adjoints and kernel arguments
struct AdjScalars
{
  double tb, dtb_carry, tsb, ltsb, Cb, Lb;  // adjoints of t, dt (carried), tsDate, lastTsDate, computed, last time step
  double th[3];                             // accumulated dL/dA, dL/dB, dL/dn
};

struct AdjArgs
{
  ElemArgs ea;                      // shape constants, conn, fpos, ne (other fields unused)
  const unsigned char *bc, *deg;
  const double *mass;
  long long nn, ne;
  // forward tape: state after increment k (x, v, a, du: [nn][3]; S [9][nip ne]; evp [nip ne]; p [ne]) and a_{k+1}
  const double *x, *v, *a, *du, *S, *evp, *p, *a1;
  const StepState *st0, *st1;       // time state after increments k and k+1
  // adjoint state, updated in place from increment k+1 to k
  double *xb, *ub, *vb, *ab;        // [nn][dim]
  double *Sb, *evb, *pb;            // [9][nip ne], [nip ne], [ne]
  double *fb;                       // [nn][dim] adjoint of the assembled internal force
  double *slot_x, *slot_u;          // [maxdeg][nn][dim] per (element, node) contributions, element order
  double *dt_part, *th_part;        // per-block partial sums
  int ndt_part, nth_part;
  double *am_len;                   // per-block Courant arg-min of x_k: length / cs and element
  long long *am_idx;
  int nam_part;
  AdjScalars *sc;
  double K, G, A, B, n;
  double alphaM, beta, gamma, safety, omegaS, cs;
  int freq;
  long long k;                      // this launch reverses increment k+1
  // Johnson-Cook temperature factor and adiabatic heating (law 2 without the
  // rate term): point temperature after increment k and its adjoint; the
  // factor's constants are ea.T0, ea.Tm, ea.m, ea.chi_rho_cp
  int jc;
  const double *T;                  // [nip ne] or null
  double *Tb;                       // [nip ne] or null
  int dim;                          // 2: axisymmetric quad4 (4 points), 3: hex8 (8 points)
};

}  // namespace dnl
