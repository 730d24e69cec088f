// Host-side definition of the single-domain handle (dnlg_handle of
// include/dnlg.h), shared by the C ABI translation units.
#pragma once
#include "../../include/dnlg.h"
#include "dnlg_internal.hpp"
#include "dnlg_history.hpp"

#include <string>
#include <vector>

namespace dnl {
struct AdjStore;  // checkpoints and tape of the adjoint (dnlg_adjoint_api.cpp)
struct PulseCtl;  // power-iteration state (dnlg_pulse.hpp)
}  // namespace dnl
using namespace dnl;

struct dnlg_handle
{
  int dim = 0, nen = 0, nip = 0;
  long long nn = 0, ne = 0;
  hipStream_t stream = nullptr;
  // device buffers
  double *x = nullptr, *v = nullptr, *a = nullptr, *du = nullptr, *d = nullptr, *mass = nullptr;
  double *fint = nullptr, *fe = nullptr, *S = nullptr, *evp = nullptr, *p = nullptr;
  double *eps = nullptr, *epsplas = nullptr, *deps = nullptr, *me = nullptr, *part = nullptr, *ke_part = nullptr;
  double *bcs_v = nullptr;
  double *crit = nullptr;  // Domain::computeCourantTimeStep result of dnlg_courant_time_step
  double cs = 0.0;         // Element::getSoundSpeed of the material (the Courant kernel's argument)
  int *conn = nullptr, *fpos = nullptr;
  unsigned char *deg = nullptr;
  int maxdeg = 0;
  std::vector<int> fpos_host;  // (element, local node) -> slot*nn + node
  unsigned char *bc = nullptr, *bcs = nullptr;
  StepState *st = nullptr;
  bool has_bcs = false;
  int form = DNLG_FORM_FP64;
  double *T = nullptr;                       // FP64 point temperature (hardening 2)
  float *S32 = nullptr, *evp32 = nullptr, *T32 = nullptr, *p32 = nullptr, *fe32 = nullptr;  // mixed form
  ElemArgs32 ea32{};
  // options
  bool ke_on = false, store_fint = true, track_d = true, track_strain = false;
  // host-side sequencing
  bool mass_done = false, pending = false, begun = false;
  long long inc = 0;
  int freq = 10;
  dnlg_material mat{};
  double hn[4] = {0, 0, 0, 0};
  ElemArgs ea{};
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_elem, ev_nodal;  // start/stop pairs
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  AdjStore *adj = nullptr;  // checkpoints of dnlg_adjoint_forward, freed by dnlg_destroy
  // Pulse time step (Domain::getMaximumFrequency), allocated at its first use
  int ts_method = 0;
  double *pulse_ke = nullptr, *pulse_y = nullptr, *pulse_val = nullptr, *pulse_init = nullptr,
         *pulse_part = nullptr;
  // endStep history lines (.plot): device ring, lines drained to the host at every synchronisation
  double *plot_ring = nullptr;
  long long plot_drained = 0;
  std::vector<double> plot_host;  // [rows][4] not yet read by dnlg_plot_rows
  long long drain_inc = 0;        // increment count at the last drain of the rings
  // HistoryFile series: device tables and rings, rows drained per file
  std::vector<HistFile> hf;       // host copy of the tables (ring pointers, first/count)
  std::vector<HistItem> hitems;
  HistFile *hf_dev = nullptr;
  HistItem *hi_dev = nullptr;
  std::vector<int2> hel;          // element fields: (element, local node) pairs of the items
  int2 *hel_dev = nullptr;
  std::vector<long long> hf_drained;
  std::vector<std::vector<double>> hf_rows;
  double init_temp = 0.0;
  PulseCtl *pulse_ctl = nullptr;
  long long pulse_calls = 0, pulse_iters = 0, pulse_iters_last = 0;
  std::string err;
};

#define DNLG_HIP(h, call)                                                                        \
  do                                                                                             \
  {                                                                                              \
    hipError_t _e = (call);                                                                      \
    if (_e != hipSuccess)                                                                        \
    {                                                                                            \
      (h)->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #call;               \
      return DNLG_E_HIP;                                                                         \
    }                                                                                            \
  } while (0)

template <typename T>
inline hipError_t dalloc(T **p, size_t n)
{
  return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

