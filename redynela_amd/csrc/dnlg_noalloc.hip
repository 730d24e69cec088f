// Domain::computeCourantTimeStep for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops): the same two
// kernels as dnlg_courant_time_step, with the per-block minima in a
// file-scope device array instead of a handle's buffer.
//
// Each GPU that runs this holds its own copy of `scratch` until the process
// ends.  Its address on a device is taken once, with that device current, at
// the entry point's first call there, and host code uses only that address:
// the array's name in host code is a host-side stand-in, not a device address.
// The partial minima of a call overwrite the first npart entries and the
// reduction reads only those, so nothing needs clearing between calls.  Calls
// that overlap on one GPU (other streams, other threads) share the array and
// race.
#include "dnlg_internal.hpp"
#include "../../include/dnlg.h"

#include <cmath>
#include <mutex>
#include <string>

namespace dnl {
void launch_courant_f64(int dim, const ElemArgs &a, double cs, double *part, hipStream_t s);
void launch_min_reduce(const double *part, int npart, double *out, hipStream_t s);
unsigned courant_blocks(long long ne);
}  // namespace dnl

using namespace dnl;

static constexpr long long kScratchDoubles = 65536;  // 512 KiB: block minima of up to 65536 * 256 elements
static constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_mu;
double *g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device
thread_local std::string g_noalloc_err;

// the device address of this device's `scratch` (current device), taken once
int scratch_address(int dev, double **out)
{
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev])
  {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
    {
      g_noalloc_err = std::string("dnlg_courant_time_step_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e);
      return DNLG_E_HIP;
    }
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return DNLG_OK;
}
}  // namespace

extern "C" const char *dnlg_noalloc_last_error(void) { return g_noalloc_err.c_str(); }

extern "C" int dnlg_courant_time_step_noalloc(int dim, long long n_elems, const double *coords, const int *conn,
                                              const dnlg_material *mat, double *crit, void *stream)
{
  g_noalloc_err.clear();
  if ((dim != DNLG_DIM_AXI && dim != DNLG_DIM_3D) || n_elems <= 0 || !coords || !conn || !mat || !crit)
  {
    g_noalloc_err = "dnlg_courant_time_step_noalloc: bad arguments";
    return DNLG_E_ARG;
  }
  if (n_elems > kScratchDoubles * 256)
  {
    g_noalloc_err = "dnlg_courant_time_step_noalloc: more than 16777216 elements (the scratch holds 65536 block minima)";
    return DNLG_E_ARG;
  }
  if (!(mat->young > 0 && mat->density > 0 && mat->poisson > -1.0 && mat->poisson < 0.5))
  {
    g_noalloc_err = "dnlg_courant_time_step_noalloc: material constants out of range";
    return DNLG_E_ARG;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices)
  {
    g_noalloc_err = "dnlg_courant_time_step_noalloc: no usable current device";
    return DNLG_E_HIP;
  }
  double *part = nullptr;
  const int rc = scratch_address(dev, &part);
  if (rc) return rc;
  // Element::getSoundSpeed with every nodal density the material's, the
  // expression dnlg_create evaluates (FE/Element.C:607-621)
  const int nen = dim == DNLG_DIM_3D ? 8 : 4;
  const double E = mat->young, Nu = mat->poisson;
  double ro = 0.;
  for (int i = 0; i < nen; i++) ro += mat->density;
  ro /= nen;
  const double cs = sqrt((E * (1 - Nu)) / (ro * (1 + Nu) * (1 - 2 * Nu)));
  ElemArgs a{};
  a.conn = conn;
  a.x = coords;
  a.ne = n_elems;
  if (dim == DNLG_DIM_3D)
  {
    // closed forms of the hex8 dN0 magnitudes (FE/ElHex8n3D.C:81-87), as dnlg_create sets them
    a.c0 = (2.0 - sqrt(3.0)) / 12.0;
    a.c1 = 1.0 / 12.0;
    a.c2 = (2.0 + sqrt(3.0)) / 12.0;
  }
  const hipStream_t s = (hipStream_t)stream;
  launch_courant_f64(dim, a, cs, part, s);
  launch_min_reduce(part, (int)courant_blocks(n_elems), crit, s);
  e = hipGetLastError();
  if (e != hipSuccess)
  {
    g_noalloc_err = std::string("dnlg_courant_time_step_noalloc: ") + hipGetErrorString(e);
    return DNLG_E_HIP;
  }
  return DNLG_OK;
}
