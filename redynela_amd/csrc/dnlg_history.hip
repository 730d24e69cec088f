// HistoryFile::save on the device (FE/HistoryFile.C:114-141), called from
// ExplicitSolverCH::endStep through Domain::writeHistoryFiles
// (FE/ExplicitSolverCH.C:463-464, FE/Domain.C:1515-1522).
//
// At endStep the time is already updated and the node coordinates moved,
// but the swap of the nodal fields comes after: the values Node::getNodalValue
// reads from Current (FE/Node.C:288-297) are those of the previous increment.
// The fused nodal kernel overwrites them in place, so the series are taken
// in two launches around it:
//   k_hist_pre  (after initStep, before the element pass): decide which files
//               write a row at tm = t + dt, the time endStep will reach, and
//               copy the Current-field values;
//   k_hist_post (after the nodal pass): the moved coordinates, then advance
//               nextTime and the row count of the files that wrote.
#include "dnlg_history.hpp"

namespace dnl {

__device__ __forceinline__ double vec_value(const double *v, int comp)
{
#pragma clang fp contract(off)
  if (comp == 0) return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);  // Vec3D::norm, LA/Vec3D.h:380-384
  return v[comp - 1];
}

__global__ __launch_bounds__(256) void k_hist_pre(HistArgs h)
{
#pragma clang fp contract(off)
  StepState *st = h.st;
  if (!st->running || st->status) return;
  const double tm = st->t + st->dt;  // the endStep update of this increment
  for (int f = threadIdx.x; f < h.nfiles; f += 256)
  {
    HistFile &F = h.files[f];
    int due = 1;
    if ((F.stop > 0) && (tm >= F.stop + F.every)) due = 0;
    if (tm < F.next) due = 0;
    F.due = due;
    if (due) F.ring[(F.rows % F.cap) * (1 + F.count)] = tm;
  }
  __syncthreads();
  // the Current field before its swap: the original one after an even number of swaps
  const bool original = ((st->increment - 1) % 2) == 0;
  for (int k = threadIdx.x; k < h.nitems; k += 256)
  {
    const HistItem it = h.items[k];
    int f = 0;
    while (f + 1 < h.nfiles && h.files[f + 1].first <= k) f++;
    const HistFile &F = h.files[f];
    if (!F.due) continue;
    double *row = F.ring + (F.rows % F.cap) * (1 + F.count);
    const long long n3 = 3 * (long long)it.node;
    const double zero3[3] = {0.0, 0.0, 0.0};
    double val = 0.0;
    switch (it.field)
    {
    case HF_DENSITY: val = original ? h.density : 0.0; break;
    case HF_TEMPERATURE: val = original ? h.T0 : 0.0; break;
    case HF_DENSITY_INC: case HF_ENERGY: case HF_ENERGY_INC: val = 0.0; break;
    case HF_SPEED: val = vec_value(h.v + n3, it.comp); break;
    case HF_SPEED_INC: val = vec_value(h.a + n3, it.comp); break;
    case HF_FORCE: val = vec_value(zero3, it.comp); break;
    case HF_DISPLACEMENT_INC: val = vec_value(h.du + n3, it.comp); break;
    case HF_DISPLACEMENT: val = vec_value(h.d + n3, it.comp); break;
    case HF_MASS: val = h.mass[it.node]; break;
    case HF_TEMPERATURE_INIT: val = h.T0; break;
    default: continue;  // coordinates: after the nodal pass
    }
    row[1 + (k - F.first)] = val;
  }
}

__global__ __launch_bounds__(256) void k_hist_post(HistArgs h)
{
#pragma clang fp contract(off)
  StepState *st = h.st;
  if (!st->running || st->status) return;
  for (int k = threadIdx.x; k < h.nitems; k += 256)
  {
    const HistItem it = h.items[k];
    if (it.field != HF_COORDS) continue;
    int f = 0;
    while (f + 1 < h.nfiles && h.files[f + 1].first <= k) f++;
    const HistFile &F = h.files[f];
    if (!F.due) continue;
    double *row = F.ring + (F.rows % F.cap) * (1 + F.count);
    row[1 + (k - F.first)] = vec_value(h.x + 3 * (long long)it.node, it.comp);
  }
  __syncthreads();
  for (int f = threadIdx.x; f < h.nfiles; f += 256)
  {
    HistFile &F = h.files[f];
    if (!F.due) continue;
    F.next += F.every;
    F.rows++;
    F.due = 0;
  }
}

void launch_hist_pre(const HistArgs &h, hipStream_t s)
{
  hipLaunchKernelGGL(k_hist_pre, dim3(1), dim3(256), 0, s, h);
}

void launch_hist_post(const HistArgs &h, hipStream_t s)
{
  hipLaunchKernelGGL(k_hist_post, dim3(1), dim3(256), 0, s, h);
}

}  // namespace dnl
