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
//   k_hist_post (after the nodal pass): the moved coordinates and the
//               element fields (the point state of this increment), then
//               advance nextTime and the row count of the files that wrote.
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

// Node::getNodalValue of an element field (FE/Node.C:350-418): each element
// of the node gives get_IntegrationValue_atNode (FE/Element.C:1180-1198), the
// sum over the node's elements of coef(loc, pt) * (the point's tensor), over
// their count; Node::getNodalValue adds that same tensor once per element of
// the node and divides by the count again.  Component 0: the von Mises
// equivalent of a stress (Tensor2::misesEquivalent, BT/Math.h:87-88) or the
// norm sqrt(2/3 s:s) of a strain (Tensor2::equivalent, LA/Tensor2.h:359-372);
// 10 i + j: the entry (i, j), 1-based.
__device__ __forceinline__ double element_field_value(const HistArgs &h, const HistItem &it)
{
#pragma clang fp contract(off)
  double V[9];
#pragma unroll
  for (int c = 0; c < 9; c++) V[c] = 0.0;
  const double *src = it.field == HF_STRAIN ? h.eps : it.field == HF_STRAIN_INC ? h.deps
                      : it.field == HF_PLASTIC_STRAIN ? h.epsplas : h.S;
  for (int j = 0; j < it.el_cnt; j++)
  {
    const int2 el = h.hel[it.el_off + j];
    for (int pt = 0; pt < h.nip; pt++)
    {
      const long long ip = (long long)el.x * h.nip + pt;
      const double c = h.coef[el.y][pt];
      const double pe = it.field == HF_STRESS ? h.p[el.x] : 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++)
      {
        // Sig = S + p I as the element pass forms it
        const double t = it.field == HF_STRESS ? src[k * h.nipt + ip] + pe * ((k % 4 == 0) ? 1.0 : 0.0)
                                               : src[k * h.nipt + ip];
        V[k] += c * t;
      }
    }
  }
  const double nb = (double)it.el_cnt;
  double V1[9];
#pragma unroll
  for (int k = 0; k < 9; k++)
  {
    const double v = V[k] / nb;
    double s = 0.0;
    for (int j = 0; j < it.el_cnt; j++) s += v;
    V1[k] = s / nb;
  }
  if (it.comp == 0)
  {
    if (it.field == HF_STRESS)
    {
      const double a = V1[0] - V1[4], b = V1[0] - V1[8], c = V1[4] - V1[8];
      return sqrt(((a * a) + (b * b) + (c * c) + 6.0 * ((V1[1] * V1[1]) + (V1[2] * V1[2]) + (V1[5] * V1[5]))) / 2.0);
    }
    return sqrt(2. * (V1[0] * V1[0] + V1[1] * V1[1] + V1[2] * V1[2] + V1[3] * V1[3] + V1[4] * V1[4] +
                      V1[5] * V1[5] + V1[6] * V1[6] + V1[7] * V1[7] + V1[8] * V1[8]) / 3.);
  }
  return V1[3 * (it.comp / 10 - 1) + (it.comp % 10 - 1)];
}

__global__ __launch_bounds__(256) void k_hist_post(HistArgs h)
{
#pragma clang fp contract(off)
  StepState *st = h.st;
  if (!st->running || st->status) return;
  for (int k = threadIdx.x; k < h.nitems; k += 256)
  {
    const HistItem it = h.items[k];
    if (it.field != HF_COORDS && it.field < HF_STRESS) continue;
    int f = 0;
    while (f + 1 < h.nfiles && h.files[f + 1].first <= k) f++;
    const HistFile &F = h.files[f];
    if (!F.due) continue;
    double *row = F.ring + (F.rows % F.cap) * (1 + F.count);
    row[1 + (k - F.first)] = it.field == HF_COORDS ? vec_value(h.x + 3 * (long long)it.node, it.comp)
                                                   : element_field_value(h, it);
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
