// HistoryFile nodal series (FE/HistoryFile.C:114-141, FE/Node.C:326-418) on
// the device: per file a ring of rows (time, one value per item).
#pragma once
#include <hip/hip_runtime.h>
#include "dnlg_internal.hpp"

namespace dnl {

// Node::getNodalValue field names (FE/Node.C:328-342), the subset this path keeps
enum : int
{
  HF_DENSITY = 0, HF_DENSITY_INC, HF_ENERGY, HF_ENERGY_INC, HF_SPEED, HF_SPEED_INC, HF_FORCE, HF_TEMPERATURE,
  HF_DISPLACEMENT_INC, HF_DISPLACEMENT, HF_COORDS, HF_MASS, HF_TEMPERATURE_INIT,
  // element fields, extrapolated from the integration points to the node (FE/Node.C:350-418)
  HF_STRESS, HF_STRAIN, HF_STRAIN_INC, HF_PLASTIC_STRAIN, HF_NFIELDS
};

struct HistFile
{
  double next, every, stop;  // HistoryFile nextTime, saveTime, stopTime
  double *ring;              // [cap][1 + count]
  long long rows;            // rows written so far
  int first, count;          // its items
  int cap, due;              // ring rows; this increment writes a row
};

struct HistItem
{
  int node, field, comp;
  int el_off, el_cnt;  // element fields: the node's (element, local node) pairs in element order
  int pad;
};

struct HistArgs
{
  HistFile *files;
  const HistItem *items;
  int nfiles, nitems;
  const double *x, *v, *a, *du, *d, *mass;
  StepState *st;
  double density, T0;  // the Current NodalField's ro and T before its first swap
  // element fields (FP64 point state, point-major [9][ne * nip], p per element)
  const int2 *hel;     // (element, local node) pairs of the items
  const double *S, *p, *eps, *epsplas, *deps;
  long long nipt;
  int nip, nen;
  // ElementData::nodes[loc].integrationPointsToNode(pt) of the element family
  // (FE/ElHex8n3D.C:41-72, FE/ElQua4nAx.C:45-48), [loc][pt]
  double coef[8][8];
};

void launch_hist_pre(const HistArgs &h, hipStream_t s);
void launch_hist_post(const HistArgs &h, hipStream_t s);

}  // namespace dnl
