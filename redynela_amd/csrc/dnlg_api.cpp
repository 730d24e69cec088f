// C ABI of libdnlgpu: handle life cycle, device memory, step issuing.
//
// The handle holds one Domain of the reference (FE/Domain.h:59-132) on one
// GPU: SoA/AoS device arrays, the node -> (element, local node) CSR used by
// the deterministic gathers, and the device-resident time-step state.  The
// call sequence of ExplicitSolverCH::solve (FE/ExplicitSolverCH.C:115-278) is
// issued as kernels on the handle's stream:
//   preamble  : [mass once] element pass, Fint gather, Courant + forced dt
//   each step : [Courant when the previous increment % freq == 0] k_dt
//               (endStep of the previous increment, computeTimeStep, initStep)
//               element kernel, fused nodal kernel
// With the Pulse method a recomputing step splits k_dt around the power
// iteration (endStep; Domain::getMaximumFrequency; computeTimeStep and
// initStep), and the host waits for the iteration to converge.
// and a final k_dt at every synchronisation point closes the last increment.
#include "dnlg_handle.hpp"
#include "dnlg_pulse.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace dnl {
void launch_elem_f64(int dim, const ElemArgs &a, bool speed_bc, hipStream_t s);
void launch_mass_f64(int dim, const ElemArgs &a, double rho, const double *hn, double *me, hipStream_t s);
void launch_gather_scalar(const unsigned char *deg, const double *src, double *dst, long long nn, hipStream_t s);
void launch_gather_fint(int dim, const unsigned char *deg, const double *fe, const float *fe32, double *fint,
                        long long nn, hipStream_t s);
void launch_elem_f32(int dim, const ElemArgs32 &a, hipStream_t s);
void launch_nodal_f64(int dim, const NodalArgs &n, bool track_d, bool has_bcs, hipStream_t s);
void launch_courant_f64(int dim, const ElemArgs &a, double cs, double *part, hipStream_t s);
void launch_min_reduce(const double *part, int npart, double *out, hipStream_t s);
void launch_dt(StepState *st, const double *part, int npart, int recompute, const double *ke_part, int nke, int mode,
               hipStream_t s);
unsigned nodal_blocks(long long nn);
void adj_free(AdjStore *s);
unsigned courant_blocks(long long ne);
}  // namespace dnl

using namespace dnl;

static thread_local std::string g_err;

extern "C" int dnlg_abi_version(void) { return DNLG_ABI_VERSION; }

extern "C" const char *dnlg_last_error(dnlg_handle *h)
{
  if (!h) return g_err.c_str();
  return h->err.c_str();
}

static void fill_elem_args(dnlg_handle *h)
{
  ElemArgs &a = h->ea;
  a.conn = h->conn;
  a.x = h->x;
  a.du = h->du;
  a.S = h->S;
  a.evp = h->evp;
  a.p = h->p;
  a.eps = h->track_strain ? h->eps : nullptr;
  a.epsplas = h->track_strain ? h->epsplas : nullptr;
  a.deps = h->track_strain ? h->deps : nullptr;
  a.fe = h->fe;
  a.fpos = h->fpos;
  a.bcs = h->bcs;
  a.bcs_v = h->bcs_v;
  a.st = h->st;
  a.ne = h->ne;
  const double E = h->mat.young, nu = h->mat.poisson;
  a.K = E / (3. * (1. - 2. * nu));   // FE/Element.C:232
  a.G = E / (2. * (1. + nu));        // Material::mu, FE/Material.h:167-172
  a.A = h->mat.A;
  a.B = h->mat.B;
  a.n = h->mat.n;
  a.nm1 = h->mat.n - 1.;
  a.yield = h->mat.hardening ? 1 : 0;
  const dnlg_material &M = h->mat;
  const double chi = (M.hardening == 2 && M.cp > 0.0) ? M.taylor_quinney / (M.density * M.cp) : 0.0;
  a.law = M.hardening == 2 ? ((M.C == 0.0 && chi == 0.0) ? 1 : 2) : (M.hardening ? 1 : 0);
  a.C = M.C;
  a.epsd0 = M.epsd0;
  a.m = M.m;
  a.T0 = M.T0;
  a.Tm = M.Tm;
  a.chi_rho_cp = chi;
  a.T = h->T;
  ElemArgs32 &b = h->ea32;
  b.conn = h->conn;
  b.x = h->x;
  b.du = h->du;
  b.S = h->S32;
  b.evp = h->evp32;
  b.T = (a.law == 2) ? h->T32 : nullptr;
  b.p = h->p32;
  b.fe = h->fe32;
  b.fpos = h->fpos;
  b.st = h->st;
  b.ne = h->ne;
  b.K = (float)a.K;
  b.G = (float)a.G;
  b.A = (float)M.A;
  b.B = (float)M.B;
  b.n = (float)M.n;
  b.C = (float)M.C;
  b.epsd0 = (float)M.epsd0;
  b.m = (float)M.m;
  b.T0 = (float)M.T0;
  b.Tm = (float)M.Tm;
  b.chi_rho_cp = (float)chi;
  b.law = a.law;
  if (h->dim == 3)
  {
    // closed forms of the hex8 tables (FE/ElHex8n3D.C:81-87)
    a.c0 = (2.0 - sqrt(3.0)) / 12.0;
    a.c1 = 1.0 / 12.0;
    a.c2 = (2.0 + sqrt(3.0)) / 12.0;
    h->hn[0] = (9.0 - 5.0 * sqrt(3.0)) / 36.0;
    h->hn[1] = (3.0 - sqrt(3.0)) / 36.0;
    h->hn[2] = (3.0 + sqrt(3.0)) / 36.0;
    h->hn[3] = (9.0 + 5.0 * sqrt(3.0)) / 36.0;
  }
  else
  {
    // FE/ElQua4nAx.C:53-57
    a.c0 = (3.0 - sqrt(3.0)) / 12.0;
    a.c1 = (3.0 + sqrt(3.0)) / 12.0;
    a.c2 = 0.0;
  }
  a.n0 = (2.0 - sqrt(3.0)) / 6.0;
  a.n1 = 1.0 / 6.0;
  a.n2 = (2.0 + sqrt(3.0)) / 6.0;
  a.n3 = 0.0;
}

static NodalArgs nodal_args(dnlg_handle *h)
{
  NodalArgs n{};
  n.x = h->x;
  n.v = h->v;
  n.a = h->a;
  n.du = h->du;
  n.d = h->d;
  n.mass = h->mass;
  n.bc = h->bc;
  n.bcs = h->bcs;
  n.bcs_v = h->bcs_v;
  n.deg = h->deg;
  n.fe = h->fe;
  n.fe32 = h->form == DNLG_FORM_MIXED ? h->fe32 : nullptr;
  n.fint = h->store_fint ? h->fint : nullptr;
  n.ke_part = h->ke_on ? h->ke_part : nullptr;
  n.st = h->st;
  n.nn = h->nn;
  return n;
}

static void free_all(dnlg_handle *h)
{
  for (const HistFile &f : h->hf)
    if (f.ring) (void)hipFree(f.ring);
  if (h->hf_dev) (void)hipFree(h->hf_dev);
  if (h->hi_dev) (void)hipFree(h->hi_dev);
  if (h->hel_dev) (void)hipFree(h->hel_dev);
  adj_free(h->adj);
  h->adj = nullptr;
  void *ptrs[] = {h->x, h->v, h->a, h->du, h->d, h->mass, h->fint, h->fe, h->S, h->evp, h->p, h->eps, h->epsplas, h->deps,
                  h->me, h->part, h->ke_part, h->bcs_v, h->conn, h->fpos, h->deg, h->bc, h->bcs, h->st, h->T,
                  h->S32, h->evp32, h->T32, h->p32, h->fe32, h->pulse_ke, h->pulse_val, h->pulse_init,
                  h->pulse_part, h->pulse_ctl, h->pulse_y == h->fe ? nullptr : h->pulse_y, h->plot_ring,
                  h->crit};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

extern "C" int dnlg_create(const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bc *bc,
                           const dnlg_solver_params *solver, dnlg_handle **out)
{
  if (!out) { g_err = "dnlg_create: null output handle"; return DNLG_E_ARG; }
  *out = nullptr;
  if (!mesh || !mat || !solver || !mesh->coords || !mesh->conn || (mesh->dim != 2 && mesh->dim != 3) ||
      mesh->n_nodes <= 0 || mesh->n_elems <= 0)
  {
    g_err = "dnlg_create: bad mesh / material / solver arguments";
    return DNLG_E_ARG;
  }
  if (!(mat->young > 0 && mat->density > 0 && mat->poisson > -1.0 && mat->poisson < 0.5))
  {
    g_err = "dnlg_create: material constants out of range";
    return DNLG_E_ARG;
  }
  if (!(solver->rho_b >= 0.0 && solver->rho_b <= 1.0))
  {
    g_err = "ExplicitSolverCH::setDissipation: dissipation value out of range [0.0, 1.0]";  // FE/ExplicitSolverCH.C:362-363
    return DNLG_E_ARG;
  }
  if (solver->form != DNLG_FORM_FP64 && solver->form != DNLG_FORM_MIXED)
  {
    g_err = "dnlg_create: unknown arithmetic form";
    return DNLG_E_ARG;
  }
  if (mat->hardening < 0 || mat->hardening > 2 ||
      (mat->hardening == 2 && (mat->C != 0.0 || mat->taylor_quinney != 0.0) &&
       !(mat->epsd0 > 0.0 && mat->Tm > mat->T0 && mat->cp >= 0.0)))
  {
    g_err = "dnlg_create: hardening law parameters out of range";
    return DNLG_E_ARG;
  }
  const int nen = mesh->dim == 3 ? 8 : 4;
  const long long nn = mesh->n_nodes, ne = mesh->n_elems;
  if (ne * nen > 2147483647LL || nn > 2147483647LL)
  {
    g_err = "dnlg_create: mesh too large for 32-bit element-node indices";
    return DNLG_E_ARG;
  }
  for (long long k = 0; k < ne * nen; k++)
    if (mesh->conn[k] < 0 || mesh->conn[k] >= nn)
    {
      g_err = "dnlg_create: connectivity references a node outside [0, n_nodes)";
      return DNLG_E_ARG;
    }
  dnlg_handle *h = new dnlg_handle();
  h->dim = mesh->dim;
  h->nen = nen;
  h->nip = nen;
  h->nn = nn;
  h->ne = ne;
  h->mat = *mat;
  h->form = solver->form;
  h->freq = solver->ts_freq > 0 ? solver->ts_freq : 10;
  h->has_bcs = bc && bc->speed_mask && bc->speed;

  // element force slots: node i's s-th element (in element order, the order in
  // which Domain::computeInternalForces case 0 adds into Fint) writes its
  // contribution to slot s; slots are [maxdeg][nn] so a node's gather reads
  // contiguous memory across a wave
  std::vector<unsigned char> deg(nn, 0);
  h->fpos_host.assign(ne * nen, 0);
  {
    std::vector<int> cnt(nn, 0);
    for (long long k = 0; k < ne * nen; k++) cnt[mesh->conn[k]]++;
    int md = 0;
    for (long long i = 0; i < nn; i++) md = cnt[i] > md ? cnt[i] : md;
    if (md > 255 || (long long)md * nn > 2147483647LL)
    {
      g_err = "dnlg_create: node degree too large";
      delete h;
      return DNLG_E_ARG;
    }
    h->maxdeg = md;
    for (long long e = 0; e < ne; e++)
      for (int k = 0; k < nen; k++)
      {
        const int node = mesh->conn[e * nen + k];
        h->fpos_host[e * nen + k] = (int)((long long)deg[node] * nn + node);
        deg[node]++;
      }
  }
  const long long nipt = ne * h->nip;
  std::vector<double> zero3(nn * 3, 0.0);
  std::vector<unsigned char> zb(nn, 0);
  std::vector<double> evp0(nipt, 1e-10);  // IntegrationPoint::flush, FE/IntegrationPoint.C:117

#define CK(call)                                                                  \
  do                                                                              \
  {                                                                               \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
    {                                                                             \
      g_err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #call;   \
      free_all(h);                                                                \
      delete h;                                                                   \
      return DNLG_E_HIP;                                                          \
    }                                                                             \
  } while (0)

  CK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  CK(dalloc(&h->x, nn * 3));
  CK(dalloc(&h->v, nn * 3));
  CK(dalloc(&h->a, nn * 3));
  CK(dalloc(&h->du, nn * 3));
  CK(dalloc(&h->d, nn * 3));
  CK(dalloc(&h->mass, nn));
  CK(dalloc(&h->fint, nn * h->dim));
  const bool jc = mat->hardening == 2;
  if (h->form == DNLG_FORM_FP64)
  {
    CK(dalloc(&h->fe, (long long)h->maxdeg * nn * h->dim));
    CK(dalloc(&h->S, nipt * 9));
    CK(dalloc(&h->evp, nipt));
    CK(dalloc(&h->p, ne));
    if (jc)
    {
      std::vector<double> t0(nipt, mat->T0);
      CK(dalloc(&h->T, nipt));
      CK(hipMemcpy(h->T, t0.data(), sizeof(double) * nipt, hipMemcpyHostToDevice));
    }
    CK(hipMemset(h->S, 0, sizeof(double) * nipt * 9));
    CK(hipMemcpy(h->evp, evp0.data(), sizeof(double) * nipt, hipMemcpyHostToDevice));
    CK(hipMemset(h->p, 0, sizeof(double) * ne));
  }
  else
  {
    std::vector<float> e32(nipt, 1e-10f);
    CK(dalloc(&h->fe32, (long long)h->maxdeg * nn * h->dim));
    CK(dalloc(&h->S32, nipt * 6));
    CK(dalloc(&h->evp32, nipt));
    CK(dalloc(&h->p32, ne));
    CK(hipMemset(h->S32, 0, sizeof(float) * nipt * 6));
    CK(hipMemcpy(h->evp32, e32.data(), sizeof(float) * nipt, hipMemcpyHostToDevice));
    CK(hipMemset(h->p32, 0, sizeof(float) * ne));
    if (jc)
    {
      std::vector<float> t0(nipt, (float)mat->T0);
      CK(dalloc(&h->T32, nipt));
      CK(hipMemcpy(h->T32, t0.data(), sizeof(float) * nipt, hipMemcpyHostToDevice));
    }
  }
  CK(dalloc(&h->me, (long long)h->maxdeg * nn));
  CK(dalloc(&h->part, courant_blocks(ne)));
  CK(dalloc(&h->crit, 1));
  CK(dalloc(&h->ke_part, nodal_blocks(nn)));
  CK(dalloc(&h->conn, ne * nen));
  CK(dalloc(&h->fpos, ne * nen));
  CK(dalloc(&h->deg, nn));
  CK(dalloc(&h->bc, nn));
  CK(dalloc(&h->st, 1));
  if (h->has_bcs)
  {
    CK(dalloc(&h->bcs, nn));
    CK(dalloc(&h->bcs_v, nn * 3));
    CK(hipMemcpy(h->bcs, bc->speed_mask, nn, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->bcs_v, bc->speed, sizeof(double) * nn * 3, hipMemcpyHostToDevice));
  }
  CK(hipMemcpy(h->x, mesh->coords, sizeof(double) * nn * 3, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->v, (bc && bc->v0) ? bc->v0 : zero3.data(), sizeof(double) * nn * 3, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->a, zero3.data(), sizeof(double) * nn * 3, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->du, zero3.data(), sizeof(double) * nn * 3, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->d, zero3.data(), sizeof(double) * nn * 3, hipMemcpyHostToDevice));
  CK(hipMemset(h->fint, 0, sizeof(double) * nn * h->dim));
  CK(hipMemcpy(h->conn, mesh->conn, sizeof(int) * ne * nen, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->fpos, h->fpos_host.data(), sizeof(int) * ne * nen, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->deg, deg.data(), nn, hipMemcpyHostToDevice));
  CK(hipMemcpy(h->bc, (bc && bc->restrain) ? bc->restrain : zb.data(), nn, hipMemcpyHostToDevice));

  // scalar state: Solver() defaults (FE/Solver.C:45-76) and
  // ExplicitSolverCH::computeIntegrationParameters (FE/ExplicitSolverCH.C:336-343)
  StepState s{};
  const double rb = solver->rho_b;
  s.alphaM = (2.0 * rb - 1.0) / (1.0 + rb);
  s.beta = (5.0 - 3.0 * rb) / ((1.0 + rb) * (1.0 + rb) * (2.0 - rb));
  s.gamma = 1.5 - s.alphaM;
  s.omegaS = sqrt((12.0 * pow((1 + rb), 3) * (2.0 - rb)) / (10 + 15 * rb - rb * rb + pow(rb, 3) - pow(rb, 4)));
  s.safety = solver->safety;
  s.endTime = solver->end_time;
  s.freq = h->freq;
  // Element::getSoundSpeed with every nodal density equal to the material's (FE/Element.C:607-621)
  {
    const double E = mat->young, Nu = mat->poisson;
    double ro = 0.;
    for (int i = 0; i < nen; i++) ro += mat->density;
    ro /= nen;
    s.cs = sqrt((E * (1 - Nu)) / (ro * (1 + Nu) * (1 - 2 * Nu)));
  }
  h->cs = s.cs;
  CK(hipMemcpy(h->st, &s, sizeof s, hipMemcpyHostToDevice));
#undef CK
  fill_elem_args(h);
  *out = h;
  return DNLG_OK;
}

extern "C" int dnlg_destroy(dnlg_handle *h)
{
  if (!h) return DNLG_E_ARG;
  (void)hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
  return DNLG_OK;
}

static int finalize(dnlg_handle *h);
static int read_state(dnlg_handle *h, StepState *s);
static constexpr long long PLOT_CAP = 8192;  // history lines the device ring holds between synchronisations

extern "C" int dnlg_set_option(dnlg_handle *h, int option, int value)
{
  if (!h) return DNLG_E_ARG;
  switch (option)
  {
  case DNLG_OPT_KINETIC_ENERGY: h->ke_on = value != 0; break;
  case DNLG_OPT_STORE_FINT: h->store_fint = value != 0; break;
  case DNLG_OPT_TRACK_DISP: h->track_d = value != 0; break;
  case DNLG_OPT_TRACK_STRAIN:
    if (value && h->form != DNLG_FORM_FP64)
    {
      h->err = "dnlg_set_option: strain tracking exists in the FP64 form only";
      return DNLG_E_ARG;
    }
    if (value && !h->eps)
    {
      const long long nipt = h->ne * h->nip;
      DNLG_HIP(h, dalloc(&h->eps, nipt * 9));
      DNLG_HIP(h, dalloc(&h->epsplas, nipt * 9));
      DNLG_HIP(h, dalloc(&h->deps, nipt * 9));
      DNLG_HIP(h, hipMemset(h->eps, 0, sizeof(double) * nipt * 9));
      DNLG_HIP(h, hipMemset(h->epsplas, 0, sizeof(double) * nipt * 9));
      DNLG_HIP(h, hipMemset(h->deps, 0, sizeof(double) * nipt * 9));
    }
    h->track_strain = value != 0;
    fill_elem_args(h);
    break;
  case DNLG_OPT_TIME_STEP_METHOD:
  {
    if (value != DNLG_TS_COURANT && value != DNLG_TS_PULSE)
    {
      h->err = "setTimeStepMethod: Unknown method";
      return DNLG_E_ARG;
    }
    int rc = finalize(h);
    if (rc) return rc;
    StepState s;
    rc = read_state(h, &s);
    if (rc) return rc;
    s.ts_method = value;
    DNLG_HIP(h, hipMemcpy(h->st, &s, sizeof s, hipMemcpyHostToDevice));
    h->ts_method = value;
    break;
  }
  case DNLG_OPT_PLOT_HISTORY:
  {
    int rc = finalize(h);
    if (rc) return rc;
    StepState s;
    rc = read_state(h, &s);
    if (rc) return rc;
    if (value && !h->plot_ring)
    {
      DNLG_HIP(h, dalloc(&h->plot_ring, PLOT_CAP * 4));
      h->plot_drained = s.plot_n;
    }
    s.plot = value ? h->plot_ring : nullptr;
    s.plot_cap = (int)PLOT_CAP;
    DNLG_HIP(h, hipMemcpy(h->st, &s, sizeof s, hipMemcpyHostToDevice));
    if (value) h->ke_on = true;  // the lines carry Domain::getTotalKineticEnergy
    break;
  }
  default: h->err = "dnlg_set_option: unknown option"; return DNLG_E_ARG;
  }
  return DNLG_OK;
}

// the history lines written since the last drain, ring -> host buffer
static int drain_plot(dnlg_handle *h, long long plot_n)
{
  const long long n = plot_n - h->plot_drained;
  if (n <= 0) return DNLG_OK;
  if (n > PLOT_CAP)
  {
    h->err = "history lines lost: more increments than the ring holds between synchronisations";
    return DNLG_E_STATE;
  }
  const size_t old = h->plot_host.size();
  h->plot_host.resize(old + 4 * (size_t)n);
  const long long first = h->plot_drained % PLOT_CAP;
  const long long n1 = n < PLOT_CAP - first ? n : PLOT_CAP - first;
  DNLG_HIP(h, hipMemcpy(&h->plot_host[old], h->plot_ring + 4 * first, sizeof(double) * 4 * n1,
                        hipMemcpyDeviceToHost));
  if (n > n1)
    DNLG_HIP(h, hipMemcpy(&h->plot_host[old + 4 * n1], h->plot_ring, sizeof(double) * 4 * (n - n1),
                          hipMemcpyDeviceToHost));
  h->plot_drained = plot_n;
  return DNLG_OK;
}

// HistoryFile rows written since the last drain, rings -> host buffers
static int drain_history(dnlg_handle *h)
{
  const int nf = (int)h->hf.size();
  std::vector<HistFile> dev(nf);
  DNLG_HIP(h, hipMemcpy(dev.data(), h->hf_dev, sizeof(HistFile) * nf, hipMemcpyDeviceToHost));
  for (int f = 0; f < nf; f++)
  {
    const long long n = dev[f].rows - h->hf_drained[f];
    if (n <= 0) continue;
    if (n > dev[f].cap)
    {
      h->err = "history rows lost: more increments than the ring holds between synchronisations";
      return DNLG_E_STATE;
    }
    const long long w = 1 + dev[f].count, cap = dev[f].cap;
    std::vector<double> &out = h->hf_rows[f];
    const size_t old = out.size();
    out.resize(old + (size_t)(n * w));
    const long long first = h->hf_drained[f] % cap;
    const long long n1 = n < cap - first ? n : cap - first;
    DNLG_HIP(h, hipMemcpy(&out[old], dev[f].ring + first * w, sizeof(double) * n1 * w, hipMemcpyDeviceToHost));
    if (n > n1)
      DNLG_HIP(h, hipMemcpy(&out[old + n1 * w], dev[f].ring, sizeof(double) * (n - n1) * w, hipMemcpyDeviceToHost));
    h->hf_drained[f] = dev[f].rows;
  }
  return DNLG_OK;
}

static int read_state(dnlg_handle *h, StepState *s)
{
  DNLG_HIP(h, hipMemcpyAsync(s, h->st, sizeof *s, hipMemcpyDeviceToHost, h->stream));
  DNLG_HIP(h, hipStreamSynchronize(h->stream));
  if (h->plot_ring)
  {
    const int rc = drain_plot(h, s->plot_n);
    if (rc) return rc;
  }
  if (!h->hf.empty())
  {
    const int rc = drain_history(h);
    if (rc) return rc;
  }
  h->drain_inc = s->increment;
  return DNLG_OK;
}

static HistArgs hist_args(dnlg_handle *h)
{
  HistArgs a{};
  a.files = h->hf_dev;
  a.items = h->hi_dev;
  a.nfiles = (int)h->hf.size();
  a.nitems = (int)h->hitems.size();
  a.x = h->x;
  a.v = h->v;
  a.a = h->a;
  a.du = h->du;
  a.d = h->d;
  a.mass = h->mass;
  a.st = h->st;
  a.density = h->mat.density;
  a.T0 = h->init_temp;
  a.hel = h->hel_dev;
  a.S = h->S;
  a.p = h->p;
  a.eps = h->eps;
  a.epsplas = h->epsplas;
  a.deps = h->deps;
  a.nipt = h->ne * h->nip;
  a.nip = h->nip;
  a.nen = h->nen;
  // integrationPointsToNode of ElHex8n3D (FE/ElHex8n3D.C:33-72) and ElQua4nAx
  // (FE/ElQua4nAx.C:38-48), each coefficient the expression the reference's
  // tables evaluate
  if (h->dim == 3)
  {
    const double p5 = (5.0 + 3.0 * sqrt(3.0)) / 4.0, m5 = (5.0 - 3.0 * sqrt(3.0)) / 4.0;
    const double r1 = (sqrt(3.0) + 1.0) / 4.0, q1 = (sqrt(3.0) - 1.0) / 4.0;
    const double t[8][8] = {{p5, -r1, q1, -r1, -r1, q1, m5, q1}, {-r1, p5, -r1, q1, q1, -r1, q1, m5},
                            {q1, -r1, p5, -r1, m5, q1, -r1, q1}, {-r1, q1, -r1, p5, q1, m5, q1, -r1},
                            {-r1, q1, m5, q1, p5, -r1, q1, -r1}, {q1, -r1, q1, m5, -r1, p5, -r1, q1},
                            {m5, q1, -r1, q1, q1, -r1, p5, -r1}, {q1, m5, q1, -r1, -r1, q1, -r1, p5}};
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) a.coef[i][j] = t[i][j];
  }
  else
  {
    const double p1 = 1.0 + sqrt(3.0) / 2.0, m1 = 1.0 - sqrt(3.0) / 2.0, hf = 1.0 / 2.0;
    const double t[4][4] = {{p1, -hf, m1, -hf}, {-hf, p1, -hf, m1}, {m1, -hf, p1, -hf}, {-hf, m1, -hf, p1}};
    for (int i = 0; i < 8; i++)
      for (int j = 0; j < 8; j++) a.coef[i][j] = (i < 4 && j < 4) ? t[i][j] : 0.0;
  }
  return a;
}

static int check_status(dnlg_handle *h, const StepState &s)
{
  if (s.status == ST_OK) return DNLG_OK;
  char buf[256];
  const char *what = s.status == ST_POLAR ? "Eigen Value Computation: No convergence (polar decomposition)"
                     : s.status == ST_RR  ? "Element::computeInternalForces: Too many atempts in plastic procedure"
                     : s.status == ST_DETJ ? "Negative or singular Jacobian"
                     : s.status == ST_PULSE ? "Domain::getMaximumFrequency: power iteration method not converged"
                                           : "device error";
  if (s.status == ST_PULSE)
    snprintf(buf, sizeof buf, "%s", what);
  else
    snprintf(buf, sizeof buf, "%s (element %d)", what, s.err_elem);
  h->err = buf;
  return s.status == ST_POLAR ? DNLG_E_POLAR_NOCONV : s.status == ST_RR ? DNLG_E_RR_NOCONV
         : s.status == ST_PULSE ? DNLG_E_PULSE_NOCONV : DNLG_E_NEG_DETJ;
}

// Domain::getMaximumFrequency (FE/Domain.C:816-946): Ke of every element
// once, then power iterations issued in growing batches, the host reading
// the convergence flag between batches (the kernels of iterations issued
// past convergence return at once).  Buffers are allocated at the first call.
static int run_pulse(dnlg_handle *h)
{
  const int nd = h->nen * h->dim;
  const long long ndof = h->nn * h->dim;
  if (!h->pulse_ctl)
  {
    DNLG_HIP(h, dalloc(&h->pulse_ke, ((h->ne + 7) / 8) * 8 * (long long)nd * nd));
    DNLG_HIP(h, dalloc(&h->pulse_val, ndof));
    DNLG_HIP(h, dalloc(&h->pulse_init, ndof));
    DNLG_HIP(h, dalloc(&h->pulse_part, nodal_blocks(h->nn)));
    DNLG_HIP(h, dalloc(&h->pulse_ctl, 1));
    if (h->fe)
      h->pulse_y = h->fe;  // free between an increment's nodal pass and the next element pass
    else
      DNLG_HIP(h, dalloc(&h->pulse_y, (long long)h->maxdeg * h->nn * h->dim));
    // start vector ev(0) = 1, ev(i) = ev(i-1) - 2/(noddl-1), FE/Domain.C:849-856
    std::vector<double> ev(ndof);
    ev[0] = 1.0;
    for (long long i = 1; i < ndof; i++) ev[i] = ev[i - 1] - 2. / (ndof - 1);
    DNLG_HIP(h, hipMemcpy(h->pulse_init, ev.data(), sizeof(double) * ndof, hipMemcpyHostToDevice));
    PulseCtl c{};
    c.invf = 1.0;
    DNLG_HIP(h, hipMemcpy(h->pulse_ctl, &c, sizeof c, hipMemcpyHostToDevice));
  }
  PulseArgs p{};
  p.conn = h->conn;
  p.fpos = h->fpos;
  p.deg = h->deg;
  p.mass = h->mass;
  p.ev_init = h->pulse_init;
  p.ke = h->pulse_ke;
  p.y = h->pulse_y;
  p.val = h->pulse_val;
  p.part = h->pulse_part;
  p.ctl = h->pulse_ctl;
  p.st = h->st;
  p.ne = h->ne;
  p.nn = h->nn;
  p.ndof = ndof;
  p.E = h->mat.young;
  p.nu = h->mat.poisson;
  launch_pulse_begin(p, h->stream);
  launch_pulse_ke(h->dim, p, h->ea, h->stream);  // and the first iteration
  int batch = 0;  // a call that starts from the last iterate often converges at once
  for (;;)
  {
    for (int b = 0; b < batch; b++) launch_pulse_iteration(h->dim, p, h->stream);
    DNLG_HIP(h, hipGetLastError());
    PulseCtl c;
    StepState s;
    DNLG_HIP(h, hipMemcpyAsync(&c, h->pulse_ctl, sizeof c, hipMemcpyDeviceToHost, h->stream));
    DNLG_HIP(h, hipMemcpyAsync(&s, h->st, sizeof s, hipMemcpyDeviceToHost, h->stream));
    DNLG_HIP(h, hipStreamSynchronize(h->stream));
    if (!s.running) return DNLG_OK;  // the solve loop ended: no computeTimeStep
    if (c.done)
    {
      h->pulse_calls++;
      h->pulse_iters += c.iter;
      h->pulse_iters_last = c.iter;
      return c.failed ? check_status(h, s) : DNLG_OK;
    }
    batch = batch == 0 ? 1 : (batch < 64 ? 2 * batch : 64);
  }
}

static void launch_elem(dnlg_handle *h, bool check_running, bool speed)
{
  if (h->form == DNLG_FORM_MIXED)
  {
    ElemArgs32 b = h->ea32;
    b.check_running = check_running ? 1 : 0;
    launch_elem_f32(h->dim, b, h->stream);
  }
  else
  {
    ElemArgs a = h->ea;
    a.check_running = check_running ? 1 : 0;
    launch_elem_f64(h->dim, a, speed, h->stream);
  }
}

static hipEvent_t next_event(dnlg_handle *h)
{
  if (h->ev_used >= h->ev_pool.size())
  {
    hipEvent_t e;
    // timing events only: no system-scope fence at record, which would write
    // back the caches between the kernels they bracket and slow the next one
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    h->ev_pool.push_back(e);
  }
  return h->ev_pool[h->ev_used++];
}

// Domain::computeInternalForces(): element pass with the Current increment, then Fint
extern "C" int dnlg_compute_internal_forces(dnlg_handle *h, double *fint_out)
{
  if (!h) return DNLG_E_ARG;
  launch_elem(h, false, false);
  DNLG_HIP(h, hipGetLastError());
  launch_gather_fint(h->dim, h->deg, h->fe, h->form == DNLG_FORM_MIXED ? h->fe32 : nullptr, h->fint, h->nn,
                     h->stream);
  DNLG_HIP(h, hipGetLastError());
  if (fint_out)
    DNLG_HIP(h, hipMemcpyAsync(fint_out, h->fint, sizeof(double) * h->nn * h->dim, hipMemcpyDeviceToHost, h->stream));
  StepState s;
  int rc = read_state(h, &s);
  if (rc) return rc;
  return check_status(h, s);
}

// Domain::computeInternalForces() inside the caller's own time loop: the
// caller's coordinates and Current displacement increment replace the
// device copies, then the element pass and the Fint gather run as above
extern "C" int dnlg_compute_internal_forces_from(dnlg_handle *h, const double *coords, const double *delta_disp,
                                                 double *fint_out)
{
  if (!h) return DNLG_E_ARG;
  int rc = finalize(h);
  if (rc) return rc;
  const size_t nb = sizeof(double) * (size_t)h->nn * 3;
  if (coords) DNLG_HIP(h, hipMemcpyAsync(h->x, coords, nb, hipMemcpyHostToDevice, h->stream));
  if (delta_disp) DNLG_HIP(h, hipMemcpyAsync(h->du, delta_disp, nb, hipMemcpyHostToDevice, h->stream));
  return dnlg_compute_internal_forces(h, fint_out);
}

// Domain::computeCourantTimeStep (FE/Domain.C:1613-1646) on the handle's
// present coordinates: per-block minima of l / c into the handle's buffer,
// then their minimum
extern "C" int dnlg_courant_time_step(dnlg_handle *h, double *crit_out)
{
  if (!h || !crit_out) return DNLG_E_ARG;
  int rc = finalize(h);
  if (rc) return rc;
  launch_courant_f64(h->dim, h->ea, h->cs, h->part, h->stream);
  launch_min_reduce(h->part, (int)courant_blocks(h->ne), h->crit, h->stream);
  DNLG_HIP(h, hipGetLastError());
  DNLG_HIP(h, hipMemcpyAsync(crit_out, h->crit, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  DNLG_HIP(h, hipStreamSynchronize(h->stream));
  return DNLG_OK;
}

// close the last issued increment: endStep (+ computeTimeStep while running)
static int finalize(dnlg_handle *h)
{
  if (!h->pending) return DNLG_OK;
  const int rec = (h->inc % h->freq) == 0;
  h->pending = false;
  if (rec && h->ts_method == DNLG_TS_PULSE)
  {
    launch_dt(h->st, nullptr, 0, 0, h->ke_on ? h->ke_part : nullptr, (int)nodal_blocks(h->nn), 3, h->stream);
    DNLG_HIP(h, hipGetLastError());
    const int rc = run_pulse(h);
    if (rc) return rc;
    launch_dt(h->st, nullptr, 0, 1, nullptr, 0, 5, h->stream);
  }
  else
  {
    if (rec) launch_courant_f64(h->dim, h->ea, h->cs, h->part, h->stream);
    launch_dt(h->st, h->part, (int)courant_blocks(h->ne), rec, h->ke_on ? h->ke_part : nullptr,
              (int)nodal_blocks(h->nn), 1, h->stream);
  }
  DNLG_HIP(h, hipGetLastError());
  return DNLG_OK;
}

extern "C" int dnlg_begin(dnlg_handle *h, double up_time, int *running)
{
  if (!h) return DNLG_E_ARG;
  int rc = finalize(h);
  if (rc) return rc;
  StepState s;
  rc = read_state(h, &s);
  if (rc) return rc;
  // ExplicitSolverCH::solve: upTime, betweenTimeBounds (FE/ExplicitSolverCH.C:121-128)
  s.upTime = up_time;
  const double lim = s.endTime < up_time ? s.endTime : up_time;
  s.running = (s.t >= 0.0) && (s.t < lim);
  DNLG_HIP(h, hipMemcpyAsync(h->st, &s, sizeof s, hipMemcpyHostToDevice, h->stream));
  h->inc = s.increment;
  h->begun = s.running != 0;
  if (running) *running = s.running;
  if (!s.running) return DNLG_OK;
  // computeMassMatrix, once (FE/Domain.C:951)
  if (!h->mass_done)
  {
    launch_mass_f64(h->dim, h->ea, h->mat.density, h->hn, h->me, h->stream);
    launch_gather_scalar(h->deg, h->me, h->mass, h->nn, h->stream);
    DNLG_HIP(h, hipGetLastError());
    h->mass_done = true;
  }
  // computeInternalForces of the preamble (FE/ExplicitSolverCH.C:157)
  launch_elem(h, false, false);
  launch_gather_fint(h->dim, h->deg, h->fe, h->form == DNLG_FORM_MIXED ? h->fe32 : nullptr, h->fint, h->nn,
                     h->stream);
  // computeTimeStep(True) (FE/ExplicitSolverCH.C:172)
  if (h->ts_method == DNLG_TS_PULSE)
  {
    DNLG_HIP(h, hipGetLastError());
    rc = run_pulse(h);
    if (rc) return rc;
    launch_dt(h->st, nullptr, 0, 1, nullptr, 0, 2, h->stream);
  }
  else
  {
    launch_courant_f64(h->dim, h->ea, h->cs, h->part, h->stream);
    launch_dt(h->st, h->part, (int)courant_blocks(h->ne), 1, nullptr, 0, 2, h->stream);
  }
  DNLG_HIP(h, hipGetLastError());
  rc = read_state(h, &s);
  if (rc) return rc;
  return check_status(h, s);
}

static int issue_steps(dnlg_handle *h, int nsteps)
{
  if (!h->begun)
  {
    h->err = "dnlg_step: call dnlg_begin first";
    return DNLG_E_STATE;
  }
  const bool rings = h->plot_ring || !h->hf.empty();
  if (rings && h->inc + nsteps - h->drain_inc > PLOT_CAP)
  {
    h->err = "dnlg_step_async: the history rings hold 8192 increments between synchronisations";
    return DNLG_E_STATE;
  }
  const NodalArgs n = nodal_args(h);
  const HistArgs ha = hist_args(h);
  const int npart = (int)courant_blocks(h->ne), nke = (int)nodal_blocks(h->nn);
  for (int s = 0; s < nsteps; s++)
  {
    int rec = 0;
    if (h->pending) rec = (h->inc % h->freq) == 0;
    if (rec && h->ts_method == DNLG_TS_PULSE)
    {
      // endStep; getMaximumFrequency (host waits); computeTimeStep + initStep
      launch_dt(h->st, nullptr, 0, 0, h->ke_on ? h->ke_part : nullptr, nke, 3, h->stream);
      DNLG_HIP(h, hipGetLastError());
      const int rc = run_pulse(h);
      if (rc) return rc;
      launch_dt(h->st, nullptr, 0, 1, nullptr, 0, 4, h->stream);
    }
    else
    {
      if (rec) launch_courant_f64(h->dim, h->ea, h->cs, h->part, h->stream);
      launch_dt(h->st, h->part, npart, rec, h->ke_on ? h->ke_part : nullptr, nke, 0, h->stream);
    }
    if (ha.nfiles) launch_hist_pre(ha, h->stream);
    if (h->timing)
    {
      hipEvent_t e0 = next_event(h), e1 = next_event(h), e2 = next_event(h), e3 = next_event(h);
      if (!e0 || !e1 || !e2 || !e3)
      {
        h->err = "dnlg_step: hipEventCreate failed";
        return DNLG_E_HIP;
      }
      DNLG_HIP(h, hipEventRecord(e0, h->stream));
      launch_elem(h, true, h->has_bcs);
      DNLG_HIP(h, hipEventRecord(e1, h->stream));
      DNLG_HIP(h, hipEventRecord(e2, h->stream));
      launch_nodal_f64(h->dim, n, h->track_d, h->has_bcs, h->stream);
      DNLG_HIP(h, hipEventRecord(e3, h->stream));
      h->ev_elem.push_back(e0);
      h->ev_elem.push_back(e1);
      h->ev_nodal.push_back(e2);
      h->ev_nodal.push_back(e3);
    }
    else
    {
      launch_elem(h, true, h->has_bcs);
      launch_nodal_f64(h->dim, n, h->track_d, h->has_bcs, h->stream);
    }
    if (ha.nfiles) launch_hist_post(ha, h->stream);
    h->inc++;
    h->pending = true;
  }
  DNLG_HIP(h, hipGetLastError());
  return DNLG_OK;
}

extern "C" int dnlg_step_async(dnlg_handle *h, int nsteps)
{
  if (!h || nsteps < 0) return DNLG_E_ARG;
  return issue_steps(h, nsteps);
}

extern "C" int dnlg_synchronize(dnlg_handle *h)
{
  if (!h) return DNLG_E_ARG;
  int rc = finalize(h);
  if (rc) return rc;
  StepState s;
  rc = read_state(h, &s);
  if (rc) return rc;
  h->inc = s.increment;
  return check_status(h, s);
}

extern "C" int dnlg_step(dnlg_handle *h, int nsteps, int *running)
{
  if (!h || nsteps < 0) return DNLG_E_ARG;
  // with history rings, synchronise (and drain them) at least every half ring
  const int lim = (h->plot_ring || !h->hf.empty()) ? (int)(PLOT_CAP / 2) : nsteps;
  StepState s;
  int rc;
  do
  {
    const int k = nsteps < lim ? nsteps : lim;
    rc = issue_steps(h, k);
    if (rc) return rc;
    rc = finalize(h);
    if (rc) return rc;
    rc = read_state(h, &s);
    if (rc) return rc;
    h->inc = s.increment;
    nsteps -= k;
  } while (nsteps > 0 && s.running && s.status == ST_OK);
  if (running) *running = s.running;
  if (!s.running) h->begun = false;
  return check_status(h, s);
}

extern "C" int dnlg_history_file_add(dnlg_handle *h, double start, double stop, double every, int n_items,
                                     const long long *nodes, const int *fields, const int *comps,
                                     double init_temp, int *id)
{
  if (!h || n_items < 0 || (n_items > 0 && (!nodes || !fields || !comps)) || !(every > 0.0)) return DNLG_E_ARG;
  for (int k = 0; k < n_items; k++)
  {
    if (nodes[k] < 0 || nodes[k] >= h->nn || fields[k] < 0 || fields[k] >= HF_NFIELDS || comps[k] < 0)
    {
      h->err = "HistoryFile::append: bad node, field or component";
      return DNLG_E_ARG;
    }
    if (fields[k] >= HF_STRESS)
    {
      // element fields: component 0 (equivalent) or 10 i + j with i, j in 1..3 (FE/Node.C:362-365)
      const int c = comps[k], i = c / 10, j = c % 10;
      if (c != 0 && !(i >= 1 && i <= 3 && j >= 1 && j <= 3))
      {
        h->err = "Node::getNodalValue: No sense for component of a tensor (0, or 11 .. 33)";
        return DNLG_E_ARG;
      }
      if (h->form != DNLG_FORM_FP64)
      {
        h->err = "dnlg_history_file_add: element fields exist in the FP64 form only";
        return DNLG_E_ARG;
      }
      if (fields[k] != HF_STRESS && !h->track_strain)
      {
        // strain fields need the tracked point strains from the first pass on
        if (h->begun || h->inc != 0 || h->mass_done)
        {
          h->err = "dnlg_history_file_add: strain fields need DNLG_OPT_TRACK_STRAIN set before the first step";
          return DNLG_E_STATE;
        }
        const int rc = dnlg_set_option(h, DNLG_OPT_TRACK_STRAIN, 1);
        if (rc) return rc;
      }
      continue;
    }
    const bool vec = fields[k] == HF_SPEED || fields[k] == HF_SPEED_INC || fields[k] == HF_FORCE ||
                     fields[k] == HF_DISPLACEMENT_INC || fields[k] == HF_DISPLACEMENT || fields[k] == HF_COORDS;
    if (vec && comps[k] > 3)
    {
      h->err = "Node::getNodalValue::get: No sense for component >3 for a vectorial quantity";
      return DNLG_E_ARG;
    }
    if (!vec && comps[k] > 1)
    {
      h->err = "Node::getNodalValue::get: No sense for component >1 for a scalar quantity";
      return DNLG_E_ARG;
    }
  }
  // bring the rows and the device state of the present files up to date
  int rc = finalize(h);
  if (rc) return rc;
  StepState s;
  rc = read_state(h, &s);
  if (rc) return rc;
  if (!h->hf.empty())
    DNLG_HIP(h, hipMemcpy(h->hf.data(), h->hf_dev, sizeof(HistFile) * h->hf.size(), hipMemcpyDeviceToHost));
  HistFile f{};
  f.next = start;  // HistoryFile::setSaveTime, FE/HistoryFile.C:241-259
  f.every = every;
  f.stop = stop;
  f.first = (int)h->hitems.size();
  f.count = n_items;
  f.cap = (int)PLOT_CAP;
  DNLG_HIP(h, dalloc(&f.ring, (size_t)PLOT_CAP * (1 + n_items)));
  for (int k = 0; k < n_items; k++)
  {
    HistItem it{(int)nodes[k], fields[k], comps[k], 0, 0, 0};
    if (fields[k] >= HF_STRESS)
    {
      // the node's elements in element order (Node::addElement in Domain::initSolve,
      // FE/Domain.C:1090-1107) with the node's local number in each
      it.el_off = (int)h->hel.size();
      for (long long e = 0; e < h->ne; e++)
        for (int L = 0; L < h->nen; L++)
          if ((long long)(h->fpos_host[e * h->nen + L] % h->nn) == nodes[k]) h->hel.push_back(int2{(int)e, L});
      it.el_cnt = (int)h->hel.size() - it.el_off;
    }
    h->hitems.push_back(it);
  }
  h->hf.push_back(f);
  h->hf_drained.push_back(0);
  h->hf_rows.emplace_back();
  h->init_temp = init_temp;
  if (h->hf_dev) (void)hipFree(h->hf_dev);
  if (h->hi_dev) (void)hipFree(h->hi_dev);
  if (h->hel_dev) (void)hipFree(h->hel_dev);
  h->hf_dev = nullptr;
  h->hi_dev = nullptr;
  h->hel_dev = nullptr;
  DNLG_HIP(h, dalloc(&h->hel_dev, h->hel.size()));
  if (!h->hel.empty())
    DNLG_HIP(h, hipMemcpy(h->hel_dev, h->hel.data(), sizeof(int2) * h->hel.size(), hipMemcpyHostToDevice));
  DNLG_HIP(h, dalloc(&h->hf_dev, h->hf.size()));
  DNLG_HIP(h, dalloc(&h->hi_dev, h->hitems.size()));
  DNLG_HIP(h, hipMemcpy(h->hf_dev, h->hf.data(), sizeof(HistFile) * h->hf.size(), hipMemcpyHostToDevice));
  if (!h->hitems.empty())
    DNLG_HIP(h, hipMemcpy(h->hi_dev, h->hitems.data(), sizeof(HistItem) * h->hitems.size(), hipMemcpyHostToDevice));
  if (id) *id = (int)h->hf.size() - 1;
  return DNLG_OK;
}

extern "C" long long dnlg_history_file_rows(dnlg_handle *h, int id, double *out, long long max_rows)
{
  if (!h || id < 0 || id >= (int)h->hf.size() || max_rows < 0) return -DNLG_E_ARG;
  std::vector<double> &r = h->hf_rows[id];
  const long long w = 1 + h->hf[id].count;
  const long long avail = (long long)r.size() / w;
  if (!out) return avail;
  const long long n = avail < max_rows ? avail : max_rows;
  if (n > 0)
  {
    memcpy(out, r.data(), sizeof(double) * w * n);
    r.erase(r.begin(), r.begin() + w * n);
  }
  return n;
}

extern "C" long long dnlg_plot_rows(dnlg_handle *h, double *out, long long max_rows)
{
  if (!h || max_rows < 0) return -DNLG_E_ARG;
  const long long avail = (long long)(h->plot_host.size() / 4);
  if (!out) return avail;
  const long long n = avail < max_rows ? avail : max_rows;
  if (n > 0)
  {
    memcpy(out, h->plot_host.data(), sizeof(double) * 4 * n);
    h->plot_host.erase(h->plot_host.begin(), h->plot_host.begin() + 4 * n);
  }
  return n;
}

extern "C" int dnlg_pulse_info(dnlg_handle *h, double *out4)
{
  if (!h || !out4) return DNLG_E_ARG;
  double fmax0 = 0.0, conv = 0.0;
  if (h->pulse_ctl)
  {
    PulseCtl c;
    DNLG_HIP(h, hipMemcpyAsync(&c, h->pulse_ctl, sizeof c, hipMemcpyDeviceToHost, h->stream));
    DNLG_HIP(h, hipStreamSynchronize(h->stream));
    fmax0 = c.fmax0;
    conv = c.conv;
  }
  out4[0] = (double)h->pulse_calls;
  out4[1] = (double)h->pulse_iters_last;
  out4[2] = fmax0;
  out4[3] = conv;
  return DNLG_OK;
}

extern "C" int dnlg_solve(dnlg_handle *h, double up_time, long long *steps_done)
{
  if (!h) return DNLG_E_ARG;
  int running = 0;
  int rc = dnlg_begin(h, up_time, &running);
  if (rc) return rc;
  StepState s0;
  rc = read_state(h, &s0);
  if (rc) return rc;
  long long inc0 = s0.increment;
  // chunks of steps between host checks of the loop condition
  int chunk = 16;
  while (running)
  {
    rc = dnlg_step(h, chunk, &running);
    if (rc) return rc;
    if (chunk < 1024) chunk *= 2;
  }
  StepState s;
  rc = read_state(h, &s);
  if (rc) return rc;
  if (steps_done) *steps_done = s.increment - inc0;
  return DNLG_OK;
}

extern "C" long long dnlg_field_size(dnlg_handle *h, int field)
{
  if (!h) return -1;
  const long long nipt = h->ne * h->nip;
  const bool mx = h->form == DNLG_FORM_MIXED;
  switch (field)
  {
  case DNLG_F_COORDS: case DNLG_F_V: case DNLG_F_A: case DNLG_F_DU: case DNLG_F_D: return h->nn * 3;
  case DNLG_F_FINT: return h->nn * h->dim;
  case DNLG_F_MASS: return h->nn;
  case DNLG_F_S: return nipt * (mx ? 6 : 9);
  case DNLG_F_EVP: return nipt;
  case DNLG_F_P: return h->ne;
  case DNLG_F_EPS: case DNLG_F_EPSPLAS: case DNLG_F_DEPS: return h->eps ? nipt * 9 : 0;
  case DNLG_F_FE: return h->ne * h->nen * h->dim;
  case DNLG_F_TEMP: return (h->T || h->T32) ? nipt : 0;
  default: return -1;
  }
}

static double *field_ptr(dnlg_handle *h, int field)
{
  switch (field)
  {
  case DNLG_F_COORDS: return h->x;
  case DNLG_F_V: return h->v;
  case DNLG_F_A: return h->a;
  case DNLG_F_DU: return h->du;
  case DNLG_F_D: return h->d;
  case DNLG_F_FINT: return h->fint;
  case DNLG_F_MASS: return h->mass;
  case DNLG_F_S: return h->S;
  case DNLG_F_EVP: return h->evp;
  case DNLG_F_P: return h->p;
  case DNLG_F_EPS: return h->eps;
  case DNLG_F_EPSPLAS: return h->epsplas;
  case DNLG_F_DEPS: return h->deps;
  case DNLG_F_FE: return h->fe;
  case DNLG_F_TEMP: return h->T;
  default: return nullptr;
  }
}

static float *field_ptr32(dnlg_handle *h, int field)
{
  if (h->form != DNLG_FORM_MIXED) return nullptr;
  switch (field)
  {
  case DNLG_F_S: return h->S32;
  case DNLG_F_EVP: return h->evp32;
  case DNLG_F_P: return h->p32;
  case DNLG_F_FE: return h->fe32;
  case DNLG_F_TEMP: return h->T32;
  default: return nullptr;
  }
}

extern "C" int dnlg_get_field(dnlg_handle *h, int field, double *out, long long n)
{
  if (!h || !out) return DNLG_E_ARG;
  const long long sz = dnlg_field_size(h, field);
  float *src32 = field_ptr32(h, field);
  double *src = src32 ? nullptr : field_ptr(h, field);
  if (sz <= 0 || (!src && !src32) || n != sz)
  {
    h->err = "dnlg_get_field: unknown field or size mismatch";
    return DNLG_E_ARG;
  }
  int rc = finalize(h);
  if (rc) return rc;
  if (field == DNLG_F_FE)
  {
    // slots -> [n_elems*nen][dim] in element order
    const long long tot = (long long)h->maxdeg * h->nn * h->dim;
    std::vector<double> all(tot);
    if (src32)
    {
      std::vector<float> tmp(tot);
      DNLG_HIP(h, hipMemcpyAsync(tmp.data(), src32, sizeof(float) * tot, hipMemcpyDeviceToHost, h->stream));
      DNLG_HIP(h, hipStreamSynchronize(h->stream));
      for (long long k = 0; k < tot; k++) all[k] = (double)tmp[k];
    }
    else
    {
      DNLG_HIP(h, hipMemcpyAsync(all.data(), src, sizeof(double) * tot, hipMemcpyDeviceToHost, h->stream));
      DNLG_HIP(h, hipStreamSynchronize(h->stream));
    }
    for (long long k = 0; k < h->ne * h->nen; k++)
      for (int c = 0; c < h->dim; c++) out[k * h->dim + c] = all[(long long)h->fpos_host[k] * h->dim + c];
    return DNLG_OK;
  }
  if (src32)
  {
    std::vector<float> tmp(n);
    DNLG_HIP(h, hipMemcpyAsync(tmp.data(), src32, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
    DNLG_HIP(h, hipStreamSynchronize(h->stream));
    for (long long k = 0; k < n; k++) out[k] = (double)tmp[k];
    return DNLG_OK;
  }
  DNLG_HIP(h, hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  DNLG_HIP(h, hipStreamSynchronize(h->stream));
  return DNLG_OK;
}

extern "C" int dnlg_set_field(dnlg_handle *h, int field, const double *in, long long n)
{
  if (!h || !in) return DNLG_E_ARG;
  const long long sz = dnlg_field_size(h, field);
  float *dst32 = field_ptr32(h, field);
  double *dst = dst32 ? nullptr : field_ptr(h, field);
  if (sz <= 0 || (!dst && !dst32) || n != sz || field == DNLG_F_FE)
  {
    h->err = "dnlg_set_field: unknown field or size mismatch";
    return DNLG_E_ARG;
  }
  if (dst32)
  {
    std::vector<float> tmp(n);
    for (long long k = 0; k < n; k++) tmp[k] = (float)in[k];
    DNLG_HIP(h, hipMemcpyAsync(dst32, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    DNLG_HIP(h, hipStreamSynchronize(h->stream));
    return DNLG_OK;
  }
  DNLG_HIP(h, hipMemcpyAsync(dst, in, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  DNLG_HIP(h, hipStreamSynchronize(h->stream));
  if (field == DNLG_F_MASS) h->mass_done = true;
  return DNLG_OK;
}

extern "C" int dnlg_set_time(dnlg_handle *h, double t)
{
  if (!h) return DNLG_E_ARG;
  if (h->begun || h->inc != 0)
  {
    h->err = "dnlg_set_time: only on a handle that has not begun (a restart)";
    return DNLG_E_STATE;
  }
  StepState s;
  DNLG_HIP(h, hipMemcpy(&s, h->st, sizeof s, hipMemcpyDeviceToHost));
  s.t = t;
  DNLG_HIP(h, hipMemcpy(h->st, &s, sizeof s, hipMemcpyHostToDevice));
  return DNLG_OK;
}

extern "C" int dnlg_get_time(dnlg_handle *h, double *out)
{
  if (!h || !out) return DNLG_E_ARG;
  int rc = finalize(h);
  if (rc) return rc;
  StepState s;
  rc = read_state(h, &s);
  if (rc) return rc;
  out[0] = s.t;
  out[1] = s.dt;
  out[2] = (double)s.increment;
  out[3] = s.computedTimeStep;
  out[4] = s.ke;
  out[5] = s.alphaM;
  out[6] = s.beta;
  out[7] = s.gamma;
  out[8] = s.omegaS;
  return DNLG_OK;
}

extern "C" void *dnlg_get_stream(dnlg_handle *h) { return h ? (void *)h->stream : nullptr; }

extern "C" int dnlg_enable_kernel_timing(dnlg_handle *h, int on)
{
  if (!h) return DNLG_E_ARG;
  DNLG_HIP(h, hipStreamSynchronize(h->stream));
  h->timing = on != 0;
  h->ev_used = 0;
  h->ev_elem.clear();
  h->ev_nodal.clear();
  return DNLG_OK;
}

extern "C" int dnlg_kernel_stats(dnlg_handle *h, int which, long long *launches, double *total_ms)
{
  if (!h || (which != 0 && which != 1)) return DNLG_E_ARG;
  DNLG_HIP(h, hipStreamSynchronize(h->stream));
  const std::vector<hipEvent_t> &ev = which == 0 ? h->ev_elem : h->ev_nodal;
  double tot = 0.0;
  for (size_t k = 0; k + 1 < ev.size(); k += 2)
  {
    float ms = 0.f;
    DNLG_HIP(h, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
    tot += ms;
  }
  if (launches) *launches = (long long)(ev.size() / 2);
  if (total_ms) *total_ms = tot;
  return DNLG_OK;
}
