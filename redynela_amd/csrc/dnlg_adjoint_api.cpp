// C ABI of the adjoint: forward with checkpoints, reverse sweep with
// segment-wise recomputation.  The forward increments are the production
// kernels of the single-domain path (the same bits as dnlg_step); the state
// after every ckpt_every-th increment is copied to device memory.  The
// backward sweep restores each checkpoint in turn (last first), recomputes its
// segment while recording every increment's state, and runs the reverse
// increments of dnlg_adjoint.hip over it.
#include "dnlg_handle.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace dnl {
void launch_adjoint_step(const AdjArgs &g, hipStream_t s);
bool launch_copy_list(const double *const *src, double *const *dst, const long long *n, int count, hipStream_t s);

// one state snapshot: x, v, a, du, d ([nn][3] each), S [9][nip ne], evp [nip ne], p [ne], StepState,
// and with the Johnson-Cook temperature factor the point temperature T [nip ne]
struct Snap
{
  long long x, v, a, du, d, S, evp, p, st, T, size;  // offsets and total size, in doubles (T < 0: none)
};

struct AdjStore
{
  long long nsteps = 0;
  int every = 0, nseg = 0;
  Snap lay{};
  double *ckpt = nullptr;   // nseg snapshots: after increments 0, every, 2 every, ...
  double *tape = nullptr;   // every + 1 snapshots of the segment being reversed
  double *fin = nullptr;    // the state after the last increment
  double *xb = nullptr, *ub = nullptr, *vb = nullptr, *ab = nullptr, *Sb = nullptr, *evb = nullptr, *pb = nullptr;
  double *Tb = nullptr;  // adjoint of the point temperature (Johnson-Cook temperature factor)
  double *fb = nullptr, *slot_x = nullptr, *slot_u = nullptr, *dt_part = nullptr, *th_part = nullptr;
  double *am_len = nullptr;
  long long *am_idx = nullptr;
  AdjScalars *sc = nullptr;
  float fwd_ms = 0.f;
  size_t bytes = 0;
};

void adj_free(AdjStore *s)
{
  if (!s) return;
  void *ptrs[] = {s->ckpt, s->tape, s->fin, s->xb, s->ub, s->vb, s->ab, s->Sb, s->evb, s->pb, s->Tb, s->fb,
                  s->slot_x, s->slot_u, s->dt_part, s->th_part, s->am_len, s->am_idx, s->sc};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete s;
}
}  // namespace dnl

static_assert(sizeof(StepState) % 8 == 0, "StepState is copied as doubles");

// the arrays of one snapshot, in doubles: x, v, a, du, d, S, evp, p, StepState (, T)
static int snap_copy(dnlg_handle *h, const Snap &L, double *snap, bool to_snap)
{
  const long long n3 = h->nn * 3, nipt = h->ne * h->nip;
  double *state[10] = {h->x, h->v, h->a, h->du, h->d, h->S, h->evp, h->p, (double *)h->st, h->T};
  const long long off[10] = {L.x, L.v, L.a, L.du, L.d, L.S, L.evp, L.p, L.st, L.T};
  const long long cnt[10] = {n3, n3, n3, n3, n3, 9 * nipt, nipt, h->ne, (long long)(sizeof(StepState) / 8), nipt};
  const int count = L.T >= 0 ? 10 : 9;
  const double *src[10];
  double *dst[10];
  for (int k = 0; k < count; k++)
  {
    src[k] = to_snap ? state[k] : snap + off[k];
    dst[k] = to_snap ? snap + off[k] : state[k];
  }
  if (!launch_copy_list(src, dst, cnt, count, h->stream))
  {
    h->err = "dnlg_adjoint: snapshot array count out of range";
    return DNLG_E_ARG;
  }
  DNLG_HIP(h, hipGetLastError());
  return DNLG_OK;
}
static int snap_to(dnlg_handle *h, const Snap &L, double *dst) { return snap_copy(h, L, dst, true); }
static int snap_from(dnlg_handle *h, const Snap &L, const double *src) { return snap_copy(h, L, (double *)src, false); }

static int adj_check(dnlg_handle *h)
{
  const bool law_ok = h->ea.law == 1 || (h->ea.law == 2 && h->ea.C == 0.0 && h->T);
  if ((h->dim != DNLG_DIM_AXI && h->dim != DNLG_DIM_3D) || h->form != DNLG_FORM_FP64 || !law_ok || h->has_bcs ||
      h->ts_method != DNLG_TS_COURANT)
  {
    h->err = "dnlg_adjoint: axisymmetric or hex8 mesh, FP64 form, hardening A + B evp^n (law 1) or Johnson-Cook "
             "with the temperature factor and no rate term (law 2, C = 0), no speed BC, Courant time step";
    return DNLG_E_ARG;
  }
  return DNLG_OK;
}

extern "C" int dnlg_adjoint_forward(dnlg_handle *h, long long nsteps, int ckpt_every)
{
  if (!h || nsteps <= 0 || ckpt_every <= 0) return DNLG_E_ARG;
  int rc = adj_check(h);
  if (rc) return rc;
  StepState s0;
  DNLG_HIP(h, hipMemcpy(&s0, h->st, sizeof s0, hipMemcpyDeviceToHost));
  if (s0.increment != 0 || h->begun)
  {
    h->err = "dnlg_adjoint_forward: needs a handle that has not stepped yet";
    return DNLG_E_STATE;
  }
  // the reverse sweep takes every one of the nsteps increments as a full
  // Courant / LINEARdbl step: an end time that stops the loop early would
  // leave no-op increments behind it
  if (s0.endTime < 1e300)
  {
    h->err = "dnlg_adjoint_forward: the handle's end_time must be unbounded (1e300); run the increments wanted with nsteps";
    return DNLG_E_ARG;
  }
  adj_free(h->adj);
  h->adj = nullptr;
  AdjStore *A = new AdjStore();
  A->nsteps = nsteps;
  A->every = (int)std::min<long long>(ckpt_every, nsteps);
  A->nseg = (int)((nsteps + A->every - 1) / A->every);
  const long long n3 = h->nn * 3, nipt = h->ne * h->nip;
  Snap &L = A->lay;
  L.x = 0;
  L.v = L.x + n3;
  L.a = L.v + n3;
  L.du = L.a + n3;
  L.d = L.du + n3;
  L.S = L.d + n3;
  L.evp = L.S + 9 * nipt;
  L.p = L.evp + nipt;
  L.st = L.p + h->ne;
  const bool jc = h->ea.law == 2;
  L.T = jc ? L.st + (long long)((sizeof(StepState) + 7) / 8) : -1;
  L.size = jc ? L.T + nipt : L.st + (long long)((sizeof(StepState) + 7) / 8);
  const int nbn = (int)((h->nn + 255) / 256), nbe = (int)((nipt + 63) / 64);
  const int dim = h->dim;
  const long long nslot = (long long)h->maxdeg * h->nn * dim;
  h->adj = A;
#define ACK(call)                                                                   \
  do                                                                                \
  {                                                                                 \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess)                                                           \
    {                                                                               \
      h->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #call;    \
      adj_free(h->adj);                                                             \
      h->adj = nullptr;                                                             \
      return DNLG_E_HIP;                                                            \
    }                                                                               \
  } while (0)
  ACK(dalloc(&A->ckpt, (size_t)L.size * A->nseg));
  // with a checkpoint after every increment the checkpoints are the tape
  if (A->every > 1) ACK(dalloc(&A->tape, (size_t)L.size * (A->every + 1)));
  ACK(dalloc(&A->fin, (size_t)L.size));
  ACK(dalloc(&A->xb, h->nn * dim));
  ACK(dalloc(&A->ub, h->nn * dim));
  ACK(dalloc(&A->vb, h->nn * dim));
  ACK(dalloc(&A->ab, h->nn * dim));
  ACK(dalloc(&A->fb, h->nn * dim));
  ACK(dalloc(&A->Sb, 9 * nipt));
  ACK(dalloc(&A->evb, nipt));
  ACK(dalloc(&A->pb, h->ne));
  if (jc) ACK(dalloc(&A->Tb, nipt));
  ACK(dalloc(&A->slot_x, nslot));
  ACK(dalloc(&A->slot_u, nslot));
  ACK(dalloc(&A->dt_part, nbn));
  ACK(dalloc(&A->th_part, 3 * nbe));
  const int nba = (int)((h->ne + 255) / 256);
  ACK(dalloc(&A->am_len, nba));
  ACK(dalloc(&A->am_idx, nba));
  ACK(dalloc(&A->sc, 1));
#undef ACK
  A->bytes = sizeof(double) * ((size_t)L.size * (A->nseg + (A->every > 1 ? A->every + 1 : 0) + 1) + h->nn * 5 * dim +
                               (jc ? 11 : 10) * nipt + h->ne + 2 * nslot + nbn + 3 * nbe + nba) +
             sizeof(long long) * nba + sizeof(AdjScalars);
  hipEvent_t e0, e1;
  DNLG_HIP(h, hipEventCreate(&e0));
  DNLG_HIP(h, hipEventCreate(&e1));
  DNLG_HIP(h, hipEventRecord(e0, h->stream));
  int running = 0;
  rc = dnlg_begin(h, 1e300, &running);
  if (rc) return rc;
  for (int seg = 0; seg < A->nseg; seg++)
  {
    rc = snap_to(h, L, A->ckpt + (size_t)seg * L.size);
    if (rc) return rc;
    const long long n = std::min<long long>(A->every, nsteps - (long long)seg * A->every);
    rc = dnlg_step_async(h, (int)n);
    if (rc) return rc;
  }
  rc = snap_to(h, L, A->fin);
  if (rc) return rc;
  DNLG_HIP(h, hipEventRecord(e1, h->stream));
  rc = dnlg_synchronize(h);
  if (rc) return rc;
  StepState s1;
  DNLG_HIP(h, hipMemcpy(&s1, h->st, sizeof s1, hipMemcpyDeviceToHost));
  if (s1.increment != nsteps || !s1.running)
  {
    h->err = "dnlg_adjoint_forward: the solve loop did not run all nsteps increments";
    return DNLG_E_STATE;
  }
  DNLG_HIP(h, hipEventElapsedTime(&A->fwd_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return DNLG_OK;
}

extern "C" int dnlg_adjoint_backward(dnlg_handle *h, const double *xbar, double *grad, dnlg_adjoint_stats *stats)
{
  if (!h || !xbar || !grad) return DNLG_E_ARG;
  AdjStore *A = h->adj;
  if (!A)
  {
    h->err = "dnlg_adjoint_backward: call dnlg_adjoint_forward first";
    return DNLG_E_STATE;
  }
  const Snap &L = A->lay;
  const long long nn = h->nn, nipt = h->ne * h->nip;
  hipStream_t s = h->stream;
  // seed: adjoint of the final coordinates ((r, z) or (x, y, z)); everything else zero
  const int dim = h->dim;
  std::vector<double> seed(nn * dim);
  for (long long i = 0; i < nn; i++)
    for (int c = 0; c < dim; c++) seed[dim * i + c] = xbar[3 * i + c];
  DNLG_HIP(h, hipMemcpyAsync(A->xb, seed.data(), sizeof(double) * nn * dim, hipMemcpyHostToDevice, s));
  DNLG_HIP(h, hipMemsetAsync(A->ub, 0, sizeof(double) * nn * dim, s));
  DNLG_HIP(h, hipMemsetAsync(A->vb, 0, sizeof(double) * nn * dim, s));
  DNLG_HIP(h, hipMemsetAsync(A->ab, 0, sizeof(double) * nn * dim, s));
  DNLG_HIP(h, hipMemsetAsync(A->Sb, 0, sizeof(double) * 9 * nipt, s));
  DNLG_HIP(h, hipMemsetAsync(A->evb, 0, sizeof(double) * nipt, s));
  DNLG_HIP(h, hipMemsetAsync(A->pb, 0, sizeof(double) * h->ne, s));
  if (A->Tb) DNLG_HIP(h, hipMemsetAsync(A->Tb, 0, sizeof(double) * nipt, s));
  DNLG_HIP(h, hipMemsetAsync(A->sc, 0, sizeof(AdjScalars), s));
  StepState s0;
  DNLG_HIP(h, hipMemcpyAsync(&s0, A->ckpt + L.st, sizeof s0, hipMemcpyDeviceToHost, s));
  DNLG_HIP(h, hipStreamSynchronize(s));

  AdjArgs g{};
  g.ea = h->ea;
  g.bc = h->bc;
  g.deg = h->deg;
  g.mass = h->mass;
  g.nn = nn;
  g.ne = h->ne;
  g.dim = dim;
  g.xb = A->xb;
  g.ub = A->ub;
  g.vb = A->vb;
  g.ab = A->ab;
  g.Sb = A->Sb;
  g.evb = A->evb;
  g.pb = A->pb;
  g.jc = L.T >= 0 ? 1 : 0;
  g.Tb = A->Tb;
  g.fb = A->fb;
  g.slot_x = A->slot_x;
  g.slot_u = A->slot_u;
  g.dt_part = A->dt_part;
  g.th_part = A->th_part;
  g.ndt_part = (int)((nn + 255) / 256);
  g.nth_part = (int)((nipt + 63) / 64);
  g.am_len = A->am_len;
  g.am_idx = A->am_idx;
  g.nam_part = (int)((h->ne + 255) / 256);
  g.sc = A->sc;
  g.K = h->ea.K;
  g.G = h->ea.G;
  g.A = h->ea.A;
  g.B = h->ea.B;
  g.n = h->ea.n;
  g.alphaM = s0.alphaM;
  g.beta = s0.beta;
  g.gamma = s0.gamma;
  g.safety = s0.safety;
  g.omegaS = s0.omegaS;
  g.cs = s0.cs;
  g.freq = s0.freq;

  hipEvent_t e0, e1, e2;
  DNLG_HIP(h, hipEventCreate(&e0));
  DNLG_HIP(h, hipEventCreate(&e1));
  DNLG_HIP(h, hipEventCreate(&e2));
  DNLG_HIP(h, hipEventRecord(e0, s));
  float recompute_ms = 0.f;
  // the reverse increment k over the states after increments k and k + 1
  auto reverse = [&](long long k, const double *t0, const double *t1) {
    g.x = t0 + L.x;
    g.v = t0 + L.v;
    g.a = t0 + L.a;
    g.du = t0 + L.du;
    g.S = t0 + L.S;
    g.evp = t0 + L.evp;
    g.p = t0 + L.p;
    g.T = L.T >= 0 ? t0 + L.T : nullptr;
    g.a1 = t1 + L.a;
    g.st0 = (const StepState *)(t0 + L.st);
    g.st1 = (const StepState *)(t1 + L.st);
    g.k = k;
    launch_adjoint_step(g, s);
  };
  if (A->every == 1)
  {
    // a checkpoint after every increment: the sweep reads them directly, with
    // no recomputation (the states are the ones a recomputed segment records)
    for (long long k = A->nsteps - 1; k >= 0; k--)
      reverse(k, A->ckpt + (size_t)k * L.size, k + 1 < A->nsteps ? A->ckpt + (size_t)(k + 1) * L.size : A->fin);
    DNLG_HIP(h, hipGetLastError());
  }
  for (int seg = A->every == 1 ? -1 : A->nseg - 1; seg >= 0; seg--)
  {
    const long long k0 = (long long)seg * A->every, k1 = std::min<long long>(k0 + A->every, A->nsteps);
    // recompute the segment, recording every increment's state
    hipEvent_t r0, r1;
    DNLG_HIP(h, hipEventCreate(&r0));
    DNLG_HIP(h, hipEventCreate(&r1));
    DNLG_HIP(h, hipEventRecord(r0, s));
    int rc = snap_from(h, L, A->ckpt + (size_t)seg * L.size);
    if (rc) return rc;
    h->inc = k0;
    h->pending = k0 > 0;
    h->begun = true;
    rc = snap_to(h, L, A->tape);
    if (rc) return rc;
    for (long long j = 1; j <= k1 - k0; j++)
    {
      rc = dnlg_step_async(h, 1);
      if (rc) return rc;
      rc = snap_to(h, L, A->tape + (size_t)j * L.size);
      if (rc) return rc;
    }
    DNLG_HIP(h, hipEventRecord(r1, s));
    // reverse increments k1 .. k0+1
    for (long long k = k1 - 1; k >= k0; k--)
    {
      const double *t0 = A->tape + (size_t)(k - k0) * L.size;
      reverse(k, t0, t0 + L.size);
    }
    DNLG_HIP(h, hipGetLastError());
    DNLG_HIP(h, hipStreamSynchronize(s));
    float ms = 0.f;
    DNLG_HIP(h, hipEventElapsedTime(&ms, r0, r1));
    recompute_ms += ms;
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
  }
  DNLG_HIP(h, hipEventRecord(e1, s));
  AdjScalars sc;
  DNLG_HIP(h, hipMemcpyAsync(&sc, A->sc, sizeof sc, hipMemcpyDeviceToHost, s));
  // leave the handle in its post-forward state
  int rc = snap_from(h, L, A->fin);
  if (rc) return rc;
  h->inc = A->nsteps;
  h->pending = true;
  h->begun = true;
  DNLG_HIP(h, hipEventRecord(e2, s));
  rc = dnlg_synchronize(h);
  if (rc) return rc;
  float bwd_ms = 0.f;
  DNLG_HIP(h, hipEventElapsedTime(&bwd_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  for (int j = 0; j < 3; j++) grad[j] = sc.th[j];
  if (stats)
  {
    stats->forward_ms = A->fwd_ms;
    stats->backward_ms = bwd_ms;
    stats->recompute_ms = recompute_ms;
    stats->device_bytes = (long long)A->bytes;
    stats->checkpoint_bytes = (long long)(sizeof(double) * (size_t)L.size * A->nseg);
    stats->n_checkpoints = A->nseg;
    stats->steps = A->nsteps;
  }
  return DNLG_OK;
}
