// C ABI of the batched mode: many independent axisymmetric domains (Taylor
// bars of uneven size and parameters) advanced by one persistent kernel, one
// workgroup per domain, each with its own time step and increment count.
// There is no reference counterpart: the reference solves one Structure per
// process (FE/Structure.C:809-897).
#include "../../include/dnlg.h"
#include "dnlg_internal.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace dnl {
void launch_batch_f64(const BatchArgs &B, int nbars, hipStream_t s);
void launch_batch_f32(const BatchArgs &B, int nbars, hipStream_t s);
void launch_bstep_elem_f64(const BatchArgs &B, const int2 *blk, int nblk, int preamble, int tile, hipStream_t s);
void launch_bstep_elem_f32(const BatchArgs &B, const int2 *blk, int nblk, int preamble, int tile, hipStream_t s);
void launch_bstep_dt(const BatchArgs &B, int nb, int preamble, hipStream_t s);
void launch_bstep_nodal(const BatchArgs &B, const int2 *blk, int nblk, bool fe32, int tile, hipStream_t s);
void launch_mass_f64(int dim, const ElemArgs &a, double rho, const double *hn, double *me, hipStream_t s);
void launch_gather_scalar(const unsigned char *deg, const double *src, double *dst, long long nn, hipStream_t s);
}  // namespace dnl

using namespace dnl;

struct dnlg_batch
{
  int nb = 0, form = DNLG_FORM_FP64;
  long long NN = 0, NE = 0;
  hipStream_t stream = nullptr;
  std::vector<BarDesc> bars;
  std::vector<int> order;
  double *x = nullptr, *v = nullptr, *a = nullptr, *du = nullptr, *d = nullptr, *mass = nullptr, *me = nullptr;
  double *fe = nullptr, *S = nullptr, *evp = nullptr, *p = nullptr;
  float *fe32 = nullptr, *S32 = nullptr, *evp32 = nullptr, *p32 = nullptr;
  int *conn = nullptr, *fpos = nullptr, *order_d = nullptr;
  unsigned char *deg = nullptr, *bc = nullptr;
  BarDesc *bars_d = nullptr;
  ElemArgs *bar_ea = nullptr;
  ElemArgs32 *bar_ea32 = nullptr;
  StepState *st = nullptr;
  unsigned long long *t0 = nullptr, *t1 = nullptr;
  BatchArgs args{};
  // stepped engine: block tables (bar, first element / node), largest bars first
  // (device tables hold the blocks of the running bars only; they are
  // compacted on the host as bars finish, and the graph is re-captured)
  int2 *blk_e = nullptr, *blk_n = nullptr;
  int nblk_e = 0, nblk_n = 0;
  std::vector<std::vector<int2>> bar_be[2], bar_bn[2];  // every block of each bar, per tile
  unsigned long long *crit = nullptr;
  hipGraphExec_t graph = nullptr;  // kGraphSteps increments over the current tables
  int graph_nblk_e = -1, graph_nblk_n = -1;
  int engine = DNLG_BATCH_AUTO;
  int tile = 0;                    // block tile of the stepped engine: 0 = 64 lanes, 1 = 256 lanes
  // the autotuner's choice for DNLG_BATCH_AUTO, made at the first solve
  bool tuned = false;
  int auto_engine = DNLG_BATCH_STEPPED;
  double probe_ms[3] = {0, 0, 0};  // stepped, stepped-wide, persistent
  bool mass_done = false;
  double rho = 0;
  std::string err;
};

static thread_local std::string g_berr;

#define BHIP(call)                                                                 \
  do                                                                               \
  {                                                                                \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
    {                                                                              \
      b->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #call;   \
      return DNLG_E_HIP;                                                           \
    }                                                                              \
  } while (0)

template <typename T>
static hipError_t balloc(T **p, size_t n)
{
  return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

static constexpr int kGraphSteps = 8;

static void bfree(dnlg_batch *b)
{
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  void *ptrs[] = {b->x, b->v, b->a, b->du, b->d, b->mass, b->me, b->fe, b->S, b->evp, b->p, b->fe32, b->S32,
                  b->evp32, b->p32, b->conn, b->fpos, b->order_d, b->deg, b->bc, b->bars_d, b->st, b->t0, b->t1,
                  b->blk_e, b->blk_n, b->crit, b->bar_ea, b->bar_ea32};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

extern "C" const char *dnlg_batch_last_error(dnlg_batch *b) { return b ? b->err.c_str() : g_berr.c_str(); }

// batched-bar workload parameters: bar b draws, in this order, nr ~ U{5..20},
// v0 ~ U(150, 300) m/s, sigma_y0 ~ U(300, 500) MPa, hardening modulus ~ U(50, 300) MPa
// from std::mt19937_64(seed); U(a,b) = a + (b-a) * (draw >> 11) * 2^-53.
extern "C" int dnlg_batch_taylor_params(int n, unsigned long long seed, int *nr, double *v0, double *sy0, double *hmod)
{
  if (n < 0 || !nr || !v0 || !sy0 || !hmod) return DNLG_E_ARG;
  std::mt19937_64 g(seed);
  auto u = [&]() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); };
  for (int i = 0; i < n; i++)
  {
    nr[i] = 5 + (int)(g() % 16);
    v0[i] = 150.0 + 150.0 * u();
    sy0[i] = 300e6 + 200e6 * u();
    hmod[i] = 50e6 + 250e6 * u();
  }
  return DNLG_OK;
}

extern "C" int dnlg_batch_create(int n, const dnlg_mesh *meshes, const dnlg_material *mats, const dnlg_bc *bcs,
                                 const dnlg_solver_params *solver, dnlg_batch **out)
{
  if (!out) return DNLG_E_ARG;
  *out = nullptr;
  if (n <= 0 || !meshes || !mats || !solver || (solver->form != DNLG_FORM_FP64 && solver->form != DNLG_FORM_MIXED))
  {
    g_berr = "dnlg_batch_create: bad arguments";
    return DNLG_E_ARG;
  }
  for (int i = 0; i < n; i++)
  {
    const dnlg_mesh &m = meshes[i];
    const dnlg_material &t = mats[i];
    if (m.dim != DNLG_DIM_AXI || m.n_nodes <= 0 || m.n_elems <= 0 || !m.coords || !m.conn)
    {
      g_berr = "dnlg_batch_create: every domain must be an axisymmetric quad4 mesh";
      return DNLG_E_ARG;
    }
    if (t.young != mats[0].young || t.poisson != mats[0].poisson || t.density != mats[0].density ||
        t.hardening > 1 || t.hardening < 0)
    {
      g_berr = "dnlg_batch_create: domains share E, nu and density; hardening 0 or 1";
      return DNLG_E_ARG;
    }
    for (long long k = 0; k < m.n_elems * 4; k++)
      if (m.conn[k] < 0 || m.conn[k] >= m.n_nodes)
      {
        g_berr = "dnlg_batch_create: connectivity references a node outside its domain";
        return DNLG_E_ARG;
      }
  }
  if (!(solver->rho_b >= 0.0 && solver->rho_b <= 1.0))
  {
    g_berr = "ExplicitSolverCH::setDissipation: dissipation value out of range [0.0, 1.0]";
    return DNLG_E_ARG;
  }
  dnlg_batch *b = new dnlg_batch();
  b->nb = n;
  b->form = solver->form;
  b->rho = mats[0].density;
  b->bars.resize(n);
  long long NN = 0, NE = 0;
  for (int i = 0; i < n; i++)
  {
    BarDesc &d = b->bars[i];
    d.node0 = NN;
    d.nn = meshes[i].n_nodes;
    d.elem0 = NE;
    d.ne = meshes[i].n_elems;
    const double E = mats[i].young, nu = mats[i].poisson;
    d.K = E / (3. * (1. - 2. * nu));
    d.G = E / (2. * (1. + nu));
    d.A = mats[i].A;
    d.B = mats[i].B;
    d.n = mats[i].n;
    d.nm1 = mats[i].n - 1.;
    d.law = mats[i].hardening ? 1 : 0;
    NN += d.nn;
    NE += d.ne;
  }
  b->NN = NN;
  b->NE = NE;
  if (NE * 4 > 2147483647LL || NN * 4 > 2147483647LL)
  {
    g_berr = "dnlg_batch_create: batch too large";
    delete b;
    return DNLG_E_ARG;
  }
  // packed host arrays
  std::vector<double> X(NN * 3), V(NN * 3, 0.0);
  std::vector<int> conn(NE * 4);
  std::vector<unsigned char> bc(NN, 0), deg(NN, 0);
  for (int i = 0; i < n; i++)
  {
    const BarDesc &d = b->bars[i];
    std::copy(meshes[i].coords, meshes[i].coords + d.nn * 3, X.begin() + d.node0 * 3);
    if (bcs && bcs[i].v0) std::copy(bcs[i].v0, bcs[i].v0 + d.nn * 3, V.begin() + d.node0 * 3);
    if (bcs && bcs[i].restrain) std::copy(bcs[i].restrain, bcs[i].restrain + d.nn, bc.begin() + d.node0);
    for (long long k = 0; k < d.ne * 4; k++) conn[d.elem0 * 4 + k] = (int)(meshes[i].conn[k] + d.node0);
  }
  // element-force slots in element order per node (see dnlg_create)
  std::vector<int> fpos(NE * 4);
  int md = 0;
  {
    std::vector<int> cnt(NN, 0);
    for (long long k = 0; k < NE * 4; k++) cnt[conn[k]]++;
    for (long long i = 0; i < NN; i++) md = std::max(md, cnt[i]);
    for (long long e = 0; e < NE; e++)
      for (int k = 0; k < 4; k++)
      {
        const int node = conn[e * 4 + k];
        fpos[e * 4 + k] = (int)((long long)deg[node] * NN + node);
        deg[node]++;
      }
  }
  // biggest work first: elements x (1 / smallest element size ~ nr)
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int p, int q) { return b->bars[p].ne > b->bars[q].ne; });
  // per-domain scalar state
  std::vector<StepState> st(n);
  for (int i = 0; i < n; i++)
  {
    StepState s{};
    const double rb = solver->rho_b;
    s.alphaM = (2.0 * rb - 1.0) / (1.0 + rb);
    s.beta = (5.0 - 3.0 * rb) / ((1.0 + rb) * (1.0 + rb) * (2.0 - rb));
    s.gamma = 1.5 - s.alphaM;
    s.omegaS = sqrt((12.0 * pow((1 + rb), 3) * (2.0 - rb)) / (10 + 15 * rb - rb * rb + pow(rb, 3) - pow(rb, 4)));
    s.safety = solver->safety;
    s.endTime = solver->end_time;
    s.freq = solver->ts_freq > 0 ? solver->ts_freq : 10;
    const double E = mats[i].young, Nu = mats[i].poisson;
    double ro = 0.;
    for (int k = 0; k < 4; k++) ro += mats[i].density;
    ro /= 4;
    s.cs = sqrt((E * (1 - Nu)) / (ro * (1 + Nu) * (1 - 2 * Nu)));
    st[i] = s;
  }
  const long long nipt = NE * 4;
  std::vector<double> evp0(nipt, 1e-10);
  std::vector<float> evp32(nipt, 1e-10f);
#define BCK(call)                                                                   \
  do                                                                                \
  {                                                                                 \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess)                                                           \
    {                                                                               \
      g_berr = std::string("HIP error ") + hipGetErrorString(_e) + " at " #call;    \
      bfree(b);                                                                     \
      delete b;                                                                     \
      return DNLG_E_HIP;                                                            \
    }                                                                               \
  } while (0)
  BCK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BCK(balloc(&b->x, NN * 3));
  BCK(balloc(&b->v, NN * 3));
  BCK(balloc(&b->a, NN * 3));
  BCK(balloc(&b->du, NN * 3));
  BCK(balloc(&b->d, NN * 3));
  BCK(balloc(&b->mass, NN));
  BCK(balloc(&b->me, (long long)md * NN));
  BCK(balloc(&b->conn, NE * 4));
  BCK(balloc(&b->fpos, NE * 4));
  BCK(balloc(&b->deg, NN));
  BCK(balloc(&b->bc, NN));
  BCK(balloc(&b->bars_d, n));
  BCK(balloc(&b->order_d, n));
  BCK(balloc(&b->st, n));
  BCK(balloc(&b->t0, n));
  BCK(balloc(&b->t1, n));
  if (b->form == DNLG_FORM_FP64)
  {
    BCK(balloc(&b->fe, (long long)md * NN * 2));
    BCK(balloc(&b->S, nipt * 9));
    BCK(balloc(&b->evp, nipt));
    BCK(balloc(&b->p, NE));
    BCK(hipMemset(b->S, 0, sizeof(double) * nipt * 9));
    BCK(hipMemcpy(b->evp, evp0.data(), sizeof(double) * nipt, hipMemcpyHostToDevice));
    BCK(hipMemset(b->p, 0, sizeof(double) * NE));
  }
  else
  {
    BCK(balloc(&b->fe32, (long long)md * NN * 2));
    BCK(balloc(&b->S32, nipt * 6));
    BCK(balloc(&b->evp32, nipt));
    BCK(balloc(&b->p32, NE));
    BCK(hipMemset(b->S32, 0, sizeof(float) * nipt * 6));
    BCK(hipMemcpy(b->evp32, evp32.data(), sizeof(float) * nipt, hipMemcpyHostToDevice));
    BCK(hipMemset(b->p32, 0, sizeof(float) * NE));
  }
  std::vector<double> z(NN * 3, 0.0);
  BCK(hipMemcpy(b->x, X.data(), sizeof(double) * NN * 3, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->v, V.data(), sizeof(double) * NN * 3, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->a, z.data(), sizeof(double) * NN * 3, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->du, z.data(), sizeof(double) * NN * 3, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->d, z.data(), sizeof(double) * NN * 3, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->conn, conn.data(), sizeof(int) * NE * 4, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->fpos, fpos.data(), sizeof(int) * NE * 4, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->deg, deg.data(), NN, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->bc, bc.data(), NN, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->bars_d, b->bars.data(), sizeof(BarDesc) * n, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->order_d, b->order.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  BCK(hipMemcpy(b->st, st.data(), sizeof(StepState) * n, hipMemcpyHostToDevice));
  {
    // stepped engine block tables, per tile: blocks of 64 (tile 0) or 256 (tile 1)
    // lanes; FP64 puts 4 points of an element on 4 lanes, FP32 one element on
    // one lane; nodes one per lane
    long long tot_e = 0, tot_n = 0;
    for (int t = 0; t < 2; t++)
    {
      const int lanes = t ? 256 : 64;
      const int per_e = b->form == DNLG_FORM_FP64 ? lanes / 4 : lanes;
      b->bar_be[t].resize(n);
      b->bar_bn[t].resize(n);
      long long te = 0, tn = 0;
      for (int i = 0; i < n; i++)
      {
        const BarDesc &d = b->bars[i];
        for (long long e = 0; e < d.ne; e += per_e) b->bar_be[t][i].push_back(make_int2(i, (int)(d.elem0 + e)));
        for (long long q = 0; q < d.nn; q += lanes) b->bar_bn[t][i].push_back(make_int2(i, (int)(d.node0 + q)));
        te += (long long)b->bar_be[t][i].size();
        tn += (long long)b->bar_bn[t][i].size();
      }
      tot_e = std::max(tot_e, te);
      tot_n = std::max(tot_n, tn);
    }
    std::vector<unsigned long long> inf(n, 0x7FF0000000000000ull);
    BCK(balloc(&b->blk_e, tot_e));
    BCK(balloc(&b->blk_n, tot_n));
    BCK(balloc(&b->crit, n));
    BCK(hipMemcpy(b->crit, inf.data(), sizeof(unsigned long long) * n, hipMemcpyHostToDevice));
  }
#undef BCK
  // kernel arguments over the packed arrays
  BatchArgs &A = b->args;
  A.bars = b->bars_d;
  A.order = b->order_d;
  A.states = b->st;
  A.t0 = b->t0;
  A.t1 = b->t1;
  A.crit = b->crit;
  ElemArgs &ea = A.ea;
  ea.conn = b->conn;
  ea.x = b->x;
  ea.du = b->du;
  ea.S = b->S;
  ea.evp = b->evp;
  ea.p = b->p;
  ea.fe = b->fe;
  ea.fpos = b->fpos;
  ea.ne = NE;
  ea.c0 = (3.0 - sqrt(3.0)) / 12.0;
  ea.c1 = (3.0 + sqrt(3.0)) / 12.0;
  ea.n0 = (2.0 - sqrt(3.0)) / 6.0;
  ea.n1 = 1.0 / 6.0;
  ea.n2 = (2.0 + sqrt(3.0)) / 6.0;
  ea.law = 1;
  ElemArgs32 &e3 = A.ea32;
  e3.conn = b->conn;
  e3.x = b->x;
  e3.du = b->du;
  e3.S = b->S32;
  e3.evp = b->evp32;
  e3.T = nullptr;
  e3.p = b->p32;
  e3.fe = b->fe32;
  e3.fpos = b->fpos;
  e3.ne = NE;
  NodalArgs &na = A.na;
  na.x = b->x;
  na.v = b->v;
  na.a = b->a;
  na.du = b->du;
  na.d = b->d;
  na.mass = b->mass;
  na.bc = b->bc;
  na.deg = b->deg;
  na.fe = b->fe;
  na.fe32 = b->fe32;
  na.nn = NN;
  // per-bar element arguments: the packed arrays with the bar's state and constants
  {
    std::vector<ElemArgs> be(n);
    std::vector<ElemArgs32> be32(n);
    for (int i = 0; i < n; i++)
    {
      const BarDesc &d = b->bars[i];
      ElemArgs x = ea;
      x.st = b->st + i;
      x.K = d.K; x.G = d.G; x.A = d.A; x.B = d.B; x.n = d.n; x.nm1 = d.nm1;
      x.yield = d.law ? 1 : 0;
      x.law = d.law;
      be[i] = x;
      ElemArgs32 y = e3;
      y.st = b->st + i;
      y.K = (float)d.K; y.G = (float)d.G; y.A = (float)d.A; y.B = (float)d.B; y.n = (float)d.n;
      y.law = d.law;
      be32[i] = y;
    }
    hipError_t e1 = balloc(&b->bar_ea, n), e2 = balloc(&b->bar_ea32, n);
    if (e1 == hipSuccess) e1 = hipMemcpy(b->bar_ea, be.data(), sizeof(ElemArgs) * n, hipMemcpyHostToDevice);
    if (e2 == hipSuccess) e2 = hipMemcpy(b->bar_ea32, be32.data(), sizeof(ElemArgs32) * n, hipMemcpyHostToDevice);
    if (e1 != hipSuccess || e2 != hipSuccess)
    {
      g_berr = std::string("HIP error ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2) + " in dnlg_batch_create";
      bfree(b);
      delete b;
      return DNLG_E_HIP;
    }
    A.bar_ea = b->bar_ea;
    A.bar_ea32 = b->bar_ea32;
  }
  *out = b;
  return DNLG_OK;
}

extern "C" int dnlg_batch_destroy(dnlg_batch *b)
{
  if (!b) return DNLG_E_ARG;
  (void)hipStreamSynchronize(b->stream);
  bfree(b);
  delete b;
  return DNLG_OK;
}

static void issue_stepped(dnlg_batch *b, int preamble)
{
  const bool f64 = b->form == DNLG_FORM_FP64;
  if (f64) launch_bstep_elem_f64(b->args, b->blk_e, b->nblk_e, preamble, b->tile, b->stream);
  else launch_bstep_elem_f32(b->args, b->blk_e, b->nblk_e, preamble, b->tile, b->stream);
  launch_bstep_dt(b->args, b->nb, preamble, b->stream);
  if (!preamble) launch_bstep_nodal(b->args, b->blk_n, b->nblk_n, !f64, b->tile, b->stream);
}

// upload the block tables of the bars still running (largest bars first) and
// re-capture the step graph when the tables shrank by a quarter or more
static int compact_tables(dnlg_batch *b, const std::vector<StepState> &st, bool force)
{
  std::vector<int2> be, bn;
  for (int k = 0; k < b->nb; k++)
  {
    const int i = b->order[k];
    if (!st[i].running || st[i].status) continue;
    be.insert(be.end(), b->bar_be[b->tile][i].begin(), b->bar_be[b->tile][i].end());
    bn.insert(bn.end(), b->bar_bn[b->tile][i].begin(), b->bar_bn[b->tile][i].end());
  }
  if (!force && 4 * (long long)be.size() > 3 * (long long)b->graph_nblk_e) return DNLG_OK;
  b->nblk_e = (int)be.size();
  b->nblk_n = (int)bn.size();
  if (be.empty()) return DNLG_OK;
  BHIP(hipMemcpyAsync(b->blk_e, be.data(), sizeof(int2) * be.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(hipMemcpyAsync(b->blk_n, bn.data(), sizeof(int2) * bn.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
  if (b->graph)
  {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->graph_nblk_e = b->nblk_e;
  b->graph_nblk_n = b->nblk_n;
  if (b->engine == DNLG_BATCH_STEPPED_PLAIN) return DNLG_OK;
  hipGraph_t g;
  BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kGraphSteps; k++) issue_stepped(b, 0);
  BHIP(hipStreamEndCapture(b->stream, &g));
  const hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BHIP(e);
  return DNLG_OK;
}

// the stepped engine: the preamble, then increments in chunks (hipGraph
// replays of kGraphSteps increments) until no bar runs or max_steps are issued
static int solve_stepped(dnlg_batch *b, long long max_steps, std::vector<StepState> &st, long long *issued)
{
  int rc = compact_tables(b, st, true);
  if (rc) return rc;
  *issued = 0;
  if (b->nblk_e == 0) return DNLG_OK;
  issue_stepped(b, 1);
  BHIP(hipGetLastError());
  long long done = 0, chunk = 32;
  while (done < max_steps)
  {
    long long n = std::min(chunk, max_steps - done);
    if (done > 0)
    {
      // no more increments than the running bars need at their present time
      // step (whole graph replays); a bar whose step shrinks gets another chunk
      long long need = 0;
      for (const StepState &s : st)
        if (s.running && !s.status && s.dt > 0.0)
        {
          const double lim = s.endTime < s.upTime ? s.endTime : s.upTime;
          need = std::max(need, (long long)std::ceil((lim - s.t) / s.dt) + 1);
        }
      need = (need + kGraphSteps - 1) / kGraphSteps * kGraphSteps;
      n = std::min(n, std::max<long long>(need, kGraphSteps));
    }
    long long k = 0;
    if (b->graph)
      for (; k + kGraphSteps <= n; k += kGraphSteps) BHIP(hipGraphLaunch(b->graph, b->stream));
    for (; k < n; k++) issue_stepped(b, 0);
    BHIP(hipGetLastError());
    done += n;
    BHIP(hipMemcpyAsync(st.data(), b->st, sizeof(StepState) * b->nb, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    bool any = false;
    for (const StepState &s : st) any = any || (s.running && !s.status);
    if (!any) break;
    rc = compact_tables(b, st, false);
    if (rc) return rc;
    chunk = std::min<long long>(chunk * 2, 512);
  }
  *issued = done;
  return DNLG_OK;
}

extern "C" int dnlg_batch_set_engine(dnlg_batch *b, int engine)
{
  if (!b || engine < DNLG_BATCH_AUTO || engine > DNLG_BATCH_STEPPED_WIDE) return DNLG_E_ARG;
  if (b->graph)
  {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->engine = engine;
  return DNLG_OK;
}

// one solve(up_time) of every domain with the given engine and tile
static int run_solve(dnlg_batch *b, double up_time, long long max_steps, int engine, double *wall_us,
                     double *kernel_ms)
{
  b->tile = engine == DNLG_BATCH_STEPPED_WIDE ? 1 : 0;
  std::vector<StepState> st(b->nb);
  BHIP(hipMemcpy(st.data(), b->st, sizeof(StepState) * b->nb, hipMemcpyDeviceToHost));
  for (auto &s : st)
  {
    s.upTime = up_time;
    const double lim = s.endTime < up_time ? s.endTime : up_time;
    s.running = (s.t >= 0.0) && (s.t < lim);
  }
  BHIP(hipMemcpy(b->st, st.data(), sizeof(StepState) * b->nb, hipMemcpyHostToDevice));
  if (!b->mass_done)
  {
    // Domain::computeMassMatrix of every domain, once (FE/Domain.C:948-1005)
    double hn[4] = {0, 0, 0, 0};
    launch_mass_f64(2, b->args.ea, b->rho, hn, b->me, b->stream);
    launch_gather_scalar(b->deg, b->me, b->mass, b->NN, b->stream);
    BHIP(hipGetLastError());
    b->mass_done = true;
  }
  b->args.max_steps = max_steps;
  BHIP(hipMemsetAsync(b->t0, 0, sizeof(unsigned long long) * b->nb, b->stream));
  BHIP(hipMemsetAsync(b->t1, 0, sizeof(unsigned long long) * b->nb, b->stream));
  // the events, and the engine code that compact_tables reads (no graph for
  // stepped-plain), are put back on every way out
  struct Scope
  {
    dnlg_batch *b;
    int engine;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scope()
    {
      b->engine = engine;
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } sc{b, b->engine};
  BHIP(hipEventCreate(&sc.e0));
  BHIP(hipEventCreate(&sc.e1));
  BHIP(hipEventRecord(sc.e0, b->stream));
  long long issued = 0;
  const bool stepped = engine != DNLG_BATCH_PERSISTENT;
  b->engine = engine;
  if (!stepped)
  {
    if (b->form == DNLG_FORM_FP64) launch_batch_f64(b->args, b->nb, b->stream);
    else launch_batch_f32(b->args, b->nb, b->stream);
    BHIP(hipGetLastError());
  }
  else
  {
    const int rc = solve_stepped(b, max_steps, st, &issued);
    if (rc) return rc;
  }
  BHIP(hipEventRecord(sc.e1, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  BHIP(hipEventElapsedTime(&ms, sc.e0, sc.e1));
  if (kernel_ms) *kernel_ms = ms;
  if (wall_us)
  {
    std::vector<unsigned long long> t0(b->nb), t1(b->nb);
    BHIP(hipMemcpy(t0.data(), b->t0, sizeof(unsigned long long) * b->nb, hipMemcpyDeviceToHost));
    BHIP(hipMemcpy(t1.data(), b->t1, sizeof(unsigned long long) * b->nb, hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    BHIP(hipGetDevice(&dev));
    BHIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    for (int i = 0; i < b->nb; i++) wall_us[i] = (khz > 0 && t1[i] > t0[i]) ? (double)(t1[i] - t0[i]) * 1e3 / (double)khz : 0.0;
  }
  BHIP(hipMemcpy(st.data(), b->st, sizeof(StepState) * b->nb, hipMemcpyDeviceToHost));
  if (stepped)
    for (int i = 0; i < b->nb; i++)
      if (!st[i].status && st[i].running && issued >= max_steps) st[i].status = ST_BUDGET;
  for (int i = 0; i < b->nb; i++)
    if (st[i].status)
    {
      const int s = st[i].status;
      const char *what = s == ST_POLAR ? "Eigen Value Computation: No convergence (polar decomposition)"
                         : s == ST_RR  ? "Element::computeInternalForces: Too many atempts in plastic procedure"
                         : s == ST_DETJ ? "Negative or singular Jacobian"
                         : s == ST_BUDGET ? "increment budget (max_steps) exhausted before upTime"
                                          : "device error";
      b->err = "domain " + std::to_string(i) + ": " + what;
      if (s != ST_BUDGET) b->err += " (element " + std::to_string(st[i].err_elem - b->bars[i].elem0) + ")";
      return s == ST_POLAR ? DNLG_E_POLAR_NOCONV : s == ST_RR ? DNLG_E_RR_NOCONV : s == ST_DETJ ? DNLG_E_NEG_DETJ
                                                                                            : DNLG_E_STATE;
    }
  return DNLG_OK;
}


// state a solve changes, copied aside for the autotuner's probes
struct BatchSnapshot
{
  std::vector<std::pair<void *, size_t>> bufs;  // (device buffer, bytes)
  std::vector<void *> copies;
};

static int snapshot_take(dnlg_batch *b, BatchSnapshot &sn)
{
  const long long NN = b->NN, NIP = b->NE * 4;
  const bool f64 = b->form == DNLG_FORM_FP64;
  sn.bufs = {{b->x, sizeof(double) * NN * 3}, {b->v, sizeof(double) * NN * 3}, {b->a, sizeof(double) * NN * 3},
             {b->du, sizeof(double) * NN * 3}, {b->d, sizeof(double) * NN * 3},
             {b->st, sizeof(StepState) * b->nb}, {b->crit, sizeof(unsigned long long) * b->nb},
             {b->t0, sizeof(unsigned long long) * b->nb}, {b->t1, sizeof(unsigned long long) * b->nb}};
  if (f64)
  {
    sn.bufs.push_back({b->S, sizeof(double) * NIP * 9});
    sn.bufs.push_back({b->evp, sizeof(double) * NIP});
    sn.bufs.push_back({b->p, sizeof(double) * b->NE});
  }
  else
  {
    sn.bufs.push_back({b->S32, sizeof(float) * NIP * 6});
    sn.bufs.push_back({b->evp32, sizeof(float) * NIP});
    sn.bufs.push_back({b->p32, sizeof(float) * b->NE});
  }
  for (auto &q : sn.bufs)
  {
    void *c = nullptr;
    BHIP(hipMalloc(&c, q.second));
    sn.copies.push_back(c);
    BHIP(hipMemcpyAsync(c, q.first, q.second, hipMemcpyDeviceToDevice, b->stream));
  }
  BHIP(hipStreamSynchronize(b->stream));
  return DNLG_OK;
}

static int snapshot_restore(dnlg_batch *b, const BatchSnapshot &sn)
{
  for (size_t k = 0; k < sn.bufs.size(); k++)
    BHIP(hipMemcpyAsync(sn.bufs[k].first, sn.copies[k], sn.bufs[k].second, hipMemcpyDeviceToDevice, b->stream));
  BHIP(hipStreamSynchronize(b->stream));
  return DNLG_OK;
}

static void snapshot_free(BatchSnapshot &sn)
{
  for (void *c : sn.copies) (void)hipFree(c);
  sn.copies.clear();
}

// The autotuner behind DNLG_BATCH_AUTO: at the first solve, each candidate
// (stepped engine with 64-lane blocks, stepped with 256-lane blocks,
// persistent) runs the first kProbeSteps increments from a snapshot of the
// state, which is restored after every probe; the candidate with the least
// device time is kept for this and later solves.  The candidates give the
// same results bit for bit (a block never straddles two bars and every bar's
// arithmetic and order are fixed), so the choice changes only the time.
static constexpr long long kProbeSteps = 48;

static int autotune(dnlg_batch *b, double up_time)
{
  BatchSnapshot sn;
  int rc = snapshot_take(b, sn);
  if (rc) return rc;
  const int cand_engine[3] = {DNLG_BATCH_STEPPED, DNLG_BATCH_STEPPED_WIDE, DNLG_BATCH_PERSISTENT};
  int best = 0;
  bool failed = false;
  for (int c = 0; c < 3 && !failed; c++)
  {
    double ms = 0.0;
    const int prc = run_solve(b, up_time, kProbeSteps, cand_engine[c], nullptr, &ms);
    // a probe ends on its increment budget: that status is part of the probe
    // only.  Another error (an element error within the first increments)
    // ends the tuning; the solve then meets it with the stepped engine and
    // reports it as an untuned solve would.
    failed = prc && prc != DNLG_E_STATE;
    rc = snapshot_restore(b, sn);
    if (rc) break;
    b->probe_ms[c] = ms;
    if (!failed && ms < b->probe_ms[best]) best = c;
  }
  snapshot_free(sn);
  if (rc) return rc;
  b->auto_engine = failed ? DNLG_BATCH_STEPPED : cand_engine[best];
  b->tuned = true;
  b->err.clear();
  return DNLG_OK;
}

extern "C" int dnlg_batch_solve(dnlg_batch *b, double up_time, long long max_steps, double *wall_us, double *kernel_ms)
{
  if (!b || max_steps <= 0) return DNLG_E_ARG;
  if (b->engine == DNLG_BATCH_AUTO && !b->tuned)
  {
    const int rc = autotune(b, up_time);
    if (rc) return rc;
  }
  return run_solve(b, up_time, max_steps, b->engine == DNLG_BATCH_AUTO ? b->auto_engine : b->engine, wall_us,
                   kernel_ms);
}

extern "C" int dnlg_batch_get_tuning(dnlg_batch *b, int *engine, double *probe_ms)
{
  if (!b) return DNLG_E_ARG;
  if (!b->tuned)
  {
    b->err = "dnlg_batch_get_tuning: the autotuner has not run (engine DNLG_BATCH_AUTO, before the first solve)";
    return DNLG_E_STATE;
  }
  if (engine) *engine = b->auto_engine;
  if (probe_ms)
    for (int c = 0; c < 3; c++) probe_ms[c] = b->probe_ms[c];
  return DNLG_OK;
}

extern "C" int dnlg_batch_get_time(dnlg_batch *b, int i, double *out)
{
  if (!b || i < 0 || i >= b->nb || !out) return DNLG_E_ARG;
  StepState s;
  BHIP(hipMemcpy(&s, b->st + i, sizeof s, hipMemcpyDeviceToHost));
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

extern "C" int dnlg_batch_get_field(dnlg_batch *b, int i, int field, double *out, long long n)
{
  if (!b || i < 0 || i >= b->nb || !out) return DNLG_E_ARG;
  const BarDesc &d = b->bars[i];
  const double *src = nullptr;
  const float *src32 = nullptr;
  long long off = 0, cnt = 0, stride = 0, comps = 1;
  switch (field)
  {
  case DNLG_F_COORDS: src = b->x; off = d.node0 * 3; cnt = d.nn * 3; break;
  case DNLG_F_V: src = b->v; off = d.node0 * 3; cnt = d.nn * 3; break;
  case DNLG_F_A: src = b->a; off = d.node0 * 3; cnt = d.nn * 3; break;
  case DNLG_F_DU: src = b->du; off = d.node0 * 3; cnt = d.nn * 3; break;
  case DNLG_F_D: src = b->d; off = d.node0 * 3; cnt = d.nn * 3; break;
  case DNLG_F_MASS: src = b->mass; off = d.node0; cnt = d.nn; break;
  case DNLG_F_EVP:
    if (b->form == DNLG_FORM_FP64) { src = b->evp; off = d.elem0 * 4; cnt = d.ne * 4; }
    else { src32 = b->evp32; off = d.elem0; cnt = d.ne; stride = b->NE; comps = 4; }  // [nip][NE]
    break;
  case DNLG_F_P:
    if (b->form == DNLG_FORM_FP64) src = b->p; else src32 = b->p32;
    off = d.elem0; cnt = d.ne;
    break;
  case DNLG_F_S:
    if (b->form == DNLG_FORM_FP64) { src = b->S; off = d.elem0 * 4; cnt = d.ne * 4; stride = b->NE * 4; comps = 9; }
    else { src32 = b->S32; off = d.elem0; cnt = d.ne; stride = b->NE; comps = 24; }  // [6][nip][NE]
    break;
  default: b->err = "dnlg_batch_get_field: unknown field"; return DNLG_E_ARG;
  }
  if (stride == 0) stride = cnt;
  if (n != cnt * comps)
  {
    b->err = "dnlg_batch_get_field: size mismatch";
    return DNLG_E_ARG;
  }
  for (long long c = 0; c < comps; c++)
  {
    if (src)
      BHIP(hipMemcpy(out + c * cnt, src + c * stride + off, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    else
    {
      std::vector<float> tmp(cnt);
      BHIP(hipMemcpy(tmp.data(), src32 + c * stride + off, sizeof(float) * cnt, hipMemcpyDeviceToHost));
      for (long long k = 0; k < cnt; k++) out[c * cnt + k] = (double)tmp[k];
    }
  }
  return DNLG_OK;
}
