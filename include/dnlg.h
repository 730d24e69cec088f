/*
 * dnlg.h -- C ABI of libdnlgpu, the MI355X (gfx950) explicit FEM core.
 *
 * It replaces, for one domain on one GPU, the hot path of pantale/ReDynELA
 * (DynELA v1.0):
 *   - Domain::computeInternalForces()          FE/Domain.h:103, FE/Domain.C:130-738
 *   - Element::computeInternalForces(Vector&)  FE/Element.h:251, FE/Element.C:164-403
 *   - Solver::solve(Real upTime) as ExplicitSolverCH::solve
 *                                              FE/Solver.h:126, FE/ExplicitSolverCH.C:115-278
 * (FE = Sources/Libraries/finiteElement.)  All calls take plain pointers and
 * sizes; host buffers belong to the caller, device memory to the handle.
 * Calls on one handle are serialized by the caller; different handles are
 * independent.  Every function returns DNLG_OK (0) or a DNLG_E_* code, and
 * dnlg_last_error() gives the message.
 */
#ifndef DNLG_H
#define DNLG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNLG_ABI_VERSION 1

/* status codes; the numerical ones mirror the reference's fatalError sites */
#define DNLG_OK 0
#define DNLG_E_ARG 1            /* bad argument (the reference: fatalError on bad input)          */
#define DNLG_E_POLAR_NOCONV 2   /* Jacobi polar decomposition > 29 rotations, LA/Tensor2.C:1257    */
#define DNLG_E_RR_NOCONV 3      /* radial return > 2000 iterations, FE/Element.C:327-330           */
#define DNLG_E_NEG_DETJ 4       /* negative (2D) or singular (3D) Jacobian, FE/Element2D.C:101-109 */
#define DNLG_E_HIP 5            /* HIP runtime error                                               */
#define DNLG_E_STATE 6          /* call out of sequence                                            */
#define DNLG_E_PULSE_NOCONV 7   /* power iteration > 1000 iterations, FE/Domain.C:936-939          */
/* Element errors (codes 2-4) are fatal in the reference (BT/Error.C:141-160).
 * Here the handle stops at the element pass that finds one: no nodal update,
 * endStep or later increment runs, and every later call reports the error
 * again.  The failing increment counts; a 2D negative Jacobian, which the
 * reference finds at the end of the previous increment, therefore leaves the
 * increment count one above the reference's, with time and nodal state
 * equal. */

/* element families: the reference's ElQua4nAx (FE/ElQua4nAx.C) and ElHex8n3D (FE/ElHex8n3D.C) */
#define DNLG_DIM_AXI 2
#define DNLG_DIM_3D 3

/* mesh: node coordinates (3 per node, z = 0 for axisymmetric) and element
 * connectivity in the reference element's local node order, 0-based node
 * indices = the reference's Node::Id after Domain::initSolve (FE/Domain.C:1159) */
typedef struct
{
  int dim;               /* DNLG_DIM_AXI (4-node quads) or DNLG_DIM_3D (8-node hexahedra) */
  long long n_nodes;
  long long n_elems;
  const double *coords;  /* [n_nodes][3] */
  const int *conn;       /* [n_elems][4 or 8] */
} dnlg_mesh;

/* Material + IsoHardElastoplastic / IsoHardJohnsonCook (FE/Material.h,
 * FE/IsoHardElastoplastic.C:171-184, FE/IsoHardJohnsonCook.C:185-198):
 *   hardening = 0: purely elastic (Material::isYield() == No)
 *   hardening = 1: sigma_y = A + B evp^n, the law both reference classes evaluate
 *   hardening = 2: Johnson-Cook with the rate and temperature factors that
 *                  FE/IsoHardJohnsonCook.h:51-68 documents but the reference
 *                  never evaluates: (A + B evp^n)[1 + C ln(epsdot/epsdot0)]
 *                  [1 - ((T-T0)/(Tm-T0))^m], adiabatic heating
 *                  dT = taylor_quinney sigma_y depsp / (density cp).
 *                  With C = 0 and taylor_quinney = 0 it is hardening = 1 exactly. */
typedef struct
{
  double young, poisson, density;
  double A, B, n;
  int hardening;
  double C, epsd0, m, T0, Tm, cp, taylor_quinney;
} dnlg_material;

/* boundary conditions: BoundaryRestrain (FE/Boundary.C:636-666) as a per-node
 * component mask (bit c = component c held), initial BoundarySpeed
 * (FE/Boundary.C:378-384) as a per-node velocity, and an optional constant
 * BoundarySpeed (FE/Boundary.C:478-524): mask + prescribed velocity.  The
 * reference drives only a speed's nonzero components, so bit c of speed_mask
 * is set where speed[c] != 0.  Speeds are applied before restraints: a
 * component both driven and held ends held, as when the reference attaches
 * the restraint last. */
typedef struct
{
  const unsigned char *restrain;   /* [n_nodes] or NULL */
  const double *v0;                /* [n_nodes][3] or NULL */
  const unsigned char *speed_mask; /* [n_nodes] or NULL */
  const double *speed;             /* [n_nodes][3] or NULL */
} dnlg_bc;

/* ExplicitSolverCH parameters (FE/ExplicitSolverCH.C:63-84, 336-343,
 * FE/Solver.C:45-76): spectral radius rho_b (1.0 = central difference),
 * time-step safety factor, Courant recompute frequency, end time. */
typedef struct
{
  double rho_b;
  double safety;
  int ts_freq;
  double end_time;
  int form;          /* DNLG_FORM_FP64 or DNLG_FORM_MIXED */
} dnlg_solver_params;

/* arithmetic forms of the element pass */
#define DNLG_FORM_FP64 0   /* the reference's arithmetic, operation for operation */
#define DNLG_FORM_MIXED 1  /* element kinematics, stress update and B^T sigma in FP32; nodal state and time in FP64 */

typedef struct dnlg_handle dnlg_handle;

/* state fields for dnlg_get_field / dnlg_set_field */
#define DNLG_F_COORDS 0   /* [n_nodes][3] Node::coords                              */
#define DNLG_F_V 1        /* [n_nodes][3] Current->mat_speed                        */
#define DNLG_F_A 2        /* [n_nodes][3] Current->dmat_speed                       */
#define DNLG_F_DU 3       /* [n_nodes][3] Current->delta_disp                       */
#define DNLG_F_D 4        /* [n_nodes][3] Current->disp                             */
#define DNLG_F_FINT 5     /* [n_nodes][dim] Domain::Fint, node-major                */
#define DNLG_F_MASS 6     /* [n_nodes] Node::mass                                   */
#define DNLG_F_S 7        /* [9][n_elems*nip] IntegrationPoint::S, point-major      */
#define DNLG_F_EVP 8      /* [n_elems*nip] IntegrationPoint::evp                    */
#define DNLG_F_P 9        /* [n_elems] IntegrationPoint::p (equal at all points)    */
#define DNLG_F_EPS 10     /* [9][n_elems*nip] IntegrationPoint::Eps (if tracked)    */
#define DNLG_F_EPSPLAS 11 /* [9][n_elems*nip] IntegrationPoint::EpsPlas (if tracked)*/
#define DNLG_F_FE 12      /* [n_elems*nen][dim] element nodal forces of the last pass */
#define DNLG_F_TEMP 13    /* [n_elems*nip] integration-point temperature (hardening = 2)  */
#define DNLG_F_DEPS 14    /* [9][n_elems*nip] IntegrationPoint::dEps of the last pass (if tracked) */
/* In DNLG_FORM_MIXED, DNLG_F_S is [6][nip][n_elems] (xx,yy,zz,xy,yz,xz) and
 * DNLG_F_EVP / DNLG_F_TEMP are [nip][n_elems]; values are widened to double. */

/* options for dnlg_set_option */
#define DNLG_OPT_KINETIC_ENERGY 1 /* compute Domain::getTotalKineticEnergy every step (default 0) */
#define DNLG_OPT_STORE_FINT 2     /* keep Domain::Fint in device memory every step (default 1)    */
#define DNLG_OPT_TRACK_DISP 3     /* keep NodalField::disp (default 1)                            */
#define DNLG_OPT_TRACK_STRAIN 4   /* keep IntegrationPoint Eps, EpsPlas; set before the first step (default 0) */
#define DNLG_OPT_TIME_STEP_METHOD 5 /* Solver::setTimeStepMethod (FE/Solver.C:312-327): DNLG_TS_COURANT
                                       (the default of this ABI) or DNLG_TS_PULSE (the default of the
                                       reference's ExplicitSolverCH, FE/ExplicitSolverCH.C:83)       */
#define DNLG_TS_COURANT 0
#define DNLG_TS_PULSE 1
#define DNLG_OPT_PLOT_HISTORY 6   /* keep the endStep history line of every increment (the .plot file,
                                     FE/ExplicitSolverCH.C:459-461; turns kinetic energy on); read with
                                     dnlg_plot_rows.  The device ring holds 8192 lines: dnlg_step_async
                                     refuses to run further ahead of the last synchronisation */

int dnlg_abi_version(void);

/* Structure/Domain set-up (Domain::initSolve, FE/Domain.C:1018-1226): copies
 * the mesh, material, BCs and solver parameters into device memory. */
int dnlg_create(const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bc *bc,
                const dnlg_solver_params *solver, dnlg_handle **out);
int dnlg_destroy(dnlg_handle *h);
int dnlg_set_option(dnlg_handle *h, int option, int value);

/* Domain::computeInternalForces(): stress update of every integration point
 * with the Current displacement increment and assembly of Domain::Fint.
 * fint_out ([n_nodes][dim], node-major as the reference's Fint) may be NULL. */
int dnlg_compute_internal_forces(dnlg_handle *h, double *fint_out);

/* Domain::computeInternalForces() as ExplicitSolverCH::initIteration calls it
 * inside the reference's own time loop (FE/ExplicitSolverCH.C:471-481, after
 * a host computePredictions): the caller's node coordinates (Node::coords)
 * and Current displacement increment (Current->delta_disp), [n_nodes][3]
 * host arrays, replace the handle's device copies before the pass (NULL keeps
 * the device copy).  The integration-point state stays in device memory, as
 * it stays in the reference's IntegrationPoints between calls.  fint_out as
 * above. */
int dnlg_compute_internal_forces_from(dnlg_handle *h, const double *coords, const double *delta_disp,
                                      double *fint_out);

/* Domain::computeCourantTimeStep() (FE/Domain.h:128, FE/Domain.C:1613-1646)
 * on the handle's present coordinates: crit = min over elements of
 * getLength() / getSoundSpeed(), the value Solver::computeTimeStep turns into
 * fmax = 2 / crit (FE/Solver.C:228-231).  Hexahedra use the build-defined
 * length V / A_max (the reference's ElHex8n3D::getLength exits). */
int dnlg_courant_time_step(dnlg_handle *h, double *crit);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the crit of
 * dnlg_courant_time_step, bit for bit, for a mesh in caller-owned DEVICE
 * memory -- coords [n_elems' nodes][3] doubles, conn [n_elems][4 or 8] ints,
 * crit one double written on the device -- with no handle.  Only young,
 * poisson and density of *mat (a host struct) are read.  The kernels run on
 * `stream` (a hipStream_t of the current device, NULL for the null stream)
 * and the call returns without waiting.  It allocates nothing: the per-block
 * minima go to a 512 KiB array that each GPU holds for the life of the
 * process, so n_elems is at most 16,777,216 (DNLG_E_ARG above that).  Calls
 * that overlap on one GPU share that array and race.  Errors:
 * dnlg_noalloc_last_error(). */
int dnlg_courant_time_step_noalloc(int dim, long long n_elems, const double *coords, const int *conn,
                                   const dnlg_material *mat, double *crit, void *stream);
const char *dnlg_noalloc_last_error(void);

/* ExplicitSolverCH::solve(upTime) (FE/ExplicitSolverCH.C:115-278) in parts:
 * dnlg_begin runs the solve() preamble (mass matrix once, internal forces,
 * forced time step); dnlg_step runs up to nsteps passes of the time loop and
 * reports whether the loop continues; dnlg_solve does the whole call. */
int dnlg_begin(dnlg_handle *h, double up_time, int *running);
int dnlg_step(dnlg_handle *h, int nsteps, int *running);
int dnlg_solve(dnlg_handle *h, double up_time, long long *steps_done);

/* issue nsteps passes of the time loop on the handle's stream without
 * waiting (no host synchronisation); dnlg_synchronize waits and checks errors */
int dnlg_step_async(dnlg_handle *h, int nsteps);
int dnlg_synchronize(dnlg_handle *h);

/* state exchange with caller-owned host buffers (n = number of doubles) */
int dnlg_get_field(dnlg_handle *h, int field, double *out, long long n);
int dnlg_set_field(dnlg_handle *h, int field, const double *in, long long n);
long long dnlg_field_size(dnlg_handle *h, int field);

/* time scalars: out[0] t, [1] dt, [2] increment, [3] computed time step,
 * [4] kinetic energy of the last step, [5] alphaM, [6] beta, [7] gamma, [8] OmegaS */
int dnlg_get_time(dnlg_handle *h, double *out);
/* Structure::readData on a restart (FE/Structure.C:915-931, driven by
 * Sources/Applications/DynELA_solve/main.C:106-130): the current time of a
 * handle that has not begun.  With dnlg_set_field of the coordinates, the
 * Current nodal fields and the point state read from a result file, the next
 * dnlg_begin continues as a restarted reference run does: mass from the
 * restored coordinates, a fresh time-step history, increment count 0. */
int dnlg_set_time(dnlg_handle *h, double t);

/* Pulse time step (Domain::getMaximumFrequency, FE/Domain.C:816-946; no
 * accessor in the reference): out[0] calls so far, [1] iterations of the
 * last call, [2] fmax0 (the last fmax, Domain::fmax0), [3] its convergence
 * measure.  A step that recomputes the Pulse time step waits on the host
 * for the power iteration, also under dnlg_step_async. */
int dnlg_pulse_info(dnlg_handle *h, double *out4);

/* endStep history lines (DNLG_OPT_PLOT_HISTORY): up to max_rows lines of
 * 4 doubles (time, time step, computed time step, kinetic energy), oldest
 * first, are moved into out; returns their number (the number waiting when
 * out is NULL), or -DNLG_E_ARG.  FE/ExplicitSolverCH.C:459-461 writes the
 * same four values with "%8.4E %8.4E %8.4E %8.4E\n" (no accessor in the reference). */
long long dnlg_plot_rows(dnlg_handle *h, double *out, long long max_rows);

/* HistoryFile nodal series (FE/HistoryFile.C:114-141; Structure::addHistoryFile,
 * FE/Structure.C:601-613): at each endStep whose time tm reaches the file's
 * next save time (from `start`, every `every`, up to stop + every when
 * stop > 0) a row (tm, value of each item) is kept.  Items are (node index,
 * field, component) with Node::getNodalValue's fields (FE/Node.C:326-418) as
 * DNLG_HF_* codes and its components: 0 the norm of a vector, 1..3 one
 * component (element fields: below); vector fields are read from the Current nodal field as the
 * reference does at endStep (the previous increment's values), coordinates
 * after the update.  init_temp is the material's initial temperature
 * (Material::setInitTemperature) for the temperature fields.  Rows are read
 * with dnlg_history_file_rows (1 + n_items doubles each); the rings hold
 * 8192 rows between synchronisations, as the .plot ring. */
#define DNLG_HF_DENSITY 0
#define DNLG_HF_DENSITY_INC 1
#define DNLG_HF_ENERGY 2
#define DNLG_HF_ENERGY_INC 3
#define DNLG_HF_SPEED 4
#define DNLG_HF_SPEED_INC 5
#define DNLG_HF_FORCE 6
#define DNLG_HF_TEMPERATURE 7
#define DNLG_HF_DISPLACEMENT_INC 8
#define DNLG_HF_DISPLACEMENT 9
#define DNLG_HF_COORDS 10
#define DNLG_HF_MASS 11
#define DNLG_HF_TEMPERATURE_INIT 12
/* element fields (FE/Node.C:350-418, FP64 form): the points' tensors
 * extrapolated to the node with the element's integrationPointsToNode table
 * and averaged over the node's elements (FE/Element.C:1180-1198);
 * component 0 is the von Mises equivalent of the stress or the norm
 * sqrt(2/3 e:e) of a strain, 10 i + j (i, j in 1..3) the entry (i, j).  The
 * strain fields need the tracked point strains: on a handle that has not
 * begun they turn DNLG_OPT_TRACK_STRAIN on. */
#define DNLG_HF_STRESS 13
#define DNLG_HF_STRAIN 14
#define DNLG_HF_STRAIN_INC 15
#define DNLG_HF_PLASTIC_STRAIN 16
int dnlg_history_file_add(dnlg_handle *h, double start, double stop, double every, int n_items,
                          const long long *nodes, const int *fields, const int *comps, double init_temp, int *id);
long long dnlg_history_file_rows(dnlg_handle *h, int id, double *out, long long max_rows);

/* HIP stream of the handle (hipStream_t), for callers that order work on it */
void *dnlg_get_stream(dnlg_handle *h);

/* kernel timing: when enabled, HIP events bracket every element-kernel launch
 * of dnlg_step/dnlg_step_async on the handle's stream; dnlg_kernel_stats
 * returns the launch count and the summed milliseconds of the element kernel
 * (which = 0) and of the nodal kernel (which = 1) since the last reset. */
int dnlg_enable_kernel_timing(dnlg_handle *h, int on);
int dnlg_kernel_stats(dnlg_handle *h, int which, long long *launches, double *total_ms);

const char *dnlg_last_error(dnlg_handle *h);

/* ---- mesh generators (host only; no GPU call) ----
 * Axisymmetric Taylor bar, the recipe of Samples/Taylor/Taylor_Axi_Parametric.dyn:
 * (nr+1)(nz+1) nodes at (i R/nr, j L/nz), nr*nz quads, base (j = 0) restrained
 * along the axis direction (bit 1), axis nodes (i = 0) restrained along r and z
 * (bits 0, 2), initial velocity (0, -v0, 0). */
int dnlg_mesh_axi_bar(int nr, int nz, double radius, double length, double v0, double *coords, int *conn,
                      unsigned char *restrain, double *vel);
/* 3D O-grid cylinder: a central n x n block of half-width s_frac*radius and four
 * ring blocks of n x m elements, nz layers along z; (n+1)^2 + 4 n m nodes and
 * n^2 + 4 n m elements per layer.  Nodes on z = 0 restrained along z (bit 2),
 * initial velocity (0, 0, -v0). */
int dnlg_mesh_ogrid_sizes(int n, int m, int nz, long long *n_nodes, long long *n_elems);
int dnlg_mesh_ogrid_bar(int n, int m, int nz, double radius, double length, double s_frac, double v0,
                        double *coords, int *conn, unsigned char *restrain, double *vel);

/* ---- batched mode: many independent axisymmetric domains, one launch ----
 * Each domain is the whole of one ExplicitSolverCH::solve(upTime)
 * (FE/ExplicitSolverCH.C:115-278) with its own mesh, hardening parameters,
 * initial velocity, time step and increment count; the reference runs one
 * Structure per process (FE/Structure.C:809-897), so a caller replaces a loop
 * of those runs with one dnlg_batch_solve.  Every domain is a quad4
 * axisymmetric mesh (dim 2); E, nu and density are shared; hardening 0 or 1;
 * no speed BCs.  Domain i of a batch follows, bit for bit in the FP64 form,
 * the run of dnlg_create + dnlg_solve on the same inputs. */
typedef struct dnlg_batch dnlg_batch;
int dnlg_batch_create(int n_domains, const dnlg_mesh *meshes, const dnlg_material *materials, const dnlg_bc *bcs,
                      const dnlg_solver_params *solver, dnlg_batch **out);
int dnlg_batch_destroy(dnlg_batch *b);
/* engines: PERSISTENT = one workgroup carries one domain through all of its
 * increments in a single launch; STEPPED = one increment of every running
 * domain per group of launches (element pass, per-domain time step, nodal
 * phase) in 64-lane blocks, replayed from a captured hipGraph; STEPPED_WIDE =
 * the same with 256-lane blocks; AUTO (default) = at the first solve, each of
 * STEPPED, STEPPED_WIDE and PERSISTENT runs the first 48 increments from a
 * copy of the state, the state is put back, and the fastest is kept for this
 * and later solves of the batch.  Results are bit-identical across engines. */
#define DNLG_BATCH_AUTO 0
#define DNLG_BATCH_PERSISTENT 1
#define DNLG_BATCH_STEPPED 2
#define DNLG_BATCH_STEPPED_PLAIN 3 /* STEPPED launched increment by increment, without graphs (for tracing tools) */
#define DNLG_BATCH_STEPPED_WIDE 4
int dnlg_batch_set_engine(dnlg_batch *b, int engine);
/* No counterpart in the reference (its runs are one Structure per process):
 * the AUTO engine's choice (one of the codes above) and the device time in ms
 * of its three probes, in the order STEPPED, STEPPED_WIDE, PERSISTENT.
 * DNLG_E_STATE before the first solve of an AUTO batch or for another engine. */
int dnlg_batch_get_tuning(dnlg_batch *b, int *engine, double *probe_ms3);
/* runs every domain to up_time (or its end_time); at most max_steps increments
 * per domain (DNLG_E_STATE if a domain needs more).  A domain stops at its
 * first element error with the failing increment counted and its nodal state
 * and time as before that increment (the reference stops inside
 * computeInternalForces); the other domains run on, and the error of the
 * lowest-numbered failed domain is returned.  wall_us[i] (optional) gets
 * domain i's device wall time from the start of its solve to the end of its
 * last increment, kernel_ms (optional) the duration of the whole solve on the
 * device stream. */
int dnlg_batch_solve(dnlg_batch *b, double up_time, long long max_steps, double *wall_us, double *kernel_ms);
/* time state of domain i, the 9 values of dnlg_get_time */
int dnlg_batch_get_time(dnlg_batch *b, int i, double *out9);
/* a field of domain i: DNLG_F_COORDS/V/A/DU/D ([nn][3]), MASS ([nn]), P ([ne]),
 * EVP ([ne][4] in the FP64 form, [4][ne] in the mixed one), S ([9][ne*4] FP64,
 * [6][4][ne] mixed); n must equal the size */
int dnlg_batch_get_field(dnlg_batch *b, int i, int field, double *out, long long n);
const char *dnlg_batch_last_error(dnlg_batch *b);
/* the parameter draw of the batched Taylor workload: for domain i, in this
 * order, nr ~ U{5..20}, v0 ~ U(150, 300) m/s, sigma_y0 ~ U(300, 500) MPa,
 * hardening modulus ~ U(50, 300) MPa, from std::mt19937_64(seed); an integer
 * draw g gives nr = 5 + g % 16 and a real draw U(a, b) = a + (b - a) (g >> 11) 2^-53 */
int dnlg_batch_taylor_params(int n, unsigned long long seed, int *nr, double *v0, double *sy0, double *hmod);

/* ---- adjoint: gradient of a function of the final coordinates w.r.t. (A, B, n) ----
 * No reference counterpart.  For an axisymmetric or hex8 FP64 handle with
 * hardening A + B evp^n, or the Johnson-Cook law with its temperature factor
 * and adiabatic heating but no rate term (law 2, C = 0), no speed BC, Courant
 * time step, that has not stepped yet:
 * dnlg_adjoint_forward runs the solve() preamble and nsteps increments
 * (the same bits as dnlg_begin + dnlg_step) and keeps the state after every
 * ckpt_every-th increment in device memory; dnlg_adjoint_backward takes the
 * adjoint xbar ([n_nodes][3]; axisymmetric: the third component ignored) of
 * the final node coordinates and returns grad = (dL/dA, dL/dB, dL/dn).  The
 * reverse sweep restores each checkpoint, recomputes its segment and
 * reverses it: the Chung-Hulbert update, the restraints, the element-order
 * assembly, the element kinematics (closed-form 2x2 polar decomposition;
 * in 3D the exact eigen-decomposition of F^T F), the radial return with its
 * elastic/plastic switch (and the point temperature), and the time step
 * (Courant arg-min element, LINEARdbl extrapolation).  With ckpt_every = 1 every increment's state is
 * kept (nsteps snapshots) and the sweep reverses over them with no
 * recomputation: memory for time, the same gradient bit for bit.  After it
 * the handle holds the post-forward state. */
typedef struct
{
  double forward_ms;      /* dnlg_adjoint_forward, device time incl. checkpoint copies  */
  double backward_ms;     /* dnlg_adjoint_backward, device time                         */
  double recompute_ms;    /* part of backward_ms spent recomputing segments             */
  long long device_bytes; /* checkpoints, segment tape and adjoint state                */
  long long checkpoint_bytes;
  int n_checkpoints;
  long long steps;
} dnlg_adjoint_stats;
int dnlg_adjoint_forward(dnlg_handle *h, long long nsteps, int ckpt_every);
int dnlg_adjoint_backward(dnlg_handle *h, const double *xbar, double *grad, dnlg_adjoint_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* DNLG_H */
