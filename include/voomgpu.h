/*
 * voomgpu.h -- C ABI of the MI355X monodomain time step (libvoomgpu.so).
 *
 * One handle = one cardiac body (mixed P1-tet + Purkinje-bar mesh, one cell
 * model family) and its Crank-Nicolson/Strang solver, resident in the HBM of
 * the current HIP device.  All calls on one handle must come from one host
 * thread; handles are independent.  Work is ordered on the handle's own HIP
 * stream (vg_get_stream).  Every function returns VG_OK or a negative error
 * code and never exits the process; vg_last_error(h) gives the message.
 *
 * Reference interfaces each entry point replaces (paths relative to
 * voomEP/src/):
 *   vg_create            CardiacBody::New (Body/CardiacBody.cc:975-1101) +
 *                        CardiacBody::Compute (:78-187) + EPSolver::New
 *                        (Solvers/EPSolver.cc:339-361) +
 *                        EP_CrankNicholson::Initialize (EP_CrankNicholson.cc:5-76)
 *                        + EPSolver::computeCapacitance (EPSolver.cc:214-219)
 *   vg_set_stimulus_*    CardiacBody::setStimulusCurrent x4 (CardiacBody.cc:243-324)
 *   vg_step              EP_CrankNicholson::Solve (EP_CrankNicholson.cc:78-156)
 *   vg_get/set_voltage   EPSolver::InitializeVoltage / _Voltage view
 *                        (EPSolver.h:187-307), EPSolver::WriteOutput payload
 *                        (EPSolver.cc:116-136)
 *   vg_get/set_states    CardiacBody::getStateVariables / setStateVariables
 *                        (CardiacBody.cc:438-463)
 *   vg_ecg_*             computeGradOneByR .. computeECG
 *                        (Applications/CardiacEP/ComputeECG.h:30-189)
 *   vg_assemble_csr      CardiacBody::getDiffusionMatrix / getCapacitanceMatrix /
 *                        computeLumpedCapacitance (CardiacBody.cc:32-67,348-391)
 *   vg_cells_*           IonicModel::Compute_Ion (IonicModel/IonicModel.h:33-34)
 *                        on independent cells
 */
#ifndef VOOMGPU_H
#define VOOMGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VG_OK 0
#define VG_ERR_ARG (-1)
#define VG_ERR_HIP (-2)
#define VG_ERR_NOT_CONVERGED (-3)
#define VG_ERR_NONFINITE (-4)
#define VG_ERR_ALLOC (-5)
#define VG_ERR_UNSUPPORTED (-6)

/* Cell models, numbered like CardiacBody::New's CellModelType switch
 * (CardiacBody.cc:988-1093). */
#define VG_MODEL_LR 0
/* Luo-Rudy with the 14 gate and background values interpolated from
 * LuoRudyGateTable's 0.001 mV table over (-90, 200) mV (CellModelType
 * LUORUDYWITHGATE; LRGateTable.h, LR.cc:71-79). */
#define VG_MODEL_LR_GATE 1
#define VG_MODEL_TT 2
/* Mahajan 2008 rabbit ventricle: each node's constants and initial state by
 * its tissue type 0-8 (vg_mesh.tissue; epi/myo/endo x apex/center/base). */
#define VG_MODEL_MJ 3
/* Mahajan with the failing-heart constant set (CellModelType MAHAJAN_FAIL,
 * Mahajan_fail.cc + SetUpMahajan_failParameters, SetUpIonicConstants.h:254-315):
 * same states and initial values, other constants. */
#define VG_MODEL_MJ_FAIL 4
/* Corrias rabbit Purkinje (IonicModel/Purkinje.cc); in a tissue the model of
 * every Purkinje node (tissue 11/12), here also selectable for vg_cells_*. */
#define VG_MODEL_PURKINJE 10

/* Arithmetic of the hot path. */
#define VG_PREC_FP64 0  /* every array and operation in FP64 */
#define VG_PREC_MIXED 1 /* FP32 states, matrices and vectors; FP64 CG dots,
                           residual norms and ECG sums (Mahajan: Ca-cycling
                           and Na_i states accumulated in FP64) */

/* Ionic integrators. */
#define VG_INTEG_REFERENCE 0   /* the reference's scheme: LR Rush-Larsen + FE;
                                  tT backward-Euler gates + Newton (J solve;
                                  VG_QUIRK_NEWTON_JT for the reference's J^T) */
#define VG_INTEG_RUSH_LARSEN 1 /* tT: Rush-Larsen gates + forward-Euler
                                  concentrations, one rate evaluation */

/* Reference behaviours kept behind flags (vg_model.quirks). */
#define VG_QUIRK_V0_ZERO 1u       /* start LR/tT at V = 0 like the driver
                                     (EPSimulator.cc:463, IonicModel.h:42) */
#define VG_QUIRK_PMJ_STICKY 2u    /* never reset the PMJ override
                                     (Mesh/MixedMesh3D.cc:14,31-34) */
#define VG_QUIRK_BAR_1X_VOLUME 4u /* integrate bars with pi r^2 L instead of
                                     the reference's 2 pi r^2 L
                                     (MixedMesh3D.cc:97-107) */
#define VG_QUIRK_NEWTON_JT 8u     /* tT reference integrator: solve each Newton
                                     step with J^T like the reference
                                     (Tusscher.cc:172-181); leaves the physical
                                     range during an upstroke at dt = 0.02 */
#define VG_QUIRK_MJ_INT_ABS 16u   /* Mahajan: read the sub-step test
                                     abs(dV/dt) > 20 (Mahajan.cc:611) through
                                     int abs(), i.e. |dV/dt| >= 21 */

/* Per-node conductance scalings of the ten Tusscher model (vg_model). */
#define VG_TT_SCALE_GNA 0
#define VG_TT_SCALE_GKR 1
#define VG_TT_SCALE_GKS 2
#define VG_TT_SCALE_GCAL 3
#define VG_TT_SCALE_GTO 4
#define VG_TT_NSCALE 5

typedef struct vg_handle vg_handle;
typedef struct vg_cells vg_cells;

/* Mesh, 0-based.  conn has 4 slots per element; bars (elem_nn 2) use the
 * first two.  fibers: 9 numbers (f, s, n) per element or NULL (isotropic).
 * tissue: per-node type (11/12 = LV/RV Purkinje) or NULL.  PMJ overrides are
 * keyed by node like the reference's .angle file: an element whose first node
 * is a PMJ node takes cosines and length from pmj_data (4 per PMJ). */
typedef struct {
  int64_t n_nodes;
  const double *xyz;
  int64_t n_elems;
  const int8_t *elem_nn;
  const int32_t *conn;
  const double *fibers;
  const int8_t *tissue;
  int64_t n_pmj;
  const int32_t *pmj_node;
  const double *pmj_data;
  double radius; /* bar radius (cm), TWODRADIUS in the reference decks */
} vg_mesh;

typedef struct {
  int cell_model;     /* VG_MODEL_* */
  int precision;      /* VG_PREC_* */
  int integrator;     /* VG_INTEG_* */
  double D[3];        /* tissue diffusivity along f, s, n (cm^2/ms) */
  double Dp[3];       /* Purkinje diffusivity */
  double DpLeft[3];   /* LV Purkinje diffusivity */
  double Cm;          /* uF/cm^2 */
  double dt;          /* diffusion step (ms) */
  double dt_ionic;    /* ionic sub-step (ms) */
  double cg_tol;      /* stop at ||r||_D <= cg_tol ||K_rhs V||_D (see VG_X0_*) */
  int cg_maxit;
  unsigned quirks;    /* VG_QUIRK_* */
  int n_param_sets;   /* tT: number of scaling sets (0 = all ones) */
  const double *param_scales;    /* [n_param_sets * VG_TT_NSCALE] */
  const uint16_t *param_index;   /* [n_nodes] set of each node, or NULL */
  int sell_sigma;     /* rows sorted by length in windows of this many rows
                         (0 or 1 = keep the mesh order) */
  int cg_blocks;      /* workgroups of each SpMV/CG launch, rounded up to a
                         multiple of 8 (0 = the library's choice) */
  int cg_x0;          /* VG_X0_*: initial guess of each half-step solve */
  /* Batching: n_instances > 1 independent bodies in one handle (no shared
   * elements, no Purkinje nodes); instance[n_nodes] gives each node's body.
   * Every body runs its own CG to its own tolerance; launches are shared. */
  int n_instances;
  const int32_t *instance;
  /* 1: after a few host-driven steps, capture the whole step (both solves
   * with a fixed number of CG launches each, learned from those steps with a
   * margin; launches past convergence return at once) in a hipGraph and
   * replay it with no host synchronisation per step.  A solve still running
   * when its captured launches run out passes its last iterate on; it is
   * counted in vg_stats.cg_truncated (and cg_not_converged), and the host,
   * which reads that count after every replay without synchronising,
   * re-captures with twice the launches before the next step. */
  int graph;
  /* Diffusive smoothing factor alpha (EPSimulator's SMOOTHING keyword,
   * EP_CrankNicholson.cc:31-36): A and B get +-alpha (4/dt) K on top of
   * +-dt/4 K.  0 = off. */
  double smoothing;
  /* Step scheme (EPSolver::New's Method, EPSolver.cc:339-361): VG_SCHEME_CN
   * (default), VG_SCHEME_BE (backward-Euler half-steps with the lumped
   * capacitance, EP_BackwardEuler.cc) or VG_SCHEME_EXPLICIT (forward-Euler
   * half-steps in strides of dt_diffusion, EP_ExplicitSolver.cc).  The BE
   * and explicit schemes advance the cells through
   * CardiacBody::getVoltageFromIonicCurrent's per-node sub-steps
   * (CardiacBody.cc:819-902). */
  int scheme;
  double dt_diffusion; /* explicit stride (ms); 0 = 0.1, Solve's default argument */
  /* Row order of the myocardium rows (VG_ORDER_*).  No reference
   * counterpart: Epetra keeps the partition's node order.  Only the storage
   * order changes; every array in and out of the library stays in mesh
   * order. */
  int row_order;
  /* Every this many steps (0 = never), recompute after each of the step's
   * two solves the true residual ||K_rhs V - A_hat x||_D / ||K_rhs V||_D in
   * FP64 arithmetic from the stored operator, and keep the largest in
   * vg_stats.max_true_relres.  The stop test uses the CG recurrence
   * residual; in mixed precision the two drift apart.  Counterpart of
   * AztecOO's TrueResidual() the reference prints after each solve
   * (EP_CrankNicholson.cc:104-109).  With graph = 1 a checked step runs
   * host-driven (the same step, bit for bit). */
  int true_residual_every;
} vg_model;

#define VG_ORDER_AUTO 0      /* mesh order, unless some operator row has a column
                                more than 32767 rows away (a slice on 32-bit
                                columns) and the geometric order has fewer such
                                rows */
#define VG_ORDER_MESH 1      /* always the mesh order */
#define VG_ORDER_GEOMETRIC 2 /* always the geometric order: nodes sorted on a
                                grid of half the median shortest edge, the
                                longest extent slowest, the shortest fastest */

#define VG_SCHEME_CN 0
#define VG_SCHEME_BE 1
#define VG_SCHEME_EXPLICIT 2

/* Initial guess of the CG solves.  Either way a solve stops when
 * ||r||_D <= cg_tol ||K_rhs V||_D, the reference's ||r_0|| when it starts from
 * x_0 = V^n (EP_CrankNicholson.cc:104): the same accuracy target. */
#define VG_X0_ZERO 0     /* increment 0 (= the reference's x_0 = V^n): the
                            reference's Krylov iterates */
#define VG_X0_PREVIOUS 1 /* the previous increment of the same half-step */
#define VG_X0_LAST 2     /* the increment of the last solve (either half-step) */
#define VG_X0_EXTRAPOLATE 3 /* each half-step's increment extrapolated from its
                               own last three, quadratic in time (the first
                               solves use as many as they have) */
#define VG_X0_EXTRAPOLATE_LINEAR 4 /* the same from its last two, linear */

typedef struct {
  int64_t steps;
  int64_t cg_iterations;      /* summed over all half-steps */
  int64_t cg_solves;
  int64_t cg_not_converged;   /* solves that hit cg_maxit */
  int64_t newton_iterations;  /* tT reference integrator only (device sum) */
  int64_t n_nodes;
  int64_t nnz;                /* of the assembled operator, diagonal included */
  int64_t sell_padded_nnz;    /* entries the device stores: off-diagonal, slice-padded */
  double last_cg_relres;      /* ||r||_D / ||K_rhs V||_D of the last solve */
  double sim_time;            /* ms */
  int64_t sell_wide_slices;   /* slices stored with 32-bit columns */
  int32_t row_order;          /* VG_ORDER_MESH or VG_ORDER_GEOMETRIC: the order used */
  int64_t cg_truncated;       /* captured solves still running when the captured
                                 launches ran out: their last iterate was used and
                                 the step re-captured with more (also counted in
                                 cg_not_converged) */
  double max_true_relres;     /* largest FP64 true residual of the checked solves
                                 (vg_model.true_residual_every) */
  double last_true_relres;    /* the last check's */
  int64_t true_residual_checks; /* solves checked (two per checked step) */
  int64_t sell_shared_offset_slices; /* slices whose rows share one set of column
                                        offsets (no per-row columns stored) */
} vg_stats;

const char *vg_version(void);

/* On failure *out may still receive a handle that carries the message for
 * vg_last_error; release it with vg_destroy either way. */
int vg_create(const vg_mesh *mesh, const vg_model *model, vg_handle **out);
void vg_destroy(vg_handle *h);
const char *vg_last_error(const vg_handle *h);
int vg_get_stream(vg_handle *h, void **hip_stream);

int vg_set_stimulus_nodes(vg_handle *h, const int32_t *nodes, int64_t n, double istim);
int vg_set_stimulus_sphere(vg_handle *h, const double center[3], double radius,
                           double istim);
int vg_set_stimulus_axis(vg_handle *h, int axis, double limit, double istim);
int vg_set_stimulus_plane(vg_handle *h, const double origin[3], const double normal[3],
                          double istim);
int vg_clear_stimulus(vg_handle *h);

int vg_step(vg_handle *h, int nsteps);
int vg_synchronize(vg_handle *h);
/* Changes vg_model.true_residual_every between steps. */
int vg_set_true_residual_check(vg_handle *h, int every);

/* Launch shape, changeable between steps (an autotuner times candidates):
 * the total CG grid (rounded up to a multiple of 8; 0 = default) and the
 * captured-step mode of vg_model.graph.  No reference counterpart: the
 * reference has no launch shapes. */
int vg_set_cg_blocks(vg_handle *h, int cg_blocks);
int vg_get_cg_blocks(vg_handle *h, int *cg_blocks);
int vg_set_graph(vg_handle *h, int enable);
/* How the CG's kernels read the operator's values: -1 the library's choice
 * (nontemporal loads when one iteration's streams exceed the 256 MB Infinity
 * Cache), 0 plain loads, 1 nontemporal loads.  Results are the same bit for
 * bit either way; only the cache policy differs.  No reference counterpart:
 * the reference has no device caches to steer. */
int vg_set_value_streaming(vg_handle *h, int mode);
int vg_get_value_streaming(vg_handle *h, int *on);

int vg_get_voltage(vg_handle *h, double *V);
int vg_set_voltage(vg_handle *h, const double *V);
/* Uniform layout [n_nodes * nvar] (nvar of the body's model); only for
 * meshes without Purkinje nodes. */
int vg_get_states(vg_handle *h, int *nvar, double *states);
int vg_set_states(vg_handle *h, int nvar, const double *states);
/* Per-node layout of getStateVariables(numVar, data): num_var[n_nodes] and
 * the node vectors packed in node order (Purkinje nodes carry 22 values).
 * Call with data NULL to get num_var first. */
int vg_get_states_ragged(vg_handle *h, int32_t *num_var, double *data);
int vg_set_states_ragged(vg_handle *h, const int32_t *num_var, const double *data);
int vg_get_lumped_capacitance(vg_handle *h, double *CL);

int vg_activation_enable(vg_handle *h, double threshold_mV);
int vg_get_activation(vg_handle *h, double *t_act);

/* The lead field is built from the same mesh the handle was created with
 * (ComputeECG builds its own Mesh the same way, ComputeECG.cc:76-79). */
int vg_ecg_setup(vg_handle *h, const vg_mesh *mesh, const double *electrodes,
                 int n_electrodes, const double D_ecg[3], double radius);
int vg_ecg(vg_handle *h, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the pseudo-ECG of vg_ecg, of the current V, written to
 * the caller's device array d_out[ne] on the handle's stream (vg_get_stream),
 * stream-ordered, without synchronising, allocating or freeing anything.  Its
 * partial sums live in one file-scope device array per GPU (16 leads x 1024
 * blocks of FP64, 128 KiB, held until the process ends) that every call on
 * that GPU shares: calls that overlap on one GPU (handles on different
 * streams) race.  VG_ERR_ARG when the leads (at most 16) or the handle's
 * node grid (at most 1024 blocks) exceed that array. */
int vg_ecg_device(vg_handle *h, double *d_out);
int vg_ecg_record(vg_handle *h, int steps_per_frame, int max_frames);
int vg_ecg_frames(vg_handle *h, double *out, int *n_frames);

int vg_profile_enable(vg_handle *h, int enable);
int vg_profile_get(vg_handle *h, int kernel_class, double *total_ms, int64_t *launches);
#define VG_KC_RHS 0      /* r_0 = K_rhs V */
#define VG_KC_CG_ITER 1  /* fused CG iteration (update + SpMV), one per launch;
                            timed per solve over its launches back to back
                            (launch boundaries included), counted per launch */
#define VG_KC_CG_START 2 /* w_0 = A r_0 */
#define VG_KC_IONIC 3    /* ionic update, Purkinje kernel included */
#define VG_KC_APPLY 4
#define VG_KC_PK_PRESOLVE 5 /* Purkinje rows of the initial guess, solved locally */
#define VG_KC_COUNT 6

int vg_stats_get(vg_handle *h, vg_stats *out);

/* Host-only assembly (no device needed): CSR of K (with Cm), M and the
 * lumped capacitance.  Call with NULL arrays to get n and nnz first. */
int vg_assemble_csr(const vg_mesh *mesh, const vg_model *model, int64_t *n,
                    int64_t *nnz, int64_t *row_ptr, int32_t *col, double *K,
                    double *M, double *CL);

/* Independent cells (IonicModel adapter): one Compute_Ion per cell per call,
 * run on the device. */
int vg_cells_create(int cell_model, int precision, int integrator, int64_t n_cells,
                    vg_cells **out);
void vg_cells_destroy(vg_cells *c);
int vg_cells_set_scales(vg_cells *c, int n_param_sets, const double *param_scales,
                        const uint16_t *param_index);
int vg_cells_set_quirks(vg_cells *c, unsigned quirks);
int vg_cells_compute_ion(vg_cells *c, double Cm, double dt, const double *V,
                         const double *istim, double *dVdt);
int vg_cells_get_states(vg_cells *c, int *nvar, double *states);
int vg_cells_set_states(vg_cells *c, int nvar, const double *states);

/* ---- the reference's files ----------------------------------------------
 * Mesh input as Mesh::New reads it (Mesh/Mesh.cc:142-416): <model>.node
 * ("n 3 x x", then "id x y z"), <model>.ele ("n x x", then "id k n1..nk",
 * 1-based; k = 4 tetrahedra or 2 bars) and, when present, <model>.fibers
 * (9 numbers per element), .purk ("n", then node ids -> tissue 11), .cell
 * ("id type" pairs) and .angle ("id c0 c1 c2 L" PMJ overrides).  Ids become
 * 0-based.  The partition files (.metis.*) are not needed on one GPU.
 * Errors without a handle go to err[err_len]. */
typedef struct vg_meshfile vg_meshfile;
int vg_mesh_read(const char *model, double radius, vg_meshfile **out, char *err, int err_len);
const vg_mesh *vg_meshfile_mesh(const vg_meshfile *m); /* owned by m */
void vg_meshfile_free(vg_meshfile *m);
/* Writes a mesh in the same files (plus single-partition .metis.npart.1 and
 * .metis.epart.1, which the reference's reader requires). */
int vg_mesh_write(const char *model, const vg_mesh *mesh, char *err, int err_len);
/* "<dir>/<prefix>_time_%08.2f.%02.0f" (EPSolver.cc:121,159). */
int vg_frame_name(char *buf, int len, const char *dir, const char *prefix, double time, int rank);
/* Voltage_time_* : raw FP64 V in node order (EPSolver::WriteOutput,
 * EPSolver.cc:116-136). */
int vg_write_voltage(vg_handle *h, const char *dir, double time, int rank);
/* InterVariables_time_* : per node int32 state count then the FP64 states
 * (EPSolver::writeRestart, EPSolver.cc:155-177). */
int vg_write_restart(vg_handle *h, const char *dir, double time, int rank);
/* Both files back into V and the states (EPSolver::readRestart,
 * EPSolver.cc:180-211). */
int vg_read_restart(vg_handle *h, const char *dir, double time, int rank);
/* Host-only lead field of the pseudo-ECG, H[ne][n_nodes] with
 * ecg_e = sum_n H[e][n] V_n: computeGradOneByR, computeDiffusionTensor and
 * computeECG (Applications/CardiacEP/ComputeECG.h:30-189) folded into one
 * matrix; no GPU needed (post-processing of written frames). */
int vg_ecg_leadfield(const vg_mesh *mesh, unsigned quirks, const double *electrodes, int ne,
                     const double D[3], double radius, double *H, char *err, int err_len);

/* ---- learned ten Tusscher ionic step ------------------------------------
 * No counterpart in the reference: a small MLP (25 inputs: V and the 18
 * other tT states at the start of the ionic update, log dt, the five
 * conductance scalings, each entered as two BF16 halves; 4 hidden layers of
 * 128, SiLU; 19 state increments)
 * trained on the GPU with BF16 MFMA kernels on samples recorded from the
 * plain Rush-Larsen model in tissue, and run in place of that model per node
 * while the node's inputs lie inside the training envelope (the plain model
 * runs the others). */
typedef struct vg_surr vg_surr;
int vg_surr_create(uint64_t seed, vg_surr **out);
void vg_surr_destroy(vg_surr *s);
const char *vg_surr_last_error(const vg_surr *s);
int vg_surr_reserve(vg_surr *s, int64_t capacity);
/* x[n][25] raw inputs, y[n][19] raw increments */
int vg_surr_add_samples(vg_surr *s, const float *x, const float *y, int64_t n);
int vg_surr_num_samples(vg_surr *s, int64_t *n);
int vg_surr_get_samples(vg_surr *s, int64_t first, int64_t n, float *x, float *y);
/* Runs nsteps steps of h (a ten Tusscher body with the Rush-Larsen
 * integrator), recording every `every`-th step the ionic update of every
 * node_stride-th myocardium row into s's samples. */
int vg_surr_collect(vg_handle *h, vg_surr *s, int nsteps, int every, int node_stride);
/* Input normalisation x' = (x - mu_x) / sd_x and input envelope [lo, hi];
 * output transform t = (asinh(y / c_y) - mu_y) / sd_y and output envelope
 * [t_lo, t_hi].  vg_surr_fit_scaling takes them from the samples: c_y a
 * tenth of the mean |y|, envelopes the samples' ranges widened by
 * envelope_margin of themselves.  A node is left to the plain model when an
 * input or a network output falls outside its envelope. */
typedef struct vg_surr_scaling {
  float mu_x[25], sd_x[25], lo[25], hi[25];
  float mu_y[19], sd_y[19], c_y[19], t_lo[19], t_hi[19];
} vg_surr_scaling;
int vg_surr_fit_scaling(vg_surr *s, double envelope_margin);
/* Upstroke fallback: rows whose V exceeds v_min while their sodium
 * inactivation gate h exceeds h_min (depolarised, the channel not yet
 * inactivated: the upstroke that sets the activation time) are left to the
 * plain model.  Off (+inf) by default; takes effect at the next attach. */
int vg_surr_set_upstroke_fallback(vg_surr *s, double v_min, double h_min);
/* Voltage-band fallback: rows with v_lo < V < v_hi (e.g. the repolarisation
 * band) are left to the plain model.  Off (an empty band) by default; takes
 * effect at the next attach. */
int vg_surr_set_band_fallback(vg_surr *s, double v_lo, double v_hi);
int vg_surr_get_scaling(vg_surr *s, vg_surr_scaling *out);
int vg_surr_set_scaling(vg_surr *s, const vg_surr_scaling *in);
int vg_surr_num_params(int64_t *n);
/* per layer W[k][n] row-major then b[n]; layers 50->128->128->128->128->19
 * (layer 0's rows: the 25 inputs' BF16 high halves, then their low halves) */
int vg_surr_get_params(vg_surr *s, float *p);
int vg_surr_set_params(vg_surr *s, const float *p);
/* mean squared error of the transformed outputs t and its gradient (n a
 * multiple of 128) */
int vg_surr_loss_grad(vg_surr *s, const float *x, const float *y, int n, double *loss,
                      float *grad);
int vg_surr_predict(vg_surr *s, const float *x, int n, float *y);
int vg_surr_train(vg_surr *s, int iters, int batch, double lr, double *loss);
/* Run h's ten Tusscher ionic update through s (NULL: the plain model);
 * s must outlive the attachment.  Weights, biases and scaling are copied at
 * attach, one copy per surrogate shared by the handles it is attached to:
 * attach again after training further. */
int vg_surr_attach(vg_handle *h, vg_surr *s);
/* node updates done by the network and by the plain model since attach */
int vg_surr_usage(vg_handle *h, int64_t *network, int64_t *fallback);

#ifdef __cplusplus
}
#endif
#endif
