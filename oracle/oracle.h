/*
 * oracle.h -- CPU FP64 restatement of voomEP's monodomain time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (ucla_cmg_amd/, include/)
 * may include, link or call this code.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, and only as the checker or as the
 * timed CPU baseline.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * voomEP/src/).  The reference itself cannot be compiled in this image: its
 * ionic translation units include voom.h, which pulls in Blitz++ and tvmet
 * (absent), and the solver needs Trilinos Epetra/AztecOO (absent).
 *
 * Parity pins (what checks each function beyond its own formula):
 *   or_assemble        closed-form P1 K_e = V G D G^T, M_e = V/20 (1+d_ab),
 *                      C_L = V/4, K*1 = 0, symmetry, sum(M) = Cm*Vol, bar
 *                      closed forms (tests/test_oracle_fe.py)
 *   or_pcg             dense numpy/scipy solve (tests/test_oracle_solver.py)
 *   or_cn_step         CN eigen-mode decay factor (tests/test_oracle_solver.py)
 *   or_tt_rates        charge conservation across dV/dt and the six
 *                      concentration equations; near-stationarity at the
 *                      reference's initial state; constants and initial
 *                      state against tests/golden/ionic_tables.json; gate
 *                      steady states at their published half points and
 *                      slopes, time constants at exact points, RyR recovery
 *                      (tests/test_oracle_gate_kinetics.py)
 *   or_tt_compute_ion  backward-Euler gate closed form; Newton fixed point
 *                      solves x - x^n - dt f(x) = 0; J and J^T linear solves
 *                      against numpy.linalg.solve
 *   or_lr_compute_ion  Rush-Larsen gate closed form at fixed V; Ca_i forward
 *                      Euler closed form with the slow inward current;
 *                      constants against the golden table; every alpha/beta
 *                      of make_parameter1 at its centre, its asymptotic
 *                      slopes and its guarded singularity's limit; I_Na
 *                      reversal and conductance; the background K currents
 *                      vanishing at E_K1 (tests/test_oracle_gate_kinetics.py);
 *                      the gate table against the formulas
 *                      (tests/test_oracle_lr_gate.py)
 *   or_pk_rates        constants, derived constants and initial state against
 *                      the reference tables; logistic steady states at their
 *                      half points and slopes, the other steady states and
 *                      time constants at exact points; guarded singularities;
 *                      charge conservation; currents vanishing at their
 *                      reversal potentials; the exchanger; the adaptive
 *                      sub-steps; AP markers (tests/test_oracle_purkinje.py)
 *   all four models    a planted-mistake harness: 41 one-spot mutations of
 *                      oracle_ionic.c (a digit, sign, centre, slope, a dropped
 *                      current, a stoichiometric factor, a state index), each
 *                      of which fails at least one pin
 *                      (tests/test_oracle_mutations.py)
 *   or_ecg_direct      brute-force numpy quadrature sum, uniform V -> 0,
 *                      dipole far-field scaling
 *   or_mj_rates        Ca and Na conservation across the flux terms; the
 *                      Markov I_CaL rows sum to -dpo/dt; gate relaxation
 *                      closed form; near-stationarity at the reference's
 *                      initial states; constants and initial states against
 *                      tests/golden/ionic_tables.json
 */
#ifndef VOOM_ORACLE_H
#define VOOM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mixed mesh: elements with 4 nodes (P1 tets) or 2 nodes (Purkinje bars),
 * 0-based connectivity, 4 slots per element (bars use the first 2). */
typedef struct {
  int64_t n_nodes;
  const double *xyz;        /* [3*n_nodes] */
  int64_t n_elems;
  const int8_t *elem_nn;    /* [n_elems] 4 or 2 */
  const int32_t *conn;      /* [4*n_elems] */
  const double *fibers;     /* [9*n_elems] rows f,s,n per element, or NULL (isotropic) */
  const int8_t *tissue;     /* [n_nodes] tissue type (11 = LV Purkinje) or NULL */
  int64_t n_pmj;
  const int32_t *pmj_node;  /* [n_pmj] 0-based node id keyed like MixedMesh3D _orient */
  const double *pmj_data;   /* [4*n_pmj] cosines(3), length */
  int pmj_sticky;           /* 1 = reproduce the never-reset PMJ flag */
  double radius;            /* bar radius (cm) */
  double D[3], Dp[3], DpLeft[3];
  double Cm;
} or_mesh;

/* CSR with K (stiffness, Cm included as in CardiacBody::Compute) and M
 * (consistent mass without Cm) sharing one sparsity pattern. */
typedef struct {
  int64_t n;
  int64_t nnz;
  int64_t *row_ptr;   /* [n+1] */
  int32_t *col;       /* [nnz] */
  double *K;          /* [nnz] */
  double *M;          /* [nnz] */
  double *CL;         /* [n] lumped capacitance */
} or_csr;

int or_assemble(const or_mesh *m, or_csr *out);
void or_csr_free(or_csr *c);

/* Element-level access for closed-form tests. */
void or_tet_element(const double x[4][3], const double Dmat[3][3], double Cm,
                    double Ke[4][4], double Me[4][4], double CLe[4]);

/* Jacobi PCG, x0 = x on entry, stop at ||r||2 <= tol*||r0||2 (AztecOO AZ_r0). */
int or_pcg(int64_t n, const int64_t *rp, const int32_t *col, const double *A,
           const double *b, double *x, double tol, int maxit, double *relres);

/* ---- ionic models ------------------------------------------------------ */
enum { OR_LR = 0, OR_TT = 2, OR_MJ = 3 };

#define OR_LR_NSTATE 8
#define OR_TT_NSTATE 19
#define OR_TT_NCONST 53
#define OR_TT_NALG 70

void or_lr_init(double *s);
double or_lr_xi(void);
double or_lr_compute_ion(double *s, double Cm, double dt, double V, double istim);
/* LUORUDYWITHGATE: the same model with the 14 gate/background values taken
 * from LuoRudyGateTable's 0.001 mV table (LRGateTable.h, LR.cc:71-79). */
double or_lr_gate_compute_ion(double *in, double Cm, double dt, double volt, double istim);
/* The 14 make_parameter1 values at v: m,h,j,d,f,X (inf, alpha+beta), x1 and
 * I_K1 + I_Kp + I_b; gate = 1 through the table (tests). */
void or_lr_tab(double v, int gate, double tab[14]);

void or_tt_constants(double *c);
void or_tt_init(double *s);
void or_tt_rates(const double *c, double *state, double *alg, double *rates);
double or_tt_compute_ion(const double *c, double *s, double Cm, double dt,
                         double V, double istim, int jt, int *newton_iters);
/* One Newton linear solve exactly as the reference forms it (J^T). */
void or_tt_newton_system(const double *c, double *s, const double *stateN,
                         const double *oldRates, double dt, double alg12,
                         double Jac[36], double B[6]);
int or_solve6_lapack_layout(const double Jac[36], double B[6]);
int or_solve6(const double Jac[36], double B[6], int transpose);

#define OR_PK_NSTATE 22
#define OR_PK_NCONST 63
#define OR_PK_NALG 69
void or_pk_constants(double *c);
void or_pk_init(double *s);
void or_pk_rates(const double *c, const double *state, double *alg, double *rates);
double or_pk_compute_ion(const double *c, double *s, double Cm, double dt, double V,
                         double istim);

#define OR_MJ_NSTATE 26
#define OR_MJ_NCONST 35
void or_mj_constants(int type, double *c);
void or_mj_fail_constants(int type, double *c); /* MAHAJAN_FAIL's set */
void or_mj_init(int type, double *s);
double or_mj_xi(void);
void or_mj_rates(const double *c, const double *s, double h, double istim_uA_per_uF,
                 double *rates);
double or_mj_compute_ion(const double *c, double *s, double Cm, double dt, double V,
                         double istim, int int_abs);

/* ---- whole-step driver ------------------------------------------------- */
typedef struct {
  int model;            /* OR_LR, OR_TT or OR_MJ */
  int64_t n;
  const or_csr *csr;
  double dt, dt_ion, Cm, cg_tol;
  int cg_maxit;
  double *V;            /* [n] */
  double *states;       /* [n*nstate] AoS like getInternalParameters */
  double *istim;        /* [n] uA/cm^3 */
  const double *tt_const; /* [n_param*53] or NULL -> reference constants */
  const uint16_t *param_idx; /* [n] or NULL; Mahajan: tissue type 0-8 (NULL: 0) */
  /* work */
  double *A, *B, *rhs, *r, *z, *p, *q, *diag;
  int64_t cg_iters_total;
  int64_t newton_iters_total;
  int nthreads;
  int tt_newton_jt;     /* 1 = solve Newton steps with J^T like the reference */
  const int8_t *pk_flag; /* [n] 1 = Purkinje node (runs the Purkinje model), or NULL */
  double *pk_states;    /* [n*22] AoS, used where pk_flag is set */
  const double *mj_const; /* [10*35] Mahajan constants by tissue type (row 9:
                             types beyond 8), or NULL (the normal set) */
  int mj_int_abs;       /* 1 = adaptive threshold read through int abs() */
  int scheme;           /* OR_SCHEME_*: the step of EPSolver::New's Method */
  int lr_gate;          /* LR: 1 = LUORUDYWITHGATE (table of LRGateTable.h) */
  double dt_diff;       /* explicit diffusion sub-step (Solve's dtDiff, 0.1 ms) */
} or_sim;

#define OR_SCHEME_CN 0       /* EP_CrankNicholson */
#define OR_SCHEME_BE 1       /* EP_BackwardEuler */
#define OR_SCHEME_EXPLICIT 2 /* EP_ExplicitSolver */

/* Sub-steps of CardiacBody::getVoltageFromIonicCurrent
 * (CardiacBody.cc:819-902) for a step dT, ionic step dt_ion and the dV/dt
 * of the first sub-step: dt_ion, then strides of dt_ion (5 dt_ion when that
 * first dV/dt truncates to int 0) up to dT, the last one clipped; a
 * zero-length final call (reached when the strides land on dT exactly)
 * changes nothing and is left out.  Returns the count (<= max). */
int or_adaptive_steps(double dT, double dt_ion, double dv_first, double *out, int max);

int or_sim_init(or_sim *s);
void or_sim_free(or_sim *s);
/* One EP_CrankNicholson::Solve. */
int or_cn_step(or_sim *s);
/* One EPSolver::Solve of the sim's scheme (CN, backward Euler, explicit). */
int or_step(or_sim *s);
/* The diffusion half of the sim's scheme alone (tests). */
void or_diffuse_half(or_sim *s);

/* ---- pseudo-ECG (ComputeECG.h:30-189) ---------------------------------- */
void or_ecg_direct(const or_mesh *m, const double Decg[3], double radius,
                   const double *V, int ne, const double *electrodes,
                   double *out);

#ifdef __cplusplus
}
#endif
#endif
