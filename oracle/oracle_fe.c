/*
 * oracle_fe.c -- FE assembly, Jacobi PCG, Crank-Nicolson Strang step and the
 * pseudo-ECG of voomEP, restated in plain FP64 C.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h for what may use it and what pins each function).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* TetQuadrature::_initialize order 2 (Quadrature/TetQuadrature.cc:16-31):
 * alpha/beta with 8 digits, weight 1/4/6. */
static const double TQ_ALPHA = 0.58541020, TQ_BETA = 0.13819660;
/* LineQuadrature order 3 (Quadrature/LineQuadrature.cc:30-35). */
static const double LQ_X = 0.577350269189626;

/* VoomMath.cc:17-52: cofactor determinant and inverse of a 3x3. */
static double det3(const double F[3][3]) {
  return F[0][0] * F[1][1] * F[2][2] + F[1][0] * F[2][1] * F[0][2] +
         F[0][1] * F[1][2] * F[2][0] - F[0][2] * F[1][1] * F[2][0] -
         F[0][1] * F[1][0] * F[2][2] - F[1][2] * F[2][1] * F[0][0];
}
static void inv3(const double a[3][3], double b[3][3]) {
  const double det = det3(a);
  b[0][0] = (a[1][1] * a[2][2] - a[2][1] * a[1][2]) / det;
  b[1][0] = (a[2][0] * a[1][2] - a[1][0] * a[2][2]) / det;
  b[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  b[0][1] = (a[2][1] * a[0][2] - a[0][1] * a[2][2]) / det;
  b[1][1] = (a[0][0] * a[2][2] - a[2][0] * a[0][2]) / det;
  b[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  b[0][2] = (a[0][1] * a[1][2] - a[1][1] * a[0][2]) / det;
  b[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  b[2][2] = (a[0][0] * a[1][1] - a[1][0] * a[0][1]) / det;
}

/* Quadrature-point data of one tet (MixedMesh3D.cc:110-139 with
 * ShapeTet4CP.cc:4-15): N[q][a], dNdx[q][a][3], w[q] = w_ref * detJ. */
static void tet_qp(const double x[4][3], double N[4][4], double dNdx[4][4][3],
                   double w[4]) {
  static const double DN[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const double pts[4][3] = {{TQ_ALPHA, TQ_BETA, TQ_BETA},
                            {TQ_BETA, TQ_ALPHA, TQ_BETA},
                            {TQ_BETA, TQ_BETA, TQ_ALPHA},
                            {TQ_BETA, TQ_BETA, TQ_BETA}};
  for (int q = 0; q < 4; q++) {
    const double s = pts[q][0], t = pts[q][1], u = pts[q][2];
    N[q][0] = 1.0 - s - t - u;
    N[q][1] = s;
    N[q][2] = t;
    N[q][3] = u;
    double dxds[3][3] = {{0}}, inv[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int c = 0; c < 4; c++) dxds[a][b] += DN[c][b] * x[c][a];
    inv3(dxds, inv);
    const double J = det3(dxds);
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 3; b++) {
        double v = 0.0;
        for (int al = 0; al < 3; al++) v += DN[a][al] * inv[al][b];
        dNdx[q][a][b] = v;
      }
    w[q] = (1.0 / 4.0 / 6.0) * J;
  }
}

/* Conductivity tensor of one element (CardiacBody.cc:138-168). */
static void cond_tensor(const double *fib, const double D[3], double C[3][3]) {
  memset(C, 0, 9 * sizeof(double));
  if (!fib) {
    for (int p = 0; p < 3; p++) C[p][p] = D[p];
    return;
  }
  for (int p = 0; p < 3; p++) {
    const double *f = fib + 3 * p;
    for (int m = 0; m < 3; m++)
      for (int n = 0; n < 3; n++) C[m][n] += D[p] * f[m] * f[n];
  }
}

void or_tet_element(const double x[4][3], const double Dmat[3][3], double Cm,
                    double Ke[4][4], double Me[4][4], double CLe[4]) {
  double N[4][4], dNdx[4][4][3], w[4];
  tet_qp(x, N, dNdx, w);
  memset(Ke, 0, 16 * sizeof(double));
  memset(Me, 0, 16 * sizeof(double));
  for (int a = 0; a < 4; a++) CLe[a] = 0.0;
  /* CardiacBody.cc:171-181 (element K and M), :51-63 (lumped C). */
  for (int q = 0; q < 4; q++) {
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        Me[a][b] += N[q][a] * N[q][b] * w[q];
        for (int m = 0; m < 3; m++)
          for (int n = 0; n < 3; n++)
            Ke[a][b] += Dmat[m][n] * dNdx[q][a][m] * dNdx[q][b][n] * w[q] * Cm;
      }
    for (int a = 0; a < 4; a++) CLe[a] += N[q][a] * Cm * w[q];
  }
}

/* Bar element (MixedMesh3D.cc:65-109, ShapeBar.cc:5-14): 2-point Gauss,
 * dN/dx = -+ang/L, weight = 1 * pi r^2 L per point (sum = 2 pi r^2 L). */
static void bar_element(const double x[2][3], const double *pmj, double radius,
                        const double Dmat[3][3], double Cm, double Ke[2][2],
                        double Me[2][2], double CLe[2]) {
  double ang[3], mag = 0.0;
  for (int j = 0; j < 3; j++) ang[j] = x[1][j] - x[0][j];
  for (int j = 0; j < 3; j++) mag += ang[j] * ang[j];
  mag = sqrt(mag);
  for (int j = 0; j < 3; j++) ang[j] /= mag;
  if (pmj) {
    ang[0] = pmj[0];
    ang[1] = pmj[1];
    ang[2] = pmj[2];
    mag = pmj[3];
  }
  memset(Ke, 0, 4 * sizeof(double));
  memset(Me, 0, 4 * sizeof(double));
  CLe[0] = CLe[1] = 0.0;
  const double xi[2] = {-LQ_X, LQ_X};
  const double fac = 2.0 / mag;
  const double w = 1.0 * M_PI * (radius * radius) * mag;
  for (int q = 0; q < 2; q++) {
    const double N[2] = {0.5 * (1.0 - xi[q]), 0.5 * (1.0 + xi[q])};
    const double DNp[2] = {-0.5, 0.5};
    double dndx[2][3];
    for (int a = 0; a < 2; a++)
      for (int j = 0; j < 3; j++) dndx[a][j] = DNp[a] * ang[j] * fac;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        Me[a][b] += N[a] * N[b] * w;
        for (int m = 0; m < 3; m++)
          for (int n = 0; n < 3; n++)
            Ke[a][b] += Dmat[m][n] * dndx[a][m] * dndx[b][n] * w * Cm;
      }
    for (int a = 0; a < 2; a++) CLe[a] += N[a] * Cm * w;
  }
}

static int cmp_i32(const void *a, const void *b) {
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

static const double *find_pmj(const or_mesh *m, int32_t node) {
  for (int64_t k = 0; k < m->n_pmj; k++)
    if (m->pmj_node[k] == node) return m->pmj_data + 4 * k;
  return NULL;
}

/* CardiacBody::Compute (CardiacBody.cc:78-187) + getDiffusionMatrix /
 * getCapacitanceMatrix (:348-391) + computeLumpedCapacitance (:32-67),
 * assembled into one CSR pattern.  Elements are visited in input order so the
 * sticky PMJ flag (MixedMesh3D.cc:14,31-34) can be reproduced. */
int or_assemble(const or_mesh *m, or_csr *out) {
  const int64_t n = m->n_nodes, ne = m->n_elems;
  int64_t *cnt = calloc(n + 1, sizeof(int64_t));
  if (!cnt) return -1;
  for (int64_t e = 0; e < ne; e++) {
    const int nn = m->elem_nn[e];
    for (int a = 0; a < nn; a++) cnt[m->conn[4 * e + a] + 1] += nn;
  }
  for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  int32_t *cand = malloc(sizeof(int32_t) * (cnt[n] ? cnt[n] : 1));
  int64_t *fill = malloc(sizeof(int64_t) * (n ? n : 1));
  for (int64_t i = 0; i < n; i++) fill[i] = cnt[i];
  for (int64_t e = 0; e < ne; e++) {
    const int nn = m->elem_nn[e];
    for (int a = 0; a < nn; a++) {
      const int32_t r = m->conn[4 * e + a];
      for (int b = 0; b < nn; b++) cand[fill[r]++] = m->conn[4 * e + b];
    }
  }
  int64_t *rp = malloc(sizeof(int64_t) * (n + 1));
  rp[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    int32_t *c = cand + cnt[i];
    const int64_t len = cnt[i + 1] - cnt[i];
    qsort(c, len, sizeof(int32_t), cmp_i32);
    int64_t u = 0;
    for (int64_t k = 0; k < len; k++)
      if (u == 0 || c[k] != c[u - 1]) c[u++] = c[k];
    rp[i + 1] = rp[i] + u;
  }
  const int64_t nnz = rp[n];
  int32_t *col = malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  for (int64_t i = 0; i < n; i++)
    memcpy(col + rp[i], cand + cnt[i], sizeof(int32_t) * (rp[i + 1] - rp[i]));
  free(cand);
  free(cnt);
  free(fill);
  double *K = calloc(nnz ? nnz : 1, sizeof(double));
  double *M = calloc(nnz ? nnz : 1, sizeof(double));
  double *CL = calloc(n ? n : 1, sizeof(double));

  const double *pmj = NULL;
  for (int64_t e = 0; e < ne; e++) {
    const int nn = m->elem_nn[e];
    const int32_t *c = m->conn + 4 * e;
    /* MixedMesh3D.cc:31-34: the flag is set when the first node is a PMJ
     * node and, in the reference, never cleared. */
    const double *hit = find_pmj(m, c[0]);
    if (hit) pmj = hit;
    else if (!m->pmj_sticky) pmj = NULL;
    double Dv[3];
    const double *fib = m->fibers ? m->fibers + 9 * e : NULL;
    if (nn == 2) {
      /* CardiacBody.cc:105-135: LV if either node has tissue type 11. */
      int inLV = m->tissue && (m->tissue[c[0]] == 11 || m->tissue[c[1]] == 11);
      if (!m->fibers) Dv[0] = Dv[1] = Dv[2] = m->Dp[0];
      else
        for (int p = 0; p < 3; p++) Dv[p] = inLV ? m->DpLeft[p] : m->Dp[p];
    } else {
      if (!m->fibers) Dv[0] = Dv[1] = Dv[2] = m->D[0];
      else
        for (int p = 0; p < 3; p++) Dv[p] = m->D[p];
    }
    double Dmat[3][3];
    cond_tensor(fib, Dv, Dmat);
    double Ke[4][4], Me[4][4], CLe[4];
    if (nn == 4) {
      double x[4][3];
      for (int a = 0; a < 4; a++)
        for (int j = 0; j < 3; j++) x[a][j] = m->xyz[3 * (int64_t)c[a] + j];
      or_tet_element(x, Dmat, m->Cm, Ke, Me, CLe);
    } else {
      double x[2][3], k2[2][2], m2[2][2], c2[2];
      for (int a = 0; a < 2; a++)
        for (int j = 0; j < 3; j++) x[a][j] = m->xyz[3 * (int64_t)c[a] + j];
      bar_element(x, pmj, m->radius, Dmat, m->Cm, k2, m2, c2);
      for (int a = 0; a < 2; a++) {
        CLe[a] = c2[a];
        for (int b = 0; b < 2; b++) {
          Ke[a][b] = k2[a][b];
          Me[a][b] = m2[a][b];
        }
      }
    }
    for (int a = 0; a < nn; a++) {
      const int32_t r = c[a];
      CL[r] += CLe[a];
      for (int b = 0; b < nn; b++) {
        int64_t k = rp[r];
        while (col[k] != c[b]) k++;
        K[k] += Ke[a][b];
        M[k] += Me[a][b];
      }
    }
  }
  out->n = n;
  out->nnz = nnz;
  out->row_ptr = rp;
  out->col = col;
  out->K = K;
  out->M = M;
  out->CL = CL;
  return 0;
}

void or_csr_free(or_csr *c) {
  free(c->row_ptr);
  free(c->col);
  free(c->K);
  free(c->M);
  free(c->CL);
  memset(c, 0, sizeof(*c));
}

static void spmv(int64_t n, const int64_t *rp, const int32_t *col,
                 const double *A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += A[k] * x[col[k]];
    y[i] = s;
  }
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* AztecOO CG + AZ_Jacobi as configured in EP_CrankNicholson.cc:58-62 and
 * called at :104/:150: x0 = current x, convergence ||r||2/||r0||2 <= tol.
 * The Jacobi preconditioner is taken as one diagonal scaling; the converged
 * solution does not depend on the variant. */
int or_pcg(int64_t n, const int64_t *rp, const int32_t *col, const double *A,
           const double *b, double *x, double tol, int maxit, double *relres) {
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
  double *dinv = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {
    double d = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
      if (col[k] == i) d = A[k];
    dinv[i] = 1.0 / d;
  }
  spmv(n, rp, col, A, x, q);
  for (int64_t i = 0; i < n; i++) r[i] = b[i] - q[i];
  const double r0 = sqrt(dot(n, r, r));
  int it = 0;
  double rn = r0;
  if (r0 > 0.0) {
    for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    for (int64_t i = 0; i < n; i++) p[i] = z[i];
    double rho = dot(n, r, z);
    while (it < maxit) {
      spmv(n, rp, col, A, p, q);
      const double alpha = rho / dot(n, p, q);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      it++;
      rn = sqrt(dot(n, r, r));
      if (rn <= tol * r0) break;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
      const double rho1 = dot(n, r, z);
      const double beta = rho1 / rho;
      rho = rho1;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
  }
  if (relres) *relres = r0 > 0.0 ? rn / r0 : 0.0;
  free(r);
  free(z);
  free(p);
  free(q);
  free(dinv);
  return it;
}

/* ---- whole-step driver --------------------------------------------------- */
int or_sim_init(or_sim *s) {
  const int64_t n = s->n, nnz = s->csr->nnz;
  s->A = malloc(sizeof(double) * nnz);
  s->B = malloc(sizeof(double) * nnz);
  s->rhs = malloc(sizeof(double) * n);
  if (!s->A || !s->B || !s->rhs) return -1;
  if (s->scheme == OR_SCHEME_BE) {
    /* EP_BackwardEuler.cc:14-27: A = (dT/2) K + diag(C_L). */
    for (int64_t i = 0; i < n; i++)
      for (int64_t k = s->csr->row_ptr[i]; k < s->csr->row_ptr[i + 1]; k++) {
        s->A[k] = s->dt / 2. * s->csr->K[k] + (s->csr->col[k] == i ? s->csr->CL[i] : 0.0);
        s->B[k] = 0.0;
      }
  } else {
    /* EP_CrankNicholson.cc:22-29: A = (dT/4) K + Cm M, B = -(dT/4) K + Cm M. */
    for (int64_t k = 0; k < nnz; k++) {
      s->A[k] = s->dt / 4. * s->csr->K[k] + s->Cm * s->csr->M[k];
      s->B[k] = -s->dt / 4. * s->csr->K[k] + s->Cm * s->csr->M[k];
    }
  }
  s->cg_iters_total = 0;
  s->newton_iters_total = 0;
#ifdef _OPENMP
  if (s->nthreads > 0) omp_set_num_threads(s->nthreads);
#endif
  return 0;
}

void or_sim_free(or_sim *s) {
  free(s->A);
  free(s->B);
  free(s->rhs);
  s->A = s->B = s->rhs = NULL;
}

static void diffusion_half(or_sim *s) {
  const or_csr *c = s->csr;
  spmv(s->n, c->row_ptr, c->col, s->B, s->V, s->rhs);
  s->cg_iters_total += or_pcg(s->n, c->row_ptr, c->col, s->A, s->rhs, s->V,
                              s->cg_tol, s->cg_maxit, NULL);
}

/* EP_CrankNicholson::Solve (EP_CrankNicholson.cc:78-156).  The lumped
 * capacitance "solve" of :124-133 is diagonal, so V += dt * dV/dt. */
int or_cn_step(or_sim *s) {
  diffusion_half(s);
  double total = 0.0;
  double dt = (s->dt_ion < s->dt) ? s->dt_ion : s->dt;
  int64_t newton = 0;
  while (total < s->dt) {
    const int64_t n = s->n;
    if (s->pk_flag) {
      /* CardiacBody::New (CardiacBody.cc:993-1002): Purkinje nodes run the
       * Purkinje model whatever the body's model is. */
      double pkc[OR_PK_NCONST];
      or_pk_constants(pkc);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) {
        if (!s->pk_flag[i]) continue;
        const double dv = or_pk_compute_ion(pkc, s->pk_states + OR_PK_NSTATE * i, s->Cm, dt,
                                            s->V[i], s->istim[i]);
        s->V[i] += dt * dv;
      }
    }
    const int8_t *pkf = s->pk_flag;
    if (s->model == OR_LR) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) {
        if (pkf && pkf[i]) continue;
        double *st = s->states + OR_LR_NSTATE * i;
        const double dv = s->lr_gate ? or_lr_gate_compute_ion(st, s->Cm, dt, s->V[i], s->istim[i])
                                     : or_lr_compute_ion(st, s->Cm, dt, s->V[i], s->istim[i]);
        s->V[i] += dt * dv;
      }
    } else if (s->model == OR_MJ) {
      /* CardiacBody::New case 3 (CardiacBody.cc:1045-1067): constants and
       * initial state by the node's tissue type. */
      double mc[10 * OR_MJ_NCONST]; /* by type; row 9: the fallback of other types */
      for (int t = 0; t < 10; t++) or_mj_constants(t, mc + OR_MJ_NCONST * t);
      const double *mt = s->mj_const ? s->mj_const : mc;
#pragma omp parallel for schedule(dynamic, 256)
      for (int64_t i = 0; i < n; i++) {
        if (pkf && pkf[i]) continue;
        const int t = s->param_idx ? s->param_idx[i] : 0; /* Mesh.cc:186,375 */
        const double dv = or_mj_compute_ion(mt + (int64_t)OR_MJ_NCONST * (t < 9 ? t : 9),
                                            s->states + OR_MJ_NSTATE * i, s->Cm, dt, s->V[i],
                                            s->istim[i], s->mj_int_abs);
        s->V[i] += dt * dv;
      }
    } else {
      double tc[OR_TT_NCONST];
      or_tt_constants(tc);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : newton)
      for (int64_t i = 0; i < n; i++) {
        if (pkf && pkf[i]) continue;
        const double *c = tc;
        if (s->tt_const && s->param_idx)
          c = s->tt_const + (int64_t)OR_TT_NCONST * s->param_idx[i];
        int it = 0;
        const double dv = or_tt_compute_ion(c, s->states + OR_TT_NSTATE * i, s->Cm, dt,
                                            s->V[i], s->istim[i], s->tt_newton_jt, &it);
        newton += it;
        s->V[i] += dt * dv;
      }
    }
    total += dt;
    dt = (total + dt < s->dt) ? dt : s->dt - total;
  }
  s->newton_iters_total += newton;
  diffusion_half(s);
  return 0;
}

int or_adaptive_steps(double dT, double dt_ion, double dv_first, double *out, int max) {
  int k = 0;
  double dt = dt_ion;
  if (!(dT / dt > 1.0)) {
    if (max > 0) out[k++] = dt; /* one stride of dt_ion, even past dT */
    return k;
  }
  if (k < max) out[k++] = dt;
  double tot = dt;
  /* int slope = iion; if (!slope) dt *= 5 (CardiacBody.cc:861-863) */
  if (fabs(dv_first) < 1.0) dt *= 5.0;
  for (;;) {
    const int last = tot + dt > dT;
    const double del = last ? dT - tot : dt;
    if (del > 0.0 && k < max) out[k++] = del;
    tot += del;
    if (last) break;
  }
  return k;
}

/* dV/dt of node i over dt through its model (states advance). */
static double node_ion(or_sim *s, int64_t i, double dt, const double *tc, const double *mt,
                       const double *pkc, int64_t *newton) {
  if (s->pk_flag && s->pk_flag[i])
    return or_pk_compute_ion(pkc, s->pk_states + OR_PK_NSTATE * i, s->Cm, dt, s->V[i], s->istim[i]);
  if (s->model == OR_LR)
    return s->lr_gate ? or_lr_gate_compute_ion(s->states + OR_LR_NSTATE * i, s->Cm, dt, s->V[i],
                                               s->istim[i])
                      : or_lr_compute_ion(s->states + OR_LR_NSTATE * i, s->Cm, dt, s->V[i],
                                          s->istim[i]);
  if (s->model == OR_MJ) {
    const int t = s->param_idx ? s->param_idx[i] : 0;
    return or_mj_compute_ion(mt + (int64_t)OR_MJ_NCONST * (t < 9 ? t : 9),
                             s->states + OR_MJ_NSTATE * i, s->Cm, dt, s->V[i], s->istim[i],
                             s->mj_int_abs);
  }
  const double *c = tc;
  if (s->tt_const && s->param_idx) c = s->tt_const + (int64_t)OR_TT_NCONST * s->param_idx[i];
  int it = 0;
  const double dv = or_tt_compute_ion(c, s->states + OR_TT_NSTATE * i, s->Cm, dt, s->V[i],
                                      s->istim[i], s->tt_newton_jt, &it);
  *newton += it;
  return dv;
}

/* CardiacBody::getVoltageFromIonicCurrent (CardiacBody.cc:819-902): each
 * node steps on its own sub-step sequence. */
static void ionic_adaptive(or_sim *s) {
  double tc[OR_TT_NCONST], pkc[OR_PK_NCONST], mc[10 * OR_MJ_NCONST];
  or_tt_constants(tc);
  or_pk_constants(pkc);
  for (int t = 0; t < 10; t++) or_mj_constants(t, mc + OR_MJ_NCONST * t);
  const double *mt = s->mj_const ? s->mj_const : mc;
  int64_t newton = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : newton)
  for (int64_t i = 0; i < s->n; i++) {
    double dt = s->dt_ion;
    double dv = node_ion(s, i, dt, tc, mt, pkc, &newton);
    s->V[i] += dt * dv;
    double seq[4096];
    const int k = or_adaptive_steps(s->dt, s->dt_ion, dv, seq, 4096);
    for (int q = 1; q < k; q++) {
      dv = node_ion(s, i, seq[q], tc, mt, pkc, &newton);
      s->V[i] += seq[q] * dv;
    }
  }
  s->newton_iters_total += newton;
}

/* One half of EP_ExplicitSolver::Solve's diffusion (EP_ExplicitSolver.cc:
 * 52-85): V += h (-K V) / C_L in strides h = min(dtDiff, dT/2) over dT/2. */
static void explicit_half(or_sim *s) {
  const or_csr *c = s->csr;
  const double TOLER = 1e-6;
  const double dtd = (s->dt_diff < 0.5 * s->dt) ? s->dt_diff : 0.5 * s->dt;
  double solve = 0.5 * s->dt, h = dtd;
  while (solve > TOLER) {
    spmv(s->n, c->row_ptr, c->col, c->K, s->V, s->rhs);
    for (int64_t i = 0; i < s->n; i++) s->V[i] += h * (-s->rhs[i]) / c->CL[i];
    solve -= h;
    h = (h < solve) ? h : solve;
  }
}

/* One half of EP_BackwardEuler::Solve (EP_BackwardEuler.cc:64-67):
 * (C_L + dT/2 K) V' = C_L V from x0 = V. */
static void be_half(or_sim *s) {
  const or_csr *c = s->csr;
  for (int64_t i = 0; i < s->n; i++) s->rhs[i] = c->CL[i] * s->V[i];
  s->cg_iters_total += or_pcg(s->n, c->row_ptr, c->col, s->A, s->rhs, s->V, s->cg_tol,
                              s->cg_maxit, NULL);
}

void or_diffuse_half(or_sim *s) {
  if (s->scheme == OR_SCHEME_BE) be_half(s);
  else if (s->scheme == OR_SCHEME_EXPLICIT) explicit_half(s);
  else diffusion_half(s);
}

int or_step(or_sim *s) {
  if (s->scheme == OR_SCHEME_CN) return or_cn_step(s);
  if (s->scheme == OR_SCHEME_BE) be_half(s);
  else explicit_half(s);
  ionic_adaptive(s);
  if (s->scheme == OR_SCHEME_BE) be_half(s);
  else explicit_half(s);
  return 0;
}

/* ---- pseudo-ECG ------------------------------------------------------------
 * computeGradOneByR / computeDiffusionTensor / computeVoltageAtQuadPoints /
 * computeECG (Applications/CardiacEP/ComputeECG.h:30-189), fused into one
 * pass over elements and quadrature points.  Elements with fewer than 9 nodes
 * (tets and bars) are included; the bar weight uses the ECG radius. */
void or_ecg_direct(const or_mesh *m, const double Decg[3], double radius,
                   const double *V, int ne, const double *el, double *out) {
  for (int e = 0; e < ne; e++) out[e] = 0.0;
  const double *pmj = NULL;
  for (int64_t k = 0; k < m->n_elems; k++) {
    const int nn = m->elem_nn[k];
    const int32_t *c = m->conn + 4 * k;
    const double *hit = find_pmj(m, c[0]);
    if (hit) pmj = hit;
    else if (!m->pmj_sticky) pmj = NULL;
    double Dmat[3][3];
    cond_tensor(m->fibers ? m->fibers + 9 * k : NULL, Decg, Dmat);
    if (!m->fibers) {
      memset(Dmat, 0, sizeof(Dmat));
      for (int p = 0; p < 3; p++) Dmat[p][p] = Decg[0];
    }
    double x[4][3];
    for (int a = 0; a < nn; a++)
      for (int j = 0; j < 3; j++) x[a][j] = m->xyz[3 * (int64_t)c[a] + j];
    double N[4][4], dNdx[4][4][3], w[4];
    int nq;
    if (nn == 4) {
      tet_qp(x, N, dNdx, w);
      nq = 4;
    } else {
      double ang[3], mag = 0.0;
      for (int j = 0; j < 3; j++) ang[j] = x[1][j] - x[0][j];
      for (int j = 0; j < 3; j++) mag += ang[j] * ang[j];
      mag = sqrt(mag);
      for (int j = 0; j < 3; j++) ang[j] /= mag;
      if (pmj) {
        ang[0] = pmj[0];
        ang[1] = pmj[1];
        ang[2] = pmj[2];
        mag = pmj[3];
      }
      const double xi[2] = {-LQ_X, LQ_X};
      for (int q = 0; q < 2; q++) {
        N[q][0] = 0.5 * (1.0 - xi[q]);
        N[q][1] = 0.5 * (1.0 + xi[q]);
        for (int j = 0; j < 3; j++) {
          dNdx[q][0][j] = -0.5 * ang[j] * (2.0 / mag);
          dNdx[q][1][j] = 0.5 * ang[j] * (2.0 / mag);
        }
        w[q] = 1.0 * M_PI * radius * radius * mag;
      }
      nq = 2;
    }
    for (int q = 0; q < nq; q++) {
      double pos[3] = {0, 0, 0}, gV[3] = {0, 0, 0};
      for (int a = 0; a < nn; a++)
        for (int j = 0; j < 3; j++) {
          pos[j] += N[q][a] * x[a][j];
          gV[j] += dNdx[q][a][j] * V[c[a]];
        }
      for (int e = 0; e < ne; e++) {
        double d2 = 0.0, g[3];
        for (int j = 0; j < 3; j++) d2 += pow(pos[j] - el[3 * e + j], 2);
        const double dist = sqrt(d2);
        if (dist < 1e-6) continue;
        for (int j = 0; j < 3; j++) g[j] = -(pos[j] - el[3 * e + j]) / pow(dist, 3);
        double acc = 0.0;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) acc += Dmat[a][b] * gV[a] * g[b] * w[q];
        out[e] += acc;
      }
    }
  }
}
