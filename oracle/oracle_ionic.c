/*
 * oracle_ionic.c -- Luo-Rudy and ten Tusscher 2006 cell models of voomEP,
 * restated in plain FP64 C with the reference's own integrators.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

/* ======================= Luo-Rudy (IonicModel/LR.cc) ======================= */
/* State layout of LuoRudy::_in (LR.cc:18-26, order fixed by the gate table
 * of make_parameter1 :166-181): 0 m, 1 h, 2 j, 3 d, 4 f, 5 X, 6 Ca_i, 7 V. */
void or_lr_init(double *s) {
  s[0] = .001;
  s[1] = 0.0001;
  s[2] = 0.99;
  s[3] = .001;
  s[4] = 0.9;
  s[5] = .001;
  s[6] = 0.0000001;
  s[7] = 0;
}

/* LR.cc:36-37: Xi = 2 (r + L) / (r L). */
double or_lr_xi(void) {
  const double r = 11e-4, L = 100E-4;
  return 2. * (r + L) / (r * L);
}

typedef struct {
  double Gna, Gsi, GK, E_K1, G_K1;
} lr_const;

/* LuoRudy::initialize (LR.cc:39-57), with the reference's modified Gsi, GK. */
static lr_const lr_constants(void) {
  lr_const c;
  const double c_K0 = 5.4, c_Ki = 145.0, RT_F = 26.73;
  c.Gna = 12.00;
  c.Gsi = 0.07;
  c.GK = 0.705 * sqrt(c_K0 / 5.4);
  c.E_K1 = RT_F * log(c_K0 / c_Ki);
  c.G_K1 = .6047 * sqrt(c_K0 / 5.4);
  return c;
}

/* LuoRudy::make_parameter1 (LR.cc:102-183). */
static void lr_table(const lr_const *C, double v, double tab[14]) {
  double alpha_m, beta_m, alpha_h, beta_h, alpha_j, beta_j;
  if (fabs(v + 47.13) < 0.001) alpha_m = 3.2;
  else alpha_m = 0.32 * (v + 47.13) / (1.0 - exp(-0.1 * (v + 47.13)));
  beta_m = 0.08 * exp(-v / 11.);
  if (v > -40.) {
    alpha_h = 0;
    alpha_j = 0;
    beta_h = 1. / (0.13 * (1 + exp(-(v + 10.66) / 11.1)));
    beta_j = 0.3 * exp(-2.535e-7 * v) / (1 + exp(-0.1 * (v + 32)));
  } else {
    alpha_h = 0.135 * exp(-(80 + v) / 6.8);
    alpha_j = (-1.2714e5 * exp(0.2444 * v) - 3.474e-5 * exp(-0.04391 * v)) *
              (v + 37.78) / (1 + exp(0.311 * (v + 79.23)));
    beta_h = 3.56 * exp(0.079 * v) + 3.1e5 * exp(0.35 * v);
    beta_j = 0.1212 * exp(-0.01052 * v) / (1 + exp(-0.1378 * (v + 40.14)));
  }
  const double alpha_d = .095 * exp(-.01 * (v - 5)) / (1 + exp(-.072 * (v - 5)));
  const double beta_d = .07 * exp(-.017 * (v + 44)) / (1 + exp(.05 * (v + 44)));
  const double alpha_f = .012 * exp(-.008 * (v + 28)) / (1 + exp(.15 * (v + 28)));
  const double beta_f = .0065 * exp(-.02 * (v + 30)) / (1 + exp(-.2 * (v + 30)));
  const double alpha_x = .0005 * exp(.083 * (v + 50)) / (1 + exp(.057 * (v + 50)));
  const double beta_x = .0013 * exp(-.06 * (v + 20)) / (1 + exp(-.04 * (v + 20)));
  const double tao_m = alpha_m + beta_m, tao_h = alpha_h + beta_h;
  const double tao_j = alpha_j + beta_j, tao_d = alpha_d + beta_d;
  const double tao_f = alpha_f + beta_f, tao_x = alpha_x + beta_x;
  double x1;
  if (v > -100) {
    if (fabs(v + 77.) < 0.001) x1 = 0.608889;
    else x1 = 2.837 * (exp(0.04 * (v + 77)) - 1) / ((v + 77) * exp(0.04 * (v + 35)));
  } else {
    x1 = 1.;
  }
  const double E_K1 = C->E_K1;
  const double alpha_k1 = 1.02 / (1 + exp(0.2385 * (v - E_K1 - 59.215)));
  const double beta1 = 0.49124 * exp(0.08032 * (v - E_K1 + 5.476));
  const double beta2 = exp(0.06175 * (v - E_K1 - 594.31));
  const double beta3 = 1 + exp(-0.5143 * (v - E_K1 + 4.753));
  const double beta_k1 = (beta1 + beta2) / beta3;
  const double k1_inf = alpha_k1 / (alpha_k1 + beta_k1);
  const double I_k1 = C->G_K1 * k1_inf * (v - E_K1);
  const double kp = 1. / (1 + exp((7.488 - v) / 5.98));
  const double I_kp = 0.0183 * kp * (v - E_K1);
  const double I_b = 0.0392 * (v + 59.87);
  tab[0] = alpha_m / tao_m;
  tab[1] = tao_m;
  tab[2] = alpha_h / tao_h;
  tab[3] = tao_h;
  tab[4] = alpha_j / tao_j;
  tab[5] = tao_j;
  tab[6] = alpha_d / (alpha_d + beta_d);
  tab[7] = tao_d;
  tab[8] = alpha_f / (alpha_f + beta_f);
  tab[9] = tao_f;
  tab[10] = alpha_x / (alpha_x + beta_x);
  tab[11] = tao_x;
  tab[12] = x1;
  tab[13] = I_k1 + I_kp + I_b;
}

/* LuoRudy::Compute_Ion (LR.cc:65-100) without the gate table. */
double or_lr_compute_ion(double *in, double Cm, double dt, double volt,
                         double istim) {
  const lr_const C = lr_constants();
  double tabz[14];
  in[7] = volt;
  lr_table(&C, in[7], tabz);
  for (int i = 0; i < 6; i++)
    in[i] = tabz[2 * i] - (tabz[2 * i] - in[i]) * exp(-dt * tabz[2 * i + 1]);
  const double Es = 7.7 - 13.0287 * log(in[6]);
  const double i_na = C.Gna * in[0] * in[0] * in[0] * in[1] * in[2] * (in[7] - 54.4);
  const double i_si = C.Gsi * in[3] * in[4] * (in[7] - Es);
  const double i_k = C.GK * tabz[12] * in[5] * (in[7] + 77.0);
  in[6] = in[6] + (-.0001 * i_si + 0.07 * (0.0001 - in[6])) * dt;
  const double i_tot = -(tabz[13] + i_k + i_na + i_si);
  return (i_tot + istim / (or_lr_xi() * Cm));
}

/* =================== ten Tusscher 2006 (IonicModel/Tusscher.cc) ============ */
/* SetUpTusscherParameters (IonicModel/SetUpIonicConstants.h:88-147). */
void or_tt_constants(double *C) {
  static const double k[OR_TT_NCONST] = {
      8314.472, 310,      96485.3415, 0.185,     0.016404,  10,       1000,
      1,        52,       0.03,       5.4,       140,       2,        5.405,
      0.153,    0.392,    14.838,     0.00029,   0.0000398, 0.000592, 0.294,
      2.724,    1,        40,         1000,      0.1,       2.5,      0.35,
      1.38,     87.5,     0.1238,     0.0005,    0.0146,    0.15,     0.045,
      0.06,     0.005,    1.5,        2.5,       1,         0.102,    0.0038,
      0.00025,  0.00036,  0.006375,   0.2,       0.001,     10,       0.3,
      0.4,      0.00025,  0.001094,   0.00005468};
  memcpy(C, k, sizeof(k));
}

/* Tusscher constructor, "modified" initial state (Tusscher.h:51-69). */
void or_tt_init(double *s) {
  static const double v[OR_TT_NSTATE] = {
      -85.423, 136.89, 8.604,  0.000126, 0.00621,  0.4712,  0.0095,
      0.00172, 0.7444, 0.7045, 0.00036,  3.373e-5, 0.7888,  0.9755,
      0.9953,  0.999998, 2.42e-8, 3.64,  0.9073};
  memcpy(s, v, sizeof(v));
}

#define S state
#define A alg
#define R rates
/* Tusscher::ComputeRates (Tusscher.cc:4-96).  alg[12] (the stimulus current)
 * is an input set by the caller and is not written here. */
void or_tt_rates(const double *C, double *S, double *A, double *R) {
  const double RTF = (C[0] * C[1]) / C[2];
  A[7] = 1.0 / (1.0 + exp((S[0] + 20.0) / 7.0));
  A[20] = 1102.5 * exp(-pow(S[0] + 27.0, 2.0) / 225.0) +
          200.0 / (1.0 + exp((13.0 - S[0]) / 10.0)) +
          180.0 / (1.0 + exp((S[0] + 30.0) / 10.0)) + 20.0;
  R[12] = (A[7] - S[12]) / A[20];
  A[8] = 0.67 / (1.0 + exp((S[0] + 35.0) / 7.0)) + 0.33;
  A[21] = 562.0 * exp(-pow(S[0] + 27.0, 2.0) / 240.0) +
          31.0 / (1.0 + exp((25.0 - S[0]) / 10.0)) +
          80.0 / (1.0 + exp((S[0] + 30.0) / 10.0));
  R[13] = (A[8] - S[13]) / A[21];
  A[9] = 0.6 / (1.0 + pow(S[10] / 0.05, 2.0)) + 0.4;
  A[22] = 80.0 / (1.0 + pow(S[10] / 0.05, 2.0)) + 2.0;
  R[14] = (A[9] - S[14]) / A[22];
  A[10] = 1.0 / (1.0 + exp((S[0] + 20.0) / 5.0));
  A[23] = 85.0 * exp(-pow(S[0] + 45.0, 2.0) / 320.0) +
          5.0 / (1.0 + exp((S[0] - 20.0) / 5.0)) + 3.0;
  R[15] = (A[10] - S[15]) / A[23];
  A[11] = 1.0 / (1.0 + exp((20.0 - S[0]) / 6.0));
  A[24] = 9.5 * exp(-pow(S[0] + 40.0, 2.0) / 1800.0) + 0.8;
  R[16] = (A[11] - S[16]) / A[24];
  A[0] = 1.0 / (1.0 + exp((-26.0 - S[0]) / 7.0));
  A[13] = 450.0 / (1.0 + exp((-45.0 - S[0]) / 10.0));
  A[26] = 6.0 / (1.0 + exp((S[0] + 30.0) / 11.5));
  A[34] = 1.0 * A[13] * A[26];
  R[4] = (A[0] - S[4]) / A[34];
  A[1] = 1.0 / (1.0 + exp((S[0] + 88.0) / 24.0));
  A[14] = 3.0 / (1.0 + exp((-60.0 - S[0]) / 20.0));
  A[27] = 1.12 / (1.0 + exp((S[0] - 60.0) / 20.0));
  A[35] = 1.0 * A[14] * A[27];
  R[5] = (A[1] - S[5]) / A[35];
  A[2] = 1.0 / (1.0 + exp((-5.0 - S[0]) / 14.0));
  A[15] = 1400.0 / pow(1.0 + exp((5.0 - S[0]) / 6.0), 1.0 / 2);
  A[28] = 1.0 / (1.0 + exp((S[0] - 35.0) / 15.0));
  A[36] = 1.0 * A[15] * A[28] + 80.0;
  R[6] = (A[2] - S[6]) / A[36];
  A[3] = 1.0 / pow(1.0 + exp((-56.86 - S[0]) / 9.03), 2.0);
  A[16] = 1.0 / (1.0 + exp((-60.0 - S[0]) / 5.0));
  A[29] = 0.1 / (1.0 + exp((S[0] + 35.0) / 5.0)) +
          0.1 / (1.0 + exp((S[0] - 50.0) / 200.0));
  A[37] = 1.0 * A[16] * A[29];
  R[7] = (A[3] - S[7]) / A[37];
  A[4] = 1.0 / pow(1.0 + exp((S[0] + 71.55) / 7.43), 2.0);
  A[17] = (S[0] < -40.0 ? 0.057 * exp(-(S[0] + 80.0) / 6.8) : 0.0);
  A[30] = (S[0] < -40.0
               ? 2.7 * exp(0.079 * S[0]) + 310000. * exp(0.3485 * S[0])
               : 0.77 / (0.13 * (1.0 + exp((S[0] + 10.66) / -11.1))));
  A[38] = 1.0 / (A[17] + A[30]);
  R[8] = (A[4] - S[8]) / A[38];
  A[5] = 1.0 / pow(1.0 + exp((S[0] + 71.55) / 7.43), 2.0);
  A[18] = (S[0] < -40.0
               ? ((-25428.0 * exp(0.2444 * S[0]) - 6.948e-06 * exp(-0.04391 * S[0])) *
                  (S[0] + 37.78) / 1.0) /
                     (1.0 + exp(0.311 * (S[0] + 79.23)))
               : 0.0);
  A[31] = (S[0] < -40.0
               ? (0.02424 * exp(-0.01052 * S[0])) / (1.0 + exp(-0.1378 * (S[0] + 40.14)))
               : (0.6 * exp(0.057 * S[0])) / (1.0 + exp(-0.1 * (S[0] + 32.0))));
  A[39] = 1.0 / (A[18] + A[31]);
  R[9] = (A[5] - S[9]) / A[39];
  A[6] = 1.0 / (1.0 + exp((-8.0 - S[0]) / 7.5));
  A[19] = 1.4 / (1.0 + exp((-35.0 - S[0]) / 13.0)) + 0.25;
  A[32] = 1.4 / (1.0 + exp((S[0] + 5.0) / 5.0));
  A[40] = 1.0 / (1.0 + exp((50.0 - S[0]) / 20.0));
  A[42] = 1.0 * A[19] * A[32] + A[40];
  R[11] = (A[6] - S[11]) / A[42];
  A[55] = (((C[21] * C[10]) / (C[10] + C[22])) * S[2] / (S[2] + C[23])) /
          (1.0 + 0.1245 * exp((-0.1 * S[0] * C[2]) / (C[0] * C[1])) +
           0.0353 * exp((-S[0] * C[2]) / (C[0] * C[1])));
  A[25] = RTF * log(C[11] / S[2]);
  A[50] = C[16] * pow(S[7], 3.0) * S[8] * S[9] * (S[0] - A[25]);
  A[51] = C[17] * (S[0] - A[25]);
  A[56] = (C[24] * (exp((C[27] * S[0] * C[2]) / (C[0] * C[1])) * pow(S[2], 3.0) * C[12] -
                    exp(((C[27] - 1.0) * S[0] * C[2]) / (C[0] * C[1])) * pow(C[11], 3.0) *
                        S[3] * C[26])) /
          ((pow(C[29], 3.0) + pow(C[11], 3.0)) * (C[28] + C[12]) *
           (1.0 + C[25] * exp(((C[27] - 1.0) * S[0] * C[2]) / (C[0] * C[1]))));
  R[2] = ((-1.0 * (A[50] + A[51] + 3.0 * A[55] + 3.0 * A[56])) / (1.0 * C[4] * C[2])) * C[3];
  A[33] = RTF * log(C[10] / S[1]);
  A[44] = 0.1 / (1.0 + exp(0.06 * ((S[0] - A[33]) - 200.0)));
  A[45] = (3.0 * exp(0.0002 * ((S[0] - A[33]) + 100.0)) + exp(0.1 * ((S[0] - A[33]) - 10.0))) /
          (1.0 + exp(-0.5 * (S[0] - A[33])));
  A[46] = A[44] / (A[44] + A[45]);
  A[47] = C[13] * A[46] * pow(C[10] / 5.4, 1.0 / 2) * (S[0] - A[33]);
  A[54] = C[20] * S[16] * S[15] * (S[0] - A[33]);
  A[48] = C[14] * pow(C[10] / 5.4, 1.0 / 2) * S[4] * S[5] * (S[0] - A[33]);
  A[41] = RTF * log((C[10] + C[9] * C[11]) / (S[1] + C[9] * S[2]));
  A[49] = C[15] * pow(S[6], 2.0) * (S[0] - A[41]);
  A[52] = ((C[18] * S[11] * S[12] * S[13] * S[14] * 4.0 * (S[0] - 15.0) * pow(C[2], 2.0)) /
           (C[0] * C[1])) *
          (0.25 * S[10] * exp((2.0 * (S[0] - 15.0) * C[2]) / (C[0] * C[1])) - C[12]) /
          (exp((2.0 * (S[0] - 15.0) * C[2]) / (C[0] * C[1])) - 1.0);
  A[43] = ((0.5 * C[0] * C[1]) / C[2]) * log(C[12] / S[3]);
  A[53] = C[19] * (S[0] - A[43]);
  A[58] = (C[32] * (S[0] - A[33])) / (1.0 + exp((25.0 - S[0]) / 5.98));
  A[57] = (C[30] * S[3]) / (S[3] + C[31]);
  R[0] = -(A[47] + A[54] + A[48] + A[49] + A[52] + A[55] + A[50] + A[51] + A[56] + A[53] +
           A[58] + A[57] + A[12]);
  R[1] = ((-1.0 * ((A[47] + A[54] + A[48] + A[49] + A[58] + A[12]) - 2.0 * A[55])) /
          (1.0 * C[4] * C[2])) *
         C[3];
  A[59] = C[44] / (1.0 + pow(C[42], 2.0) / pow(S[3], 2.0));
  A[60] = C[43] * (S[17] - S[3]);
  A[61] = C[41] * (S[10] - S[3]);
  A[63] = 1.0 / (1.0 + (C[45] * C[46]) / pow(S[3] + C[46], 2.0));
  R[3] = A[63] * ((((A[60] - A[59]) * C[51]) / C[4] + A[61]) -
                  (1.0 * ((A[53] + A[57]) - 2.0 * A[56]) * C[3]) / (2.0 * 1.0 * C[4] * C[2]));
  A[62] = C[38] - (C[38] - C[39]) / (1.0 + pow(C[37] / S[17], 2.0));
  A[65] = C[34] * A[62];
  R[18] = -A[65] * S[10] * S[18] + C[36] * (1.0 - S[18]);
  A[64] = C[33] / A[62];
  A[66] = (A[64] * pow(S[10], 2.0) * S[18]) / (C[35] + A[64] * pow(S[10], 2.0));
  A[67] = C[40] * A[66] * (S[17] - S[10]);
  A[68] = 1.0 / (1.0 + (C[47] * C[48]) / pow(S[17] + C[48], 2.0));
  R[17] = A[68] * (A[59] - (A[67] + A[60]));
  A[69] = 1.0 / (1.0 + (C[49] * C[50]) / pow(S[10] + C[50], 2.0));
  R[10] = A[69] * (((-1.0 * A[52] * C[3]) / (2.0 * 1.0 * C[52] * C[2]) + (A[67] * C[51]) / C[52]) -
                   (A[61] * C[4]) / C[52]);
}
#undef S
#undef A
#undef R

static const int TT_IND[12] = {4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16};
static const int TT_DIND[12] = {34, 35, 36, 37, 38, 39, 42, 20, 21, 22, 23, 24};
static const int TT_NR[6] = {1, 2, 3, 10, 17, 18};

/* Gaussian elimination with partial pivoting on a 6x6 system held row-major
 * in Jac (Jac[j*6+i] = dg_j/dx_i).  transpose = 1 solves the system exactly
 * as a column-major LAPACK dgesv sees that array (entry (r,c) = Jac[c*6+r],
 * i.e. J^T, the reference's behaviour); transpose = 0 solves J d = B.
 * Returns 0, or k+1 for an exact zero pivot in column k (LAPACK's INFO). */
int or_solve6(const double Jac[36], double B[6], int transpose) {
  double a[6][6];
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) a[r][c] = transpose ? Jac[c * 6 + r] : Jac[r * 6 + c];
  int info = 0;
  for (int k = 0; k < 6; k++) {
    int p = k;
    double best = fabs(a[k][k]);
    for (int r = k + 1; r < 6; r++)
      if (fabs(a[r][k]) > best) {
        best = fabs(a[r][k]);
        p = r;
      }
    if (a[p][k] == 0.0) {
      if (!info) info = k + 1;
      continue;
    }
    if (p != k) {
      for (int c = 0; c < 6; c++) {
        double t = a[k][c];
        a[k][c] = a[p][c];
        a[p][c] = t;
      }
      double t = B[k];
      B[k] = B[p];
      B[p] = t;
    }
    for (int r = k + 1; r < 6; r++) {
      const double l = a[r][k] / a[k][k];
      a[r][k] = l;
      for (int c = k + 1; c < 6; c++) a[r][c] -= l * a[k][c];
      B[r] -= l * B[k];
    }
  }
  if (info) return info;
  for (int k = 5; k >= 0; k--) {
    double s = B[k];
    for (int c = k + 1; c < 6; c++) s -= a[k][c] * B[c];
    B[k] = s / a[k][k];
  }
  return 0;
}

int or_solve6_lapack_layout(const double Jac[36], double B[6]) { return or_solve6(Jac, B, 1); }

/* The FD Jacobian and residual of one Newton pass (Tusscher.cc:157-178).
 * On return the state is restored the way the reference restores it. */
void or_tt_newton_system(const double *C, double *S, const double *stateN,
                         const double *oldRates, double dt, double alg12,
                         double Jac[36], double B[6]) {
  double alg[OR_TT_NALG], rates[OR_TT_NSTATE], fxph[6], fxmh[6];
  memset(alg, 0, sizeof(alg));
  memset(rates, 0, sizeof(rates));
  alg[12] = alg12;
  for (int i = 0; i < 6; i++) {
    const double eps = 1e-3 * S[TT_NR[i]];
    const double twoeps = 2. * eps;
    S[TT_NR[i]] += eps;
    or_tt_rates(C, S, alg, rates);
    for (int j = 0; j < 6; j++) fxph[j] = S[TT_NR[j]] - stateN[j] - dt * rates[TT_NR[j]];
    S[TT_NR[i]] -= 2 * eps;
    or_tt_rates(C, S, alg, rates);
    for (int j = 0; j < 6; j++) fxmh[j] = S[TT_NR[j]] - stateN[j] - dt * rates[TT_NR[j]];
    S[TT_NR[i]] += eps;
    for (int j = 0; j < 6; j++) Jac[j * 6 + i] = (fxph[j] - fxmh[j]) / twoeps;
  }
  for (int i = 0; i < 6; i++) B[i] = S[TT_NR[i]] - stateN[i] - dt * oldRates[i];
}

/* Tusscher::Compute_Ion + UpdateStateVariables (Tusscher.cc:116-213):
 * backward-Euler gates, then Newton on the six concentrations with a central
 * FD Jacobian, stopping at ||dx||2 <= 1e-6.  jt = 1 solves each step with
 * J^T as the reference does (row-major array handed to column-major dgesv,
 * :172-181); that iteration leaves the physical range during a normal action
 * potential at dt = 0.02 ms (Ca_i < 0, then NaN), so jt = 0 (a true Newton
 * step towards the same backward-Euler fixed point) is the default of the
 * whole-step driver.  The reference has no iteration cap; this restatement
 * stops at 500 passes and reports the count. */
double or_tt_compute_ion(const double *C, double *S, double Cm, double dt,
                         double volt, double istim, int jt, int *newton_iters) {
  double alg[OR_TT_NALG], rates[OR_TT_NSTATE];
  memset(alg, 0, sizeof(alg));
  memset(rates, 0, sizeof(rates));
  const double Xi = 2000.;
  S[0] = volt;
  alg[12] = -istim / (Cm * Xi);
  double StateN[6], oldRates[6], fxph[6], fxmh[6], Jac[36], B[6];
  for (int i = 0; i < 6; i++) StateN[i] = S[TT_NR[i]];
  or_tt_rates(C, S, alg, rates);
  for (int i = 0; i < 12; i++) {
    const double b = -1. / alg[TT_DIND[i]];
    const double a = alg[i] / alg[TT_DIND[i]];
    S[TT_IND[i]] = (S[TT_IND[i]] + dt * a) / (1 - dt * b);
  }
  or_tt_rates(C, S, alg, rates);
  for (int i = 0; i < 6; i++) oldRates[i] = rates[TT_NR[i]];
  double resid = 100.;
  int it = 0;
  while (resid > 1e-6 && it < 500) {
    for (int i = 0; i < 6; i++) {
      const double eps = 1e-3 * S[TT_NR[i]];
      const double twoeps = 2. * eps;
      S[TT_NR[i]] += eps;
      or_tt_rates(C, S, alg, rates);
      for (int j = 0; j < 6; j++) fxph[j] = S[TT_NR[j]] - StateN[j] - dt * rates[TT_NR[j]];
      S[TT_NR[i]] -= 2 * eps;
      or_tt_rates(C, S, alg, rates);
      for (int j = 0; j < 6; j++) fxmh[j] = S[TT_NR[j]] - StateN[j] - dt * rates[TT_NR[j]];
      S[TT_NR[i]] += eps;
      for (int j = 0; j < 6; j++) Jac[j * 6 + i] = (fxph[j] - fxmh[j]) / twoeps;
    }
    for (int i = 0; i < 6; i++) B[i] = S[TT_NR[i]] - StateN[i] - dt * oldRates[i];
    or_solve6(Jac, B, jt);
    for (int i = 0; i < 6; i++) S[TT_NR[i]] -= B[i];
    or_tt_rates(C, S, alg, rates);
    for (int i = 0; i < 6; i++) oldRates[i] = rates[TT_NR[i]];
    resid = 0;
    for (int i = 0; i < 6; i++) resid += B[i] * B[i];
    resid = sqrt(resid);
    it++;
  }
  if (newton_iters) *newton_iters = it;
  return rates[0];
}

/* ================= Corrias rabbit Purkinje (IonicModel/Purkinje.cc) ========== */
/* SetUpPurkinjeParameters (IonicModel/SetUpIonicConstants.h:16-85). */
void or_pk_constants(double *C) {
  static const double k[57] = {
      310.0, 96485.3415, 8314.472, 2.0,    140,     5.4,    69.0,     13266.5, 0.6,
      0.2,   0.06,       700,      9000000, 500,    0.5,    -4320.0,  1.0,     0.9,
      0.7,   5.4,        5.0,      350.0,  10.0,    3.0,    0.005,    2.0,     1140.0,
      15.0,  2.0,        20.0,     1.5,    3.0,     3,      0.001,    0.001,   0.5,
      442.2, 0.01,       0.01,     0.0001, 5.0,     0.0001, 1.5,      0.188709677,
      0.045290323, 2500.0, 0.001, 120.0,  0.000246, 1.7,    1.6,      120.0,   0.000246,
      1.7,   1.6,        10.0,     5000.0};
  memcpy(C, k, sizeof(k));
  C[57] = C[1] / (C[2] * C[0]);
  C[58] = (C[2] * C[0]) / C[1];
  C[59] = C[9] * C[7];
  C[60] = C[8] * C[7];
  C[61] = (C[8] + C[9]) * C[7];
  C[62] = C[10] * C[7];
}

/* Purkinje constructor (Purkinje.h:44-50). */
void or_pk_init(double *s) {
  static const double v[OR_PK_NSTATE] = {
      -88.34, 0.00001, 0.000032, 0.17,     6.7,      140,      0.001337, 0.01,
      0.000003, 0.1,   0.7,      0.0,      0.7,      0.000007, 0.978861, 0.000012,
      0.864489, 0.25,  1.0,      0.0,      0.011099, 0.011099};
  memcpy(s, v, sizeof(v));
}

#define S state
#define A alg
#define R rates
/* Purkinje::ComputeRates (Purkinje.cc:5-104); alg[0] (stimulus, pA) is an
 * input. */
void or_pk_rates(const double *C, const double *S, double *A, double *R) {
  A[1] = 1.0 / (1.0 + exp((S[0] + 47.8) / -5.5));
  R[6] = (A[1] - S[6]) / C[16];
  A[3] = 1.0 / (1.0 + exp((S[0] + 14.6) / -5.5));
  R[8] = (A[3] - S[8]) / C[18];
  A[6] = 1.0 / (1.0 + exp((S[0] + -7.0) / -9.0));
  R[11] = (A[6] - S[11]) / C[20];
  A[7] = 1.0 / (1.0 + exp((S[0] + 27.5) / 8.0));
  R[12] = (A[7] - S[12]) / C[21];
  A[8] = 1.0 / (1.0 + exp((S[0] + 25.0) / -5.0));
  R[13] = (A[8] - S[13]) / C[24];
  A[9] = 1.0 / (1.0 + exp((S[0] + 69.0) / 3.96));
  R[14] = (A[9] - S[14]) / C[25];
  A[10] = 1.0 / (1.0 + exp((S[0] + 30.0) / -5.0));
  R[15] = (A[10] - S[15]) / C[27];
  A[2] = 1.0 / (1.0 + exp((S[0] + 67.9) / 3.87));
  A[17] = 1.42271 * exp(-0.05119 * S[0]);
  R[7] = (A[2] - S[7]) / A[17];
  A[4] = 1.0 / (1.0 + exp((S[0] + 31.0) / 5.54));
  A[18] = 25.1 / (0.04 + 0.7 * exp(-1.0 * pow(0.028 * (S[0] + 14.5), 2.0)));
  R[9] = (A[4] - S[9]) / A[18];
  A[5] = 0.4 + 0.6 / (1.0 + pow(S[1] / 0.0001, 2.0));
  A[19] = 2.0 + 80.0 / (1.0 + pow(S[1] / 0.0001, 2.0));
  R[10] = (A[5] - S[10]) / A[19];
  A[11] = 0.1 + 0.9 / (1.0 + exp((S[0] + 75.6) / 6.3));
  A[20] = 120.0 + 1.0 * exp((S[0] + 100.0) / 25.0);
  R[16] = (A[11] - S[16]) / A[20];
  A[13] = 1.0 / (1.0 + exp((S[0] - 1.5) / -16.7));
  A[22] = (fabs(S[0] + 30.0) < 0.0145
               ? 417.946
               : 1.0 / ((7.19e-05 * (S[0] + 30.0)) / (1.0 - exp(-0.148 * (S[0] + 30.0))) +
                        (0.000131 * (S[0] + 30.0)) / (exp(0.0687 * (S[0] + 30.0)) - 1.0)));
  R[18] = (A[13] - S[18]) / A[22];
  A[14] = 1.0 / (1.0 + exp((S[0] + 109.0) / 10.0));
  A[23] = 6000.0 / (exp(-1.0 * (2.9 + 0.04 * S[0])) + exp(1.0 * (3.6 + 0.11 * S[0])));
  R[20] = (A[14] - S[20]) / A[23];
  A[15] = 1.0 / (1.0 + exp((S[0] + 109.0) / 10.0));
  A[24] = 6000.0 / (exp(-1.0 * (2.9 + 0.04 * S[0])) + exp(1.0 * (3.6 + 0.11 * S[0])));
  R[21] = (A[15] - S[21]) / A[24];
  A[12] = 1.0 / (1.0 + exp((S[0] + 50.0) / -7.5));
  A[21] = (fabs(S[0] + 7.0) > 0.001 ? (0.00138 * 1.0 * (S[0] + 7.0)) / (1.0 - exp(-0.123 * (S[0] + 7.0)))
                                    : 0.00138 / 0.123);
  A[26] = (fabs(S[0] + 10.0) > 0.001 ? (6.1e-05 * 1.0 * (S[0] + 10.0)) / (exp(0.145 * (S[0] + 10.0)) - 1.0)
                                     : 0.00061 / 0.145);
  A[29] = 1.0 / (A[21] + A[26]);
  R[17] = (A[12] - S[17]) / A[29];
  A[27] = A[13];
  A[30] = 4.0 * A[22];
  R[19] = (A[27] - S[19]) / A[30];
  A[41] = 1.0 / (1.0 + exp((92.0 + S[0]) / 10.0));
  A[42] = C[58] * log(C[5] / S[5]);
  A[43] = C[29] * pow(C[5] / 5.4, 0.8) * A[41] * (S[0] - A[42]);
  A[32] = C[58] * log(C[5] / S[5]);
  A[33] = C[22] * S[11] * S[12] * (S[0] - A[32]);
  A[34] = 1.0 / (1.0 + exp((5.0 - S[0]) / 17.0));
  A[35] = C[58] * log(C[5] / S[5]);
  A[36] = C[23] * A[34] * (S[0] - A[35]);
  A[44] = 1.0 / (1.0 + exp((33.0 + S[0]) / 22.4));
  A[45] = C[58] * log(C[5] / S[5]);
  A[46] = C[30] * pow(C[5] / 5.4, 1.0) * A[44] * S[17] * (S[0] - A[45]);
  A[47] = C[58] * log(C[5] / S[5]);
  A[48] = C[31] * S[18] * S[19] * (S[0] - A[47]);
  A[53] = C[58] * log(C[5] / S[5]);
  A[54] = C[37] * (S[0] - A[53]);
  A[50] = 1.0 / (1.0 + exp((S[0] + 80.0) / -45.0));
  A[51] = 1.0 / (1.0 + exp((S[0] + 0.0) / 125.0));
  A[52] = C[36] * A[50] * A[51] * (1.0 / (1.0 + pow(1.9 / C[5], 1.45))) *
          (1.0 / (1.0 + pow(31.98 / S[4], 1.0)));
  A[60] = C[58] * log(C[5] / S[5]);
  A[62] = C[43] * S[20] * (S[0] - A[60]);
  R[5] = ((-1.0 * A[33] + -1.0 * A[36] + -1.0 * A[46] + -1.0 * A[48] + -1.0 * A[43] +
           -1.0 * A[54] + -1.0 * A[62] + -1.0 * A[0] + 2.0 * A[52]) * 1000.0) / (C[1] * C[61]);
  A[28] = 0.5 * C[58] * log(C[3] / S[1]);
  A[31] = C[19] * S[8] * S[9] * S[10] * (S[0] - A[28]);
  A[16] = 0.5 * C[58] * log(C[3] / S[1]);
  A[25] = C[17] * S[6] * S[7] * (S[0] - A[16]);
  A[37] = C[58] * log(C[4] / S[4]);
  A[38] = C[26] * S[13] * S[14] * (S[0] - A[37]);
  A[39] = C[58] * log(C[4] / S[4]);
  A[40] = C[28] * S[15] * S[16] * (S[0] - A[39]);
  A[59] = C[40] * (1.0 / (1.0 + pow(C[41] / S[1], C[42])));
  A[57] = 0.5 * C[58] * log(C[3] / S[1]);
  A[58] = C[39] * (S[0] - A[57]);
  A[64] = C[58] * log(C[4] / S[4]);
  A[66] = C[44] * S[21] * (S[0] - A[64]);
  A[49] = (512.009 * C[33] *
           (exp((C[35] * (C[32] - 2.0) * S[0]) / C[58]) * pow(S[4], C[32]) * C[3] -
            S[1] * exp(((C[35] - 1.0) * (C[32] - 2.0) * S[0]) / C[58]) * pow(C[4], C[32]))) /
          ((1.0 + C[34] * (S[1] * pow(C[4], C[32]) + C[3] * pow(S[4], C[32]))) *
           (1.0 + S[1] / 0.0069));
  A[55] = C[58] * log(C[4] / S[4]);
  A[56] = C[38] * (S[0] - A[55]);
  R[0] = ((-1.0 * 1.0) / C[6]) *
         (A[43] + A[33] + A[36] + A[46] + A[48] + A[54] + A[52] + A[31] + A[25] + A[38] +
          A[40] + A[59] + A[58] + A[66] + A[62] + A[49] + A[56] + A[0]);
  A[61] = C[45] * (S[3] / (1.0 + pow(C[46] / S[1], 2.0)));
  A[63] = C[47] * ((pow(S[1] / C[48], C[50]) - pow(S[3] / C[49], C[50])) /
                   (1.0 + pow(S[1] / C[48], C[50]) + pow(S[3] / C[49], C[50])));
  A[65] = C[51] * ((pow(S[2] / C[52], C[54]) - pow(S[3] / C[53], C[54])) /
                   (1.0 + pow(S[2] / C[52], C[54]) + pow(S[3] / C[53], C[54])));
  A[67] = C[55] * (S[3] - S[2]);
  R[3] = ((-1.0 * A[61] + -1.0 * A[67] + A[65] + A[63]) * 1000.0) / (2.0 * C[1] * C[62]);
  R[4] = ((-1.0 * A[38] + -1.0 * A[40] + -3.0 * A[52] + -3.0 * A[49] + -1.0 * A[66] +
           -1.0 * A[56]) * 1000.0) / (C[1] * C[61]);
  A[68] = C[56] * (S[1] - S[2]);
  R[1] = ((-1.0 * A[31] + -1.0 * A[25] + -1.0 * A[58] + 1.0 * A[61] + -1.0 * A[63] +
           -1.0 * A[68] + 2.0 * A[49] + -1.0 * A[59]) * 1000.0) / (2.0 * C[1] * C[59]);
  R[2] = ((-1.0 * A[65] + A[68] + A[67]) * 1000.0) / (2.0 * C[1] * C[60]);
}
#undef S
#undef A
#undef R

/* Purkinje::Compute_Ion + UpdateStateVariables (Purkinje.cc:179-210): forward
 * Euler on all 22 states (V included), with 10 sub-steps of dt/10 when
 * |dV/dt| > 1.5 mV/ms; the last evaluated dV/dt is returned. */
double or_pk_compute_ion(const double *C, double *S, double Cm, double dt, double volt,
                         double istim) {
  double alg[OR_PK_NALG], rates[OR_PK_NSTATE];
  memset(alg, 0, sizeof(alg));
  memset(rates, 0, sizeof(rates));
  const double Xi = 2500., area = 4.612E-5;
  (void)Cm;
  S[0] = volt;
  alg[0] = -1E6 * istim * area / Xi;
  or_pk_rates(C, S, alg, rates);
  if (fabs(rates[0]) > 1.5) {
    for (int i = 0; i < 10; i++) {
      or_pk_rates(C, S, alg, rates);
      for (int j = 0; j < OR_PK_NSTATE; j++) S[j] += rates[j] * dt / 10;
    }
  } else {
    for (int j = 0; j < OR_PK_NSTATE; j++) S[j] += rates[j] * dt;
  }
  return rates[0];
}
