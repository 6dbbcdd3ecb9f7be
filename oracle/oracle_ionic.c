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
/* LuoRudyGateTable::get_parameters (LRGateTable.h:93-110, the branch without
 * _USE_GATE_SLOPE): the 14 values linearly interpolated between the
 * make_parameter1 values of the 0.001 mV grid -90 + 0.001 i.  The grid values
 * are computed where needed (they are what makeGateTable stores). */
static void lr_gate_table(const lr_const *C, double v, double tab[14]) {
  const int index = (int)((v + 90) * 1000);
  double lo[14], hi[14];
  lr_table(C, -90 + 0.001 * index, lo);
  lr_table(C, -90 + 0.001 * (index + 1), hi);
  const double delVbydT = 1000 * (v - (-90 + 0.001 * index));
  for (int i = 0; i < 14; i++) tab[i] = (hi[i] - lo[i]) * delVbydT + lo[i];
}

static double lr_compute(double *in, double Cm, double dt, double volt, double istim, int gate) {
  const lr_const C = lr_constants();
  double tabz[14];
  in[7] = volt;
  /* LR.cc:71-79: the table inside (-90, 200) mV, the formulas outside. */
  if (gate && volt > -90 && volt < 200) lr_gate_table(&C, in[7], tabz);
  else lr_table(&C, in[7], tabz);
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

double or_lr_compute_ion(double *in, double Cm, double dt, double volt, double istim) {
  return lr_compute(in, Cm, dt, volt, istim, 0);
}

void or_lr_tab(double v, int gate, double tab[14]) {
  const lr_const C = lr_constants();
  if (gate && v > -90 && v < 200) lr_gate_table(&C, v, tab);
  else lr_table(&C, v, tab);
}

double or_lr_gate_compute_ion(double *in, double Cm, double dt, double volt, double istim) {
  return lr_compute(in, Cm, dt, volt, istim, 1);
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
    /* Tusscher.cc:158 takes 1e-3 x; at x = 0 that is 0/0 -- step 1e-12 then. */
    const double eps = S[TT_NR[i]] != 0.0 ? 1e-3 * S[TT_NR[i]] : 1e-12;
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
      /* Tusscher.cc:158 takes 1e-3 x; at x = 0 that is 0/0 -- step 1e-12 then. */
    const double eps = S[TT_NR[i]] != 0.0 ? 1e-3 * S[TT_NR[i]] : 1e-12;
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

/* ---- Mahajan 2008 (UCLA rabbit ventricle) ----------------------------------
 * IonicModel/Mahajan.cc and SetUpMahajanParameters (SetUpIonicConstants.h:
 * 159-251).  State order of getInternalParameters: V, m, h, j, xr, xs1, xs2,
 * c_p, c_s, c_i, c_j, c_j', I_r, and the L-type Ca channel Markov states
 * c1, c2, xi1ca, xi1ba, xi2ca, xi2ba, then Na_i, x_tof, y_tof, x_tos, y_tos
 * and the troponin-bound Ca of myoplasm and submembrane. */
enum {
  MJ_V, MJ_M, MJ_H, MJ_J, MJ_XR, MJ_XS1, MJ_XS2, MJ_CP, MJ_CS, MJ_CI, MJ_CJ, MJ_CJP, MJ_IR,
  MJ_C1, MJ_C2, MJ_XI1CA, MJ_XI1BA, MJ_XI2CA, MJ_XI2BA, MJ_NAI, MJ_XTOF, MJ_YTOF, MJ_XTOS,
  MJ_YTOS, MJ_TROPI, MJ_TROPS
};

/* Per tissue type (epi/myo/endo x apex/center/base, SetUpIonicConstants.h:
 * 175-186): G_tof and G_Ks. */
static const double MJ_GTOF[9] = {0.11, 0.11, 0.11, 0.11, 0.11, 0.11, 0.11 * 0.85, 0.11 * 0.85,
                                  0.11 * 0.85};
static const double MJ_GKS[9] = {1.9 * 0.32,  1.4 * 0.32,  1.0 * 0.32, 0.74 * 0.32, 0.52 * 0.32,
                                 0.35 * 0.32, 0.98 * 0.32, 0.7 * 0.32, 0.5 * 0.32};

void or_mj_constants(int type, double *c) {
  for (int k = 0; k < OR_MJ_NCONST; k++) c[k] = 0.0;
  c[0] = 136.0;   /* Na_o (mM) */
  c[1] = 140.0;   /* K_i */
  c[2] = 5.4;     /* K_o */
  c[3] = 1.8;     /* Ca_o */
  c[4] = 182.0;   /* g_Ca */
  c[5] = 1.0;     /* (unused I_Ca factor) */
  c[6] = 0.04;    /* g_tos */
  if (type >= 0 && type < 9) {
    c[7] = MJ_GTOF[type];
    c[10] = MJ_GKS[type];
  } else { /* the reference's fallback: epi-base with G_Ks 0.32 */
    c[7] = 0.11;
    c[10] = 0.32;
  }
  c[8] = 0.84;    /* g_NaCa */
  c[9] = 0.0125;  /* g_Kr */
  c[11] = 0.3;    /* g_K1 */
  c[12] = 1.5;    /* g_NaK */
  c[13] = 0.4;    /* v_up */
  c[14] = 4.0;    /* tau_s */
  c[15] = 12.0;   /* g_Na */
  c[16] = 30.0;   /* tau_r */
  c[17] = 100.0;  /* tau_a */
  c[18] = 11.3;   /* a_v */
  c[19] = 90.0;   /* c* */
  c[20] = 308.0;  /* T */
  c[21] = -80.0;  /* (unused V_c) */
  c[22] = 8.314;  /* R */
  c[23] = 96.485; /* F */
  c[24] = c[23] / (c[22] * c[20]);
  c[25] = 153.0;  /* q_max */
  c[26] = 50.0;   /* release threshold */
  c[27] = c[28] = c[29] = c[30] = 1.0; /* J_leak, J_d, J_trpn, J_up multipliers */
  c[31] = c[32] = c[33] = 0.0;         /* m, h, j leaks */
  c[34] = 1.0;                         /* tau_j multiplier */
}

/* The failing-heart constant set, SetUpMahajan_failParameters
 * (SetUpIonicConstants.h:254-315): the same model with slower uptake,
 * stronger exchange, weaker I_K1 / I_to / I_Ks, a lower release threshold
 * and leaky Na gates. */
void or_mj_fail_constants(int type, double *c) {
  static const double GTOF[9] = {0.12, 0.12, 0.12, 0.12, 0.12, 0.12, 0.12 * 0.77, 0.12 * 0.77,
                                 0.12 * 0.77};
  static const double GKS[9] = {1.5, 1.3, 1.1, 1.0, 0.9, 0.8, 1.1, 1.0, 0.9};
  or_mj_constants(type, c);
  c[6] = 0.04 * 0.65;
  c[7] = (type >= 0 && type < 9 ? GTOF[type] : 0.12) * 0.67;
  c[10] = (type >= 0 && type < 9 ? GKS[type] : 1.0) * 0.1386 * 0.6;
  c[8] = 0.84 * 2.0;
  c[11] = 0.3 * 0.5;
  c[13] = 0.27;
  c[14] = 6.0;
  c[17] = 200.0;
  c[18] = 22.5;
  c[19] = 50.0;
  c[26] = 10.0;
  c[31] = c[32] = c[33] = 0.01;
}

/* Initial states by tissue type (Mahajan.cc:27-158; out-of-range types take
 * the reference's default row, equal to type 2). */
static const double MJ_INIT[9][OR_MJ_NSTATE] = {
    {-81.4359, 0.002789, 0.96679, 0.97859, 0.014994, 0.044487, 0.091992, 2.1933, 0.23277,
     0.24868, 101.6418, 95.4278, 0.009495, 3.7e-005, 0.96094, 0.003718, 6.7e-005, 0.010571,
     0.024659, 11.5909, 0.00533, 0.99047, 0.005333, 0.17488, 21.5955, 20.0206},
    {-81.3846, 0.002813, 0.96642, 0.97829, 0.015109, 0.04988, 0.10153, 2.3513, 0.23616, 0.2529,
     102.5325, 96.0961, 0.010394, 3.7e-005, 0.95815, 0.004448, 6.7e-005, 0.01261, 0.024683,
     11.442, 0.005349, 0.99012, 0.005352, 0.17323, 21.8749, 20.2219},
    {-81.3238, 0.002841, 0.96596, 0.97791, 0.015254, 0.056799, 0.11328, 2.5207, 0.23954,
     0.25736, 103.4612, 96.8014, 0.011385, 3.7e-005, 0.95449, 0.005413, 6.7e-005, 0.015284,
     0.024706, 11.2566, 0.00537, 0.98962, 0.005373, 0.17158, 22.1689, 20.4268},
    {-81.2657, 0.002868, 0.96553, 0.97752, 0.015406, 0.064019, 0.12504, 2.6684, 0.24235,
     0.26131, 104.2505, 97.4054, 0.01227, 3.7e-005, 0.95055, 0.006458, 6.8e-005, 0.018158,
     0.024723, 11.0791, 0.005391, 0.98902, 0.005394, 0.17018, 22.4276, 20.6034},
    {-81.1929, 0.002903, 0.96497, 0.97694, 0.015625, 0.073994, 0.14059, 2.8407, 0.2456,
     0.26605, 105.1459, 98.097, 0.013328, 3.7e-005, 0.94486, 0.00798, 6.8e-005, 0.022313,
     0.024742, 10.8644, 0.005417, 0.98806, 0.005421, 0.16859, 22.7359, 20.8156},
    {-81.1078, 0.002944, 0.96431, 0.97609, 0.01595, 0.087401, 0.1603, 3.0189, 0.24905,
     0.27153, 106.0478, 98.7911, 0.014455, 3.8e-005, 0.93663, 0.010192, 6.8e-005, 0.028302,
     0.024766, 10.6216, 0.005448, 0.98648, 0.005452, 0.16682, 23.0895, 21.0607},
    {-81.2626, 0.00287, 0.9655, 0.97756, 0.015387, 0.061634, 0.12278, 2.7185, 0.24183, 0.2588,
     104.2029, 97.6338, 0.01267, 3.7e-005, 0.95198, 0.006047, 6.8e-005, 0.017057, 0.024804,
     11.0783, 0.005392, 0.98941, 0.005395, 0.16991, 22.2542, 20.5438},
    {-81.1873, 0.002905, 0.96493, 0.97701, 0.015593, 0.071141, 0.13799, 2.9031, 0.24508,
     0.26334, 105.1225, 98.3761, 0.013821, 3.7e-005, 0.94672, 0.007447, 6.8e-005, 0.020885,
     0.024838, 10.8496, 0.005419, 0.9886, 0.005422, 0.16858, 22.5504, 20.7522},
    {-81.1085, 0.002943, 0.96432, 0.97631, 0.015852, 0.082505, 0.15518, 3.0765, 0.2482,
     0.26807, 105.9653, 99.0544, 0.01493, 3.8e-005, 0.94008, 0.009226, 6.8e-005, 0.025709,
     0.024873, 10.6197, 0.005448, 0.98743, 0.005451, 0.16728, 22.8568, 20.9674}};

void or_mj_init(int type, double *s) {
  const double *row = MJ_INIT[(type >= 0 && type < 9) ? type : 2];
  for (int k = 0; k < OR_MJ_NSTATE; k++) s[k] = row[k];
}

double or_mj_xi(void) { return 1.55e-4 / 2.58e-8; } /* Mahajan.cc:164 */

/* Exact exponential relaxation over h toward x_inf with time constant tau,
 * written as a rate so that the caller's forward-Euler update lands on it
 * (Mahajan.cc:271-273 and the other gates). */
static double mj_relax(double xinf, double tau, double x, double h) {
  return (xinf - (xinf - x) * exp(-h / tau) - x) / h;
}

/* Instantaneous buffering factor of calmodulin, SR, membrane and
 * sarcolemmal sites (Mahajan.cc:540-555). */
static double mj_buffer(double c) {
  const double b_cal = 24.0 * 7.0 / ((7.0 + c) * (7.0 + c));
  const double b_sr = 47.0 * 0.6 / ((0.6 + c) * (0.6 + c));
  const double b_mem = 15.0 * 0.3 / ((0.3 + c) * (0.3 + c));
  const double b_sar = 42.0 * 13.0 / ((13.0 + c) * (13.0 + c));
  return 1.0 / (1.0 + b_cal + b_sr + b_mem + b_sar);
}

/* ucla_rhsfun (Mahajan.cc:168-230) with the comp_* currents (:235-603):
 * every rate of S at gate step h; istim in uA/uF.  Returns nothing; dV/dt is
 * R[0]. */
void or_mj_rates(const double *c, const double *S, double h, double istim, double *R) {
  const double V = S[MJ_V], frt = c[24];
  const double nao = c[0], ki = c[1], ko = c[2], cao = c[3];
  const double nai = S[MJ_NAI], cs = S[MJ_CS], ci = S[MJ_CI], cj = S[MJ_CJ];

  /* I_K1, reversal from P_K : P_Na = 100 : 1 (:331-347) */
  const double ek1 = log((100.0 * ko + nao) / (100.0 * ki + nai)) / frt;
  const double a_k1 = 1.02 / (1.0 + exp(0.2385 * (V - ek1 - 59.215)));
  const double b_k1 = (0.49124 * exp(0.08032 * (V - ek1 + 5.476)) +
                       exp(0.061750 * (V - ek1 - 594.31))) /
                      (1.0 + exp(-0.5143 * (V - ek1 + 4.753)));
  const double i_k1 = c[11] * sqrt(ko / 5.4) * a_k1 / (a_k1 + b_k1) * (V - ek1);

  /* I_to, slow and fast (:352-385) */
  const double ek = log(ko / ki) / frt;
  const double x_inf = 1.0 / (1.0 + exp(-(V + 3.0) / 15.0));
  const double y_inf = 1.0 / (1.0 + exp((V + 33.5) / 10.0));
  const double tau_xs = 9.0 / (1.0 + exp((V + 3.0) / 15.0)) + 0.5;
  const double tau_ys = 3000.0 / (1.0 + exp((V + 60.0) / 10.0)) + 30.0;
  const double i_tos = c[6] * S[MJ_XTOS] * (S[MJ_YTOS] + 0.5 * y_inf) * (V - ek);
  R[MJ_XTOS] = mj_relax(x_inf, tau_xs, S[MJ_XTOS], h);
  R[MJ_YTOS] = mj_relax(y_inf, tau_ys, S[MJ_YTOS], h);
  const double tau_xf = 3.5 * exp(-(V / 30.0) * (V / 30.0)) + 1.5;
  const double tau_yf = 20.0 / (1.0 + exp((V + 33.5) / 10.0)) + 20.0;
  const double i_tof = c[7] * S[MJ_XTOF] * S[MJ_YTOF] * (V - ek);
  R[MJ_XTOF] = mj_relax(x_inf, tau_xf, S[MJ_XTOF], h);
  R[MJ_YTOF] = mj_relax(y_inf, tau_yf, S[MJ_YTOF], h);

  /* I_NaK (:390-398) */
  const double sigma = (exp(nao / 67.3) - 1.0) / 7.0;
  const double f_nak = 1.0 / (1.0 + 0.1245 * exp(-0.1 * V * frt) + 0.0365 * sigma * exp(-V * frt));
  const double i_nak = c[12] * f_nak / (1.0 + 12.0 / nai) * ko / (ko + 1.5);

  /* Na/Ca exchange flux (uM/ms), concentrations in mM (:403-425) */
  const double csm = cs / 1000.0;
  const double nai3 = nai * nai * nai, nao3 = nao * nao * nao;
  const double e_in = exp(0.35 * V * frt), e_out = exp((0.35 - 1.0) * V * frt);
  const double zw3 = nai3 * cao * e_in - nao3 * csm * e_out;
  const double zw4 = 1.0 + 0.2 * e_out;
  const double r_loss = 0.3 / cs;
  const double aloss = 1.0 / (1.0 + r_loss * r_loss * r_loss);
  const double yz1 = 1.3 * nai3 + 87.5 * 87.5 * 87.5 * csm;
  const double yz2 = 12.3 * 12.3 * 12.3 * cao * (1.0 + csm / 0.0036);
  const double rn = nai / 12.3;
  const double yz3 = 0.0036 * nao3 * (1.0 + rn * rn * rn);
  const double yz4 = nai3 * cao + nao3 * csm;
  const double j_naca = c[8] * aloss * zw3 / (zw4 * (yz1 + yz2 + yz3 + yz4));

  /* Ca cycling (:180-200) */
  const double j_d = c[28] * (cs - ci) / c[14];
  R[MJ_TROPI] = c[29] * (0.0327 * ci * (70.0 - S[MJ_TROPI]) - 0.0196 * S[MJ_TROPI]);
  R[MJ_TROPS] = 0.0327 * cs * (70.0 - S[MJ_TROPS]) - 0.0196 * S[MJ_TROPS];
  const double j_up = c[30] * c[13] * ci * ci / (ci * ci + 0.25);
  const double j_leak = c[27] * 2.069e-5 * cj * cj / (cj * cj + 2500.0) * (16.667 * cj - ci);

  /* L-type Ca channel, 7-state Markov model (:451-516) */
  const double po_inf = 1.0 / (1.0 + exp(-V / 8.0));
  const double alpha = po_inf, beta = 1.0 - po_inf;
  const double rc = 3.0 / S[MJ_CP];
  const double fca = 1.0 / (1.0 + rc * rc * rc);
  const double s1 = 0.0182688 * fca, s1t = 0.00195;
  const double k1 = 0.024168 * fca, k2 = 1.03615e-4, k1t = 0.00413, k2t = 0.00224;
  const double r1 = 0.30, r2 = 3.0;
  const double s2 = s1 * (r1 / r2) * (k2 / k1);
  const double s2t = s1t * (r1 / r2) * (k2t / k1t);
  const double poi = 1.0 / (1.0 + exp(-(V + 40.0) / 3.0));
  const double k3 = (1.0 - poi) / 3.0, k3t = k3;
  const double Pr = 1.0 - 1.0 / (1.0 + exp(-(V + 40.0) / 4.0));
  const double recov = 10.0 + 4954.0 * exp(V / 15.6);
  const double rcp = S[MJ_CP] / 6.09365;
  const double tau_ca0 = 78.0329 / (1.0 + rcp * rcp * rcp * rcp) + 0.1;
  const double tau_ca = (recov - tau_ca0) * Pr + tau_ca0;
  const double tau_ba = (recov - 450.0) * Pr + 450.0;
  const double Ps = 1.0 / (1.0 + exp(-(V + 40.0) / 11.32));
  const double k6 = fca * Ps / tau_ca, k5 = (1.0 - Ps) / tau_ca;
  const double k6t = Ps / tau_ba, k5t = (1.0 - Ps) / tau_ba;
  const double k4 = k3 * (alpha / beta) * (k1 / k2) * (k5 / k6);
  const double k4t = k3t * (alpha / beta) * (k1t / k2t) * (k5t / k6t);
  const double c1 = S[MJ_C1], c2 = S[MJ_C2];
  const double i1c = S[MJ_XI1CA], i1b = S[MJ_XI1BA], i2c = S[MJ_XI2CA], i2b = S[MJ_XI2BA];
  const double po = 1.0 - i1c - i2c - i1b - i2b - c1 - c2;
  R[MJ_C2] = beta * c1 + k5 * i2c + k5t * i2b - (k6 + k6t + alpha) * c2;
  R[MJ_C1] = alpha * c2 + k2 * i1c + k2t * i1b + r2 * po - (beta + r1 + k1t + k1) * c1;
  R[MJ_XI1CA] = k1 * c1 + k4 * i2c + s1 * po - (k3 + k2 + s2) * i1c;
  R[MJ_XI2CA] = k3 * i1c + k6 * c2 - (k5 + k4) * i2c;
  R[MJ_XI1BA] = k1t * c1 + k4t * i2b + s1t * po - (k3t + k2t + s2t) * i1b;
  R[MJ_XI2BA] = k3t * i1b + k6t * c2 - (k5t + k4t) * i2b;

  /* Ca driving force (:430-447) */
  const double za = 2.0 * V * frt;
  const double f1 = 4.0 * 0.00054 * c[23] * c[23] / (c[22] * c[20]);
  double rxa;
  if (fabs(za) < 0.001) rxa = f1 * (csm * exp(za) - 0.341 * cao) / (2.0 * frt);
  else rxa = V * f1 * (csm * exp(za) - 0.341 * cao) / (exp(za) - 1.0);
  const double j_ca = c[4] * po * rxa;

  R[MJ_CS] = mj_buffer(cs) * (50.0 * (S[MJ_IR] - j_d - j_ca + j_naca) - R[MJ_TROPS]);
  R[MJ_CI] = mj_buffer(ci) * (j_d - j_up + j_leak - R[MJ_TROPI]);
  R[MJ_CJ] = -S[MJ_IR] + j_up - j_leak;
  R[MJ_CJP] = (cj - S[MJ_CJP]) / c[17];

  /* SR release: load dependence (:560-578), sparks (:582-589), dyad (:594-603) */
  double q = 0.0;
  const double cjp = S[MJ_CJP];
  if (cjp > c[26] && cjp < c[19]) q = cjp - c[26];
  else if (cjp >= c[19]) q = c[18] * cjp + (c[19] - c[26]) - c[18] * c[19];
  if (q >= c[25]) q = c[25];
  const double Qr = cj * q / c[19];
  const double ey = exp(-0.05 * (V + 30.0));
  const double spark = 2.58079 * po * fabs(rxa) * ey / (1.0 + ey);
  R[MJ_IR] = spark * Qr - S[MJ_IR] * (1.0 - c[16] * R[MJ_CJ] / cj) / c[16];
  const double ex = exp(-0.3576 * (V + 30.0));
  const double g_sr = 26841.8 * ex / (1.0 + ex);
  R[MJ_CP] = po * Qr * fabs(rxa) * g_sr + po * 9000.0 * fabs(rxa) - (S[MJ_CP] - cs) / 0.5;

  /* I_Na, Hund-Rudy/Livshitz gates (:235-287) */
  const double ena = log(nao / nai) / frt;
  double am = 3.2;
  if (fabs(V + 47.13) > 0.001) am = 0.32 * (V + 47.13) / (1.0 - exp(-0.1 * (V + 47.13)));
  const double bm = 0.08 * exp(-V / 11.0);
  const double a = 1.0 - 1.0 / (1.0 + exp(-(V + 40.0) / 0.24));
  const double ah = a * 0.135 * exp((80.0 + V) / (-6.8));
  const double bh = (1.0 - a) / (0.13 * (1.0 + exp((V + 10.66) / (-11.1)))) +
                    a * (3.56 * exp(0.079 * V) + 3.1e5 * exp(0.35 * V));
  const double aj = a * (-1.2714e5 * exp(0.2444 * V) - 3.474e-5 * exp(-0.04391 * V)) *
                    (V + 37.78) / (1.0 + exp(0.311 * (V + 79.23)));
  const double bj = (1.0 - a) * (0.3 * exp(-2.535e-7 * V) / (1.0 + exp(-0.1 * (V + 32.0)))) +
                    a * (0.1212 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14))));
  const double tau_h = 1.0 / (ah + bh), tau_j = c[34] / (aj + bj), tau_m = 1.0 / (am + bm);
  const double mg = (S[MJ_M] + c[31]) / (1.0 + c[31]);
  const double i_na = c[15] * (S[MJ_H] + c[32]) / (1.0 + c[32]) * (S[MJ_J] + c[33]) / (1.0 + c[33]) *
                      mg * mg * mg * (V - ena);
  R[MJ_H] = mj_relax(ah * tau_h, tau_h, S[MJ_H], h);
  R[MJ_J] = mj_relax(aj * tau_j, tau_j, S[MJ_J], h);
  R[MJ_M] = mj_relax(am * tau_m, tau_m, S[MJ_M], h);

  /* I_Kr (:292-313) */
  double kr1 = 0.00138 / 0.123, kr2 = 0.00061 / 0.145;
  if (fabs(V + 7.0) > 0.001) kr1 = 0.00138 * (V + 7.0) / (1.0 - exp(-0.123 * (V + 7.0)));
  if (fabs(V + 10.0) > 0.001) kr2 = 0.00061 * (V + 10.0) / (exp(0.145 * (V + 10.0)) - 1.0);
  const double xr_inf = 1.0 / (1.0 + exp(-(V + 50.0) / 7.5));
  const double rg = 1.0 / (1.0 + exp((V + 33.0) / 22.4));
  const double i_kr = c[9] * sqrt(ko / 5.4) * S[MJ_XR] * rg * (V - ek);
  R[MJ_XR] = mj_relax(xr_inf, 1.0 / (kr1 + kr2), S[MJ_XR], h);

  /* I_Ks with Ca dependence (:318-345) */
  const double eks = log((ko + 0.018330 * nao) / (ki + 0.018330 * nai)) / frt;
  const double xs_inf = 1.0 / (1.0 + exp(-(V - 1.50) / 16.70));
  double tau_xs1;
  if (fabs(V + 30.0) < 0.001 / 0.0687)
    tau_xs1 = 1.0 / (0.0000719 / 0.148 + 0.000131 / 0.0687);
  else
    tau_xs1 = 1.0 / (0.0000719 * (V + 30.0) / (1.0 - exp(-0.148 * (V + 30.0))) +
                     0.000131 * (V + 30.0) / (exp(0.0687 * (V + 30.0)) - 1.0));
  const double rks = 0.5 / ci;
  const double g_ksx = 1.0 + 0.8 / (1.0 + rks * rks * rks);
  const double i_ks = c[10] * g_ksx * S[MJ_XS1] * S[MJ_XS2] * (V - eks);
  R[MJ_XS1] = mj_relax(xs_inf, tau_xs1, S[MJ_XS1], h);
  R[MJ_XS2] = mj_relax(xs_inf, 4.0 * tau_xs1, S[MJ_XS2], h);

  /* Na_i and V (:205-229); wca converts uM/ms to uA/uF */
  const double wca = 8.0313;
  R[MJ_NAI] = -(1.0 / wca / 1000.0) * (i_na + 3.0 * i_nak + 3.0 * wca * j_naca);
  R[MJ_V] = -(i_na + i_k1 + i_kr + i_ks + i_tos + i_tof + wca * j_naca + 2.0 * wca * j_ca + i_nak) +
            istim;
}

/* Mahajan::Compute_Ion + UpdateStateVariables (Mahajan.cc:607-644): forward
 * Euler over dt, or 10 sub-steps of dt/10 when |dV/dt| > 20 (int_abs: the
 * threshold as abs() of an int reads it, |dV/dt| >= 21); returns the last
 * dV/dt evaluated. */
double or_mj_compute_ion(const double *c, double *S, double Cm, double dt, double V,
                         double istim, int int_abs) {
  double R[OR_MJ_NSTATE];
  S[MJ_V] = V;
  const double ist = istim / (Cm * or_mj_xi());
  or_mj_rates(c, S, dt, ist, R);
  const int adapt = int_abs ? (fabs(trunc(R[MJ_V])) > 20.0) : (fabs(R[MJ_V]) > 20.0);
  if (adapt) {
    const double h = dt / 10.0;
    for (int q = 0; q < 10; q++) {
      or_mj_rates(c, S, h, ist, R);
      for (int k = 0; k < OR_MJ_NSTATE; k++) S[k] += R[k] * h;
    }
  } else {
    for (int k = 0; k < OR_MJ_NSTATE; k++) S[k] += R[k] * dt;
  }
  return R[MJ_V];
}
