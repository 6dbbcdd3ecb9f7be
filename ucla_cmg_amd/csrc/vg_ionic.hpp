// vg_ionic.hpp -- per-node ionic models for gfx950 (device code).
//
// Luo-Rudy (IonicModel/LR.cc) and ten Tusscher 2006 (IonicModel/Tusscher.cc,
// constants IonicModel/SetUpIonicConstants.h:88-147) written as register-
// resident device functions over one node's state.  T is float or double;
// FAST selects hardware exp/log approximations (used only with float).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace vg {

// The per-node step functions are forced inline: as calls (the inliner
// declines them once a kernel has two call sites of a large body) their state
// arrays would live in scratch memory and every sub-step would save and
// restore the callee's registers there.
#define VG_INLINE inline __attribute__((always_inline))

// ---- math shims --------------------------------------------------------------
template <class T, bool FAST> struct Mth;
template <bool FAST> struct Mth<double, FAST> {
  __host__ __device__ static double ex(double x) { return exp(x); }
  __host__ __device__ static double lg(double x) { return log(x); }
  __host__ __device__ static double sq(double x) { return sqrt(x); }
  __device__ static double pw(double x, double y) { return pow(x, y); }
  __device__ static double em1(double x) { return expm1(x); }
};
template <> struct Mth<float, true> {
  __device__ static float ex(float x) { return __expf(x); }
  __device__ static float lg(float x) { return __logf(x); }
  __device__ static float sq(float x) { return __fsqrt_rn(x); }
  __device__ static float pw(float x, float y) { return __powf(x, y); }
  __device__ static float em1(float x) { return expm1f(x); }
};
template <> struct Mth<float, false> {
  __device__ static float ex(float x) { return expf(x); }
  __device__ static float lg(float x) { return logf(x); }
  __device__ static float sq(float x) { return sqrtf(x); }
  __device__ static float pw(float x, float y) { return powf(x, y); }
  __device__ static float em1(float x) { return expm1f(x); }
};

// A float whose division is the hardware reciprocal (1 ulp) times the
// dividend, for the fast FP32 paths: the rate expressions hold ~100 divisions
// per node, and an IEEE-exact float division is a ten-instruction sequence
// on gfx950 (scale, reciprocal, Newton steps, fix-up).  A constant divisor is
// folded into an exact constant reciprocal.  Only the arithmetic the rate
// code uses is defined.
struct RcpF {
  float v;
  RcpF() = default;
  __device__ constexpr RcpF(float x) : v(x) {}
  __device__ constexpr RcpF(double x) : v((float)x) {}
  __device__ constexpr RcpF(int x) : v((float)x) {}
  __device__ constexpr explicit operator float() const { return v; }
  // A hidden friend, found only for RcpF arguments (a namespace-scope
  // overload would hide ::fabs from the double and float code here).
  friend __device__ inline RcpF fabs(RcpF a) { return RcpF(fabsf(a.v)); }
};
__device__ inline RcpF operator+(RcpF a, RcpF b) { return RcpF(a.v + b.v); }
__device__ inline RcpF operator-(RcpF a, RcpF b) { return RcpF(a.v - b.v); }
__device__ inline RcpF operator*(RcpF a, RcpF b) { return RcpF(a.v * b.v); }
__device__ inline RcpF operator-(RcpF a) { return RcpF(-a.v); }
__device__ inline RcpF operator/(RcpF a, RcpF b) {
  if (__builtin_constant_p(b.v)) return RcpF(a.v * (1.0f / b.v));
  return RcpF(a.v * __builtin_amdgcn_rcpf(b.v));
}
__device__ inline RcpF &operator+=(RcpF &a, RcpF b) { return a = a + b; }
__device__ inline bool operator<(RcpF a, RcpF b) { return a.v < b.v; }
__device__ inline bool operator>(RcpF a, RcpF b) { return a.v > b.v; }
__device__ inline bool operator<=(RcpF a, RcpF b) { return a.v <= b.v; }
__device__ inline bool operator>=(RcpF a, RcpF b) { return a.v >= b.v; }
template <bool FAST> struct Mth<RcpF, FAST> {
  __device__ static RcpF ex(RcpF x) { return RcpF(__expf(x.v)); }
  __device__ static RcpF lg(RcpF x) { return RcpF(__logf(x.v)); }
  __device__ static RcpF sq(RcpF x) { return RcpF(__fsqrt_rn(x.v)); }
  __device__ static RcpF pw(RcpF x, RcpF y) { return RcpF(__powf(x.v, y.v)); }
  __device__ static RcpF em1(RcpF x) { return RcpF(expm1f(x.v)); }
};
// The arithmetic type of a model's rate code: RcpF for the fast FP32 paths.
template <class T, bool FAST> struct RateT { using type = T; };
template <> struct RateT<float, true> { using type = RcpF; };

// ======================= Luo-Rudy 1991 ========================================
// State: 0 m, 1 h, 2 j, 3 d, 4 f, 5 X, 6 Ca_i, 7 V (LR.cc:18-26 with the gate
// order of make_parameter1 :166-181).
constexpr int LR_NSTATE = 8;
constexpr double LR_INIT[LR_NSTATE] = {.001, 0.0001, 0.99, .001, 0.9, .001, 0.0000001, 0};

// make_parameter1 (LR.cc:102-183; LRGateTable.h:114-190): at v, the 14
// values (inf, alpha + beta) of m, h, j, d, f, X, then x1 and
// I_K1 + I_Kp + I_b.
template <class T, bool FAST>
__host__ __device__ inline void lr_tab(T v, T tab[14]) {
  using M = Mth<T, FAST>;
  const T c_K0 = T(5.4), c_Ki = T(145.0), RT_F = T(26.73);
  const T E_K1 = RT_F * M::lg(c_K0 / c_Ki);
  const T G_K1 = T(.6047) * M::sq(c_K0 / T(5.4));
  T alpha_m, beta_m, alpha_h, beta_h, alpha_j, beta_j;
  if (fabs(v + T(47.13)) < T(0.001)) alpha_m = T(3.2);
  else alpha_m = T(0.32) * (v + T(47.13)) / (T(1.0) - M::ex(T(-0.1) * (v + T(47.13))));
  beta_m = T(0.08) * M::ex(-v / T(11.));
  if (v > T(-40.)) {
    alpha_h = T(0);
    alpha_j = T(0);
    beta_h = T(1.) / (T(0.13) * (T(1) + M::ex(-(v + T(10.66)) / T(11.1))));
    beta_j = T(0.3) * M::ex(T(-2.535e-7) * v) / (T(1) + M::ex(T(-0.1) * (v + T(32))));
  } else {
    alpha_h = T(0.135) * M::ex(-(T(80) + v) / T(6.8));
    alpha_j = (T(-1.2714e5) * M::ex(T(0.2444) * v) - T(3.474e-5) * M::ex(T(-0.04391) * v)) *
              (v + T(37.78)) / (T(1) + M::ex(T(0.311) * (v + T(79.23))));
    beta_h = T(3.56) * M::ex(T(0.079) * v) + T(3.1e5) * M::ex(T(0.35) * v);
    beta_j = T(0.1212) * M::ex(T(-0.01052) * v) / (T(1) + M::ex(T(-0.1378) * (v + T(40.14))));
  }
  const T alpha_d = T(.095) * M::ex(T(-.01) * (v - T(5))) / (T(1) + M::ex(T(-.072) * (v - T(5))));
  const T beta_d = T(.07) * M::ex(T(-.017) * (v + T(44))) / (T(1) + M::ex(T(.05) * (v + T(44))));
  const T alpha_f = T(.012) * M::ex(T(-.008) * (v + T(28))) / (T(1) + M::ex(T(.15) * (v + T(28))));
  const T beta_f = T(.0065) * M::ex(T(-.02) * (v + T(30))) / (T(1) + M::ex(T(-.2) * (v + T(30))));
  const T alpha_x = T(.0005) * M::ex(T(.083) * (v + T(50))) / (T(1) + M::ex(T(.057) * (v + T(50))));
  const T beta_x = T(.0013) * M::ex(T(-.06) * (v + T(20))) / (T(1) + M::ex(T(-.04) * (v + T(20))));
  T x1;
  if (v > T(-100)) {
    if (fabs(v + T(77.)) < T(0.001)) x1 = T(0.608889);
    else x1 = T(2.837) * (M::ex(T(0.04) * (v + T(77))) - T(1)) / ((v + T(77)) * M::ex(T(0.04) * (v + T(35))));
  } else {
    x1 = T(1.);
  }
  const T alpha_k1 = T(1.02) / (T(1) + M::ex(T(0.2385) * (v - E_K1 - T(59.215))));
  const T beta1 = T(0.49124) * M::ex(T(0.08032) * (v - E_K1 + T(5.476)));
  const T beta2 = M::ex(T(0.06175) * (v - E_K1 - T(594.31)));
  const T beta3 = T(1) + M::ex(T(-0.5143) * (v - E_K1 + T(4.753)));
  const T beta_k1 = (beta1 + beta2) / beta3;
  const T k1_inf = alpha_k1 / (alpha_k1 + beta_k1);
  const T I_k1 = G_K1 * k1_inf * (v - E_K1);
  const T kp = T(1.) / (T(1) + M::ex((T(7.488) - v) / T(5.98)));
  const T I_kp = T(0.0183) * kp * (v - E_K1);
  const T I_b = T(0.0392) * (v + T(59.87));
  const T a[6] = {alpha_m, alpha_h, alpha_j, alpha_d, alpha_f, alpha_x};
  const T b[6] = {beta_m, beta_h, beta_j, beta_d, beta_f, beta_x};
  for (int i = 0; i < 6; i++) {
    tab[2 * i] = a[i] / (a[i] + b[i]);
    tab[2 * i + 1] = a[i] + b[i];
  }
  tab[12] = x1;
  tab[13] = I_k1 + I_kp + I_b;
}

// Rows of LuoRudyGateTable (LRGateTable.h:67-76): make_parameter1 at
// -90 + 0.001 i, i < LR_TABLE_ROWS.
constexpr int LR_TABLE_ROWS = 290001;

// LuoRudy::Compute_Ion (LR.cc:65-99), evaluated in T.  With a gate table
// (LUORUDYWITHGATE) the 14 values inside (-90, 200) mV are interpolated
// linearly from it (LuoRudyGateTable::get_parameters, LRGateTable.h:93-110).
template <class T, bool FAST>
__device__ VG_INLINE T lr_compute_ion(T s[LR_NSTATE], T Cm, T dt, T v, T istim, T gna_scale,
                                   T gsi_scale, const T *__restrict__ table = nullptr) {
  using M = Mth<T, FAST>;
  const T c_K0 = T(5.4);
  const T Gna = T(12.00) * gna_scale;
  const T Gsi = T(0.07) * gsi_scale;
  const T GK = T(0.705) * M::sq(c_K0 / T(5.4));
  const T r = T(11e-4), L = T(100E-4);
  const T Xi = T(2.) * (r + L) / (r * L);
  T tab[14];
  if (table && v > T(-90) && v < T(200)) {
    const int index = (int)((v + T(90)) * T(1000));
    const T d = T(1000) * (v - (T(-90) + T(0.001) * T(index)));
    const T *lo = table + (int64_t)index * 14;
#pragma unroll
    for (int i = 0; i < 14; i++) tab[i] = (lo[14 + i] - lo[i]) * d + lo[i];
  } else {
    lr_tab<T, FAST>(v, tab);
  }
  // Rush-Larsen with the rate sums alpha+beta (LR.cc:83-85).
#pragma unroll
  for (int i = 0; i < 6; i++) s[i] = tab[2 * i] - (tab[2 * i] - s[i]) * M::ex(-dt * tab[2 * i + 1]);
  s[7] = v;
  const T Es = T(7.7) - T(13.0287) * M::lg(s[6]);
  const T i_na = Gna * s[0] * s[0] * s[0] * s[1] * s[2] * (v - T(54.4));
  const T i_si = Gsi * s[3] * s[4] * (v - Es);
  const T i_k = GK * tab[12] * s[5] * (v + T(77.0));
  s[6] = s[6] + (T(-.0001) * i_si + T(0.07) * (T(0.0001) - s[6])) * dt;
  const T i_tot = -(tab[13] + i_k + i_na + i_si);
  return i_tot + istim / (Xi * Cm);
}

// ======================= ten Tusscher 2006 =====================================
// State (CellML order used by Tusscher.cc): 0 V, 1 K_i, 2 Na_i, 3 Ca_i, 4 Xr1,
// 5 Xr2, 6 Xs, 7 m, 8 h, 9 j, 10 Ca_ss, 11 d, 12 f, 13 f2, 14 fCass, 15 s,
// 16 r, 17 Ca_SR, 18 R'.
constexpr int TT_NSTATE = 19;
constexpr double TT_INIT[TT_NSTATE] = {-85.423, 136.89, 8.604,    0.000126, 0.00621,
                                       0.4712,  0.0095, 0.00172,  0.7444,   0.7045,
                                       0.00036, 3.373e-5, 0.7888, 0.9755,   0.9953,
                                       0.999998, 2.42e-8, 3.64,   0.9073};
// SetUpTusscherParameters (SetUpIonicConstants.h:93-145).
constexpr double TTC[53] = {
    8314.472, 310,      96485.3415, 0.185,     0.016404,  10,       1000,
    1,        52,       0.03,       5.4,       140,       2,        5.405,
    0.153,    0.392,    14.838,     0.00029,   0.0000398, 0.000592, 0.294,
    2.724,    1,        40,         1000,      0.1,       2.5,      0.35,
    1.38,     87.5,     0.1238,     0.0005,    0.0146,    0.15,     0.045,
    0.06,     0.005,    1.5,        2.5,       1,         0.102,    0.0038,
    0.00025,  0.00036,  0.006375,   0.2,       0.001,     10,       0.3,
    0.4,      0.00025,  0.001094,   0.00005468};
constexpr double TT_XI = 2000.;  // Tusscher.h:72

// Gate order of UpdateStateVariables (Tusscher.cc:122,125).
constexpr int TT_GATE_STATE[12] = {4, 5, 6, 7, 8, 9, 11, 12, 13, 14, 15, 16};
constexpr int TT_NR_STATE[6] = {1, 2, 3, 10, 17, 18};

template <class T> struct TTG {  // per-node conductances (after scaling)
  T gNa, gKr, gKs, gCaL, gto;
};

template <class T> __device__ inline TTG<T> tt_conductances(const T *scale) {
  TTG<T> g;
  g.gNa = T(TTC[16]) * scale[0];
  g.gKr = T(TTC[14]) * scale[1];
  g.gKs = T(TTC[15]) * scale[2];
  g.gCaL = T(TTC[18]) * scale[3];
  g.gto = T(TTC[20]) * scale[4];
  return g;
}

// Steady states and time constants of the 12 gates (Tusscher.cc:5-55), in the
// order of TT_GATE_STATE.  Depends on V and, for fCass, Ca_ss.
template <class T, bool FAST>
__device__ inline void tt_gates(T V, T Cass, T inf[12], T tau[12]) {
  using M = Mth<T, FAST>;
  inf[7] = T(1.0) / (T(1.0) + M::ex((V + T(20.0)) / T(7.0)));
  tau[7] = T(1102.5) * M::ex(-((V + T(27.0)) * (V + T(27.0))) / T(225.0)) +
           T(200.0) / (T(1.0) + M::ex((T(13.0) - V) / T(10.0))) +
           T(180.0) / (T(1.0) + M::ex((V + T(30.0)) / T(10.0))) + T(20.0);
  inf[8] = T(0.67) / (T(1.0) + M::ex((V + T(35.0)) / T(7.0))) + T(0.33);
  tau[8] = T(562.0) * M::ex(-((V + T(27.0)) * (V + T(27.0))) / T(240.0)) +
           T(31.0) / (T(1.0) + M::ex((T(25.0) - V) / T(10.0))) +
           T(80.0) / (T(1.0) + M::ex((V + T(30.0)) / T(10.0)));
  const T q = (Cass / T(0.05)) * (Cass / T(0.05));
  inf[9] = T(0.6) / (T(1.0) + q) + T(0.4);
  tau[9] = T(80.0) / (T(1.0) + q) + T(2.0);
  inf[10] = T(1.0) / (T(1.0) + M::ex((V + T(20.0)) / T(5.0)));
  tau[10] = T(85.0) * M::ex(-((V + T(45.0)) * (V + T(45.0))) / T(320.0)) +
            T(5.0) / (T(1.0) + M::ex((V - T(20.0)) / T(5.0))) + T(3.0);
  inf[11] = T(1.0) / (T(1.0) + M::ex((T(20.0) - V) / T(6.0)));
  tau[11] = T(9.5) * M::ex(-((V + T(40.0)) * (V + T(40.0))) / T(1800.0)) + T(0.8);
  inf[0] = T(1.0) / (T(1.0) + M::ex((T(-26.0) - V) / T(7.0)));
  tau[0] = (T(450.0) / (T(1.0) + M::ex((T(-45.0) - V) / T(10.0)))) *
           (T(6.0) / (T(1.0) + M::ex((V + T(30.0)) / T(11.5))));
  inf[1] = T(1.0) / (T(1.0) + M::ex((V + T(88.0)) / T(24.0)));
  tau[1] = (T(3.0) / (T(1.0) + M::ex((T(-60.0) - V) / T(20.0)))) *
           (T(1.12) / (T(1.0) + M::ex((V - T(60.0)) / T(20.0))));
  inf[2] = T(1.0) / (T(1.0) + M::ex((T(-5.0) - V) / T(14.0)));
  tau[2] = (T(1400.0) / M::sq(T(1.0) + M::ex((T(5.0) - V) / T(6.0)))) *
               (T(1.0) / (T(1.0) + M::ex((V - T(35.0)) / T(15.0)))) + T(80.0);
  {
    const T e = T(1.0) + M::ex((T(-56.86) - V) / T(9.03));
    inf[3] = T(1.0) / (e * e);
  }
  tau[3] = (T(1.0) / (T(1.0) + M::ex((T(-60.0) - V) / T(5.0)))) *
           (T(0.1) / (T(1.0) + M::ex((V + T(35.0)) / T(5.0))) +
            T(0.1) / (T(1.0) + M::ex((V - T(50.0)) / T(200.0))));
  {
    const T e = T(1.0) + M::ex((V + T(71.55)) / T(7.43));
    inf[4] = T(1.0) / (e * e);
    inf[5] = inf[4];
  }
  if (V < T(-40.0)) {
    const T ah = T(0.057) * M::ex(-(V + T(80.0)) / T(6.8));
    const T bh = T(2.7) * M::ex(T(0.079) * V) + T(310000.) * M::ex(T(0.3485) * V);
    tau[4] = T(1.0) / (ah + bh);
    const T aj = ((T(-25428.0) * M::ex(T(0.2444) * V) - T(6.948e-06) * M::ex(T(-0.04391) * V)) *
                  (V + T(37.78))) /
                 (T(1.0) + M::ex(T(0.311) * (V + T(79.23))));
    const T bj = (T(0.02424) * M::ex(T(-0.01052) * V)) / (T(1.0) + M::ex(T(-0.1378) * (V + T(40.14))));
    tau[5] = T(1.0) / (aj + bj);
  } else {
    const T bh = T(0.77) / (T(0.13) * (T(1.0) + M::ex((V + T(10.66)) / T(-11.1))));
    tau[4] = T(1.0) / (T(0.0) + bh);
    const T bj = (T(0.6) * M::ex(T(0.057) * V)) / (T(1.0) + M::ex(T(-0.1) * (V + T(32.0))));
    tau[5] = T(1.0) / (T(0.0) + bj);
  }
  inf[6] = T(1.0) / (T(1.0) + M::ex((T(-8.0) - V) / T(7.5)));
  tau[6] = (T(1.4) / (T(1.0) + M::ex((T(-35.0) - V) / T(13.0))) + T(0.25)) *
               (T(1.4) / (T(1.0) + M::ex((V + T(5.0)) / T(5.0)))) +
           T(1.0) / (T(1.0) + M::ex((T(50.0) - V) / T(20.0)));
}

// Currents and the rates of V and the six concentrations (Tusscher.cc:56-95).
// R7[0] = dV/dt, R7[1..6] = d/dt of K_i, Na_i, Ca_i, Ca_ss, Ca_SR, R'.
// istim_term is the reference's _Algebraic[12] = -istim/(C_m Xi).
// LIMIT evaluates the I_CaL driving-force ratio through expm1 so it stays
// finite at V = 15 mV; without it the reference's expression is used verbatim.
template <class T, bool FAST, bool LIMIT>
__device__ inline void tt_currents(const T *S, T istim_term, const TTG<T> &g, T R7[7]) {
  using M = Mth<T, FAST>;
  const T R = T(TTC[0]), Tk = T(TTC[1]), F = T(TTC[2]);
  const T Cmc = T(TTC[3]), Vc = T(TTC[4]);
  const T Ko = T(TTC[10]), Nao = T(TTC[11]), Cao = T(TTC[12]);
  const T RTF = (R * Tk) / F;
  const T V = S[0];
  const T FVRT = (V * F) / (R * Tk);
  const T i_NaK = (((T(TTC[21]) * Ko) / (Ko + T(TTC[22]))) * S[2] / (S[2] + T(TTC[23]))) /
                  (T(1.0) + T(0.1245) * M::ex(T(-0.1) * FVRT) + T(0.0353) * M::ex(-FVRT));
  const T E_Na = RTF * M::lg(Nao / S[2]);
  const T i_Na = g.gNa * (S[7] * S[7] * S[7]) * S[8] * S[9] * (V - E_Na);
  const T i_b_Na = T(TTC[17]) * (V - E_Na);
  const T gam = T(TTC[27]);
  const T i_NaCa =
      (T(TTC[24]) * (M::ex(gam * FVRT) * (S[2] * S[2] * S[2]) * Cao -
                     M::ex((gam - T(1.0)) * FVRT) * (Nao * Nao * Nao) * S[3] * T(TTC[26]))) /
      ((T(TTC[29]) * T(TTC[29]) * T(TTC[29]) + Nao * Nao * Nao) * (T(TTC[28]) + Cao) *
       (T(1.0) + T(TTC[25]) * M::ex((gam - T(1.0)) * FVRT)));
  const T E_K = RTF * M::lg(Ko / S[1]);
  const T a_K1 = T(0.1) / (T(1.0) + M::ex(T(0.06) * ((V - E_K) - T(200.0))));
  const T b_K1 = (T(3.0) * M::ex(T(0.0002) * ((V - E_K) + T(100.0))) +
                  M::ex(T(0.1) * ((V - E_K) - T(10.0)))) /
                 (T(1.0) + M::ex(T(-0.5) * (V - E_K)));
  const T xK1 = a_K1 / (a_K1 + b_K1);
  const T sqKo = M::sq(Ko / T(5.4));
  const T i_K1 = T(TTC[13]) * xK1 * sqKo * (V - E_K);
  const T i_to = g.gto * S[16] * S[15] * (V - E_K);
  const T i_Kr = g.gKr * sqKo * S[4] * S[5] * (V - E_K);
  const T E_Ks = RTF * M::lg((Ko + T(TTC[9]) * Nao) / (S[1] + T(TTC[9]) * S[2]));
  const T i_Ks = g.gKs * (S[6] * S[6]) * (V - E_Ks);
  T i_CaL;
  if (LIMIT) {
    const T u = T(2.0) * (V - T(15.0)) * F / (R * Tk);
    const T eu = M::ex(u);
    const T ratio = (fabs(u) < T(1e-4)) ? (T(1.0) - T(0.5) * u) : u / M::em1(u);
    i_CaL = g.gCaL * S[11] * S[12] * S[13] * S[14] * T(2.0) * F * ratio *
            (T(0.25) * S[10] * eu - Cao);
  } else {
    const T eu = M::ex((T(2.0) * (V - T(15.0)) * F) / (R * Tk));
    i_CaL = ((g.gCaL * S[11] * S[12] * S[13] * S[14] * T(4.0) * (V - T(15.0)) * (F * F)) /
             (R * Tk)) *
            (T(0.25) * S[10] * eu - Cao) / (eu - T(1.0));
  }
  const T E_Ca = ((T(0.5) * R * Tk) / F) * M::lg(Cao / S[3]);
  const T i_b_Ca = T(TTC[19]) * (V - E_Ca);
  const T i_p_K = (T(TTC[32]) * (V - E_K)) / (T(1.0) + M::ex((T(25.0) - V) / T(5.98)));
  const T i_p_Ca = (T(TTC[30]) * S[3]) / (S[3] + T(TTC[31]));
  R7[0] = -(i_K1 + i_to + i_Kr + i_Ks + i_CaL + i_NaK + i_Na + i_b_Na + i_NaCa + i_b_Ca +
            i_p_K + i_p_Ca + istim_term);
  R7[1] = ((T(-1.0) * ((i_K1 + i_to + i_Kr + i_Ks + i_p_K + istim_term) - T(2.0) * i_NaK)) /
           (T(1.0) * Vc * F)) *
          Cmc;
  R7[2] = ((T(-1.0) * (i_Na + i_b_Na + T(3.0) * i_NaK + T(3.0) * i_NaCa)) / (T(1.0) * Vc * F)) * Cmc;
  const T Vsr = T(TTC[51]), Vss = T(TTC[52]);
  const T i_up = T(TTC[44]) / (T(1.0) + (T(TTC[42]) * T(TTC[42])) / (S[3] * S[3]));
  const T i_leak = T(TTC[43]) * (S[17] - S[3]);
  const T i_xfer = T(TTC[41]) * (S[10] - S[3]);
  const T bi = T(1.0) / (T(1.0) + (T(TTC[45]) * T(TTC[46])) / ((S[3] + T(TTC[46])) * (S[3] + T(TTC[46]))));
  R7[3] = bi * ((((i_leak - i_up) * Vsr) / Vc + i_xfer) -
                (T(1.0) * ((i_b_Ca + i_p_Ca) - T(2.0) * i_NaCa) * Cmc) / (T(2.0) * T(1.0) * Vc * F));
  const T kcasr = T(TTC[38]) - (T(TTC[38]) - T(TTC[39])) /
                                  (T(1.0) + (T(TTC[37]) / S[17]) * (T(TTC[37]) / S[17]));
  const T k2 = T(TTC[34]) * kcasr;
  R7[6] = -k2 * S[10] * S[18] + T(TTC[36]) * (T(1.0) - S[18]);
  const T k1 = T(TTC[33]) / kcasr;
  const T O = (k1 * (S[10] * S[10]) * S[18]) / (T(TTC[35]) + k1 * (S[10] * S[10]));
  const T i_rel = T(TTC[40]) * O * (S[17] - S[10]);
  const T bsr = T(1.0) / (T(1.0) + (T(TTC[47]) * T(TTC[48])) / ((S[17] + T(TTC[48])) * (S[17] + T(TTC[48]))));
  R7[5] = bsr * (i_up - (i_rel + i_leak));
  const T bss = T(1.0) / (T(1.0) + (T(TTC[49]) * T(TTC[50])) / ((S[10] + T(TTC[50])) * (S[10] + T(TTC[50]))));
  R7[4] = bss * (((T(-1.0) * i_CaL * Cmc) / (T(2.0) * T(1.0) * Vss * F) + (i_rel * Vsr) / Vss) -
                 (i_xfer * Vc) / Vss);
}

// Rush-Larsen gates + forward-Euler concentrations with one rate evaluation:
// gates first (at V and the old Ca_ss), then currents at the new gates and old
// concentrations, mirroring the reference's gates-then-concentrations order.
// Returns dV/dt; the caller advances V.
// The fast FP32 form computes in RcpF (reciprocal divisions).
template <class T, bool FAST>
__device__ VG_INLINE T tt_step_rl(T S[TT_NSTATE], T V, T istim_term, T dt, const TTG<T> &g) {
  using U = typename RateT<T, FAST>::type;
  using M = Mth<U, FAST>;
  U Su[TT_NSTATE];
  Su[0] = U(V);
#pragma unroll
  for (int k = 1; k < TT_NSTATE; k++) Su[k] = U(S[k]);
  const TTG<U> gu{U(g.gNa), U(g.gKr), U(g.gKs), U(g.gCaL), U(g.gto)};
  const U dtu = U(dt);
  U inf[12], tau[12];
  tt_gates<U, FAST>(Su[0], Su[10], inf, tau);
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const int k = TT_GATE_STATE[i];
    Su[k] = inf[i] - (inf[i] - Su[k]) * M::ex(-dtu / tau[i]);
  }
  U R7[7];
  tt_currents<U, FAST, true>(Su, U(istim_term), gu, R7);
#pragma unroll
  for (int i = 0; i < 6; i++) Su[TT_NR_STATE[i]] += dtu * R7[1 + i];
  S[0] = V;
#pragma unroll
  for (int k = 1; k < TT_NSTATE; k++) S[k] = T(Su[k]);
  return T(R7[0]);
}

// 6x6 solve with partial pivoting of J d = B for the row-major array J
// (J[j*6+i] = dg_j/dx_i), or, with transpose, of the system a column-major
// LAPACK dgesv sees for that array (entry (r,c) = J[c*6+r]); B overwritten.
// Gaussian elimination with partial pivoting on a 6x6 system, fully unrolled:
// the pivot row is swapped in by selects over the candidate rows, so the
// matrix stays in registers (an index chosen at run time would put it in
// scratch memory).  Same operations and pivot order as the plain loop.
template <class T> __device__ VG_INLINE void tt_solve6(const T J[36], T B[6], bool transpose) {
  T a[6][6];
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = 0; c < 6; c++) a[r][c] = transpose ? J[c * 6 + r] : J[r * 6 + c];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    int p = k;
    T best = fabs(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < 6; r++)
      if (fabs(a[r][k]) > best) {
        best = fabs(a[r][k]);
        p = r;
      }
#pragma unroll
    for (int r = k + 1; r < 6; r++) {
      const bool sw = r == p;
#pragma unroll
      for (int c = 0; c < 6; c++) {
        const T t = a[k][c];
        a[k][c] = sw ? a[r][c] : t;
        a[r][c] = sw ? t : a[r][c];
      }
      const T t = B[k];
      B[k] = sw ? B[r] : t;
      B[r] = sw ? t : B[r];
    }
#pragma unroll
    for (int r = k + 1; r < 6; r++) {
      const T l = a[r][k] / a[k][k];
#pragma unroll
      for (int c = k + 1; c < 6; c++) a[r][c] -= l * a[k][c];
      B[r] -= l * B[k];
    }
  }
#pragma unroll
  for (int k = 5; k >= 0; k--) {
    T s = B[k];
#pragma unroll
    for (int c = k + 1; c < 6; c++) s -= a[k][c] * B[c];
    B[k] = s / a[k][k];
  }
}

// The reference integrator (Tusscher.cc:116-195): backward-Euler gates, then
// Newton on the six concentrations with a central-difference Jacobian,
// stopping at ||dx||2 <= 1e-6 (capped at max_newton passes).  jt solves each
// step with J^T as the reference's row-major/column-major mix-up does
// (:172-181); by default the step solves J, which converges to the same
// backward-Euler solution and stays physical through the upstroke.
// In FP32 the absolute 1e-6 is below the states' rounding (K_i ~ 138 mM
// has an ulp of 1.5e-5), so once Newton reaches that floor its steps only
// wander and the loop would run to the cap: in FP32 it also ends at the
// first pass whose step is no smaller than the one before.
template <class T>
__device__ VG_INLINE T tt_step_reference(T S[TT_NSTATE], T V, T istim_term, T dt,
                                      const TTG<T> &g, int max_newton, bool jt, int *iters) {
  S[0] = V;
  T StateN[6], oldR[6], fxph[6], fxmh[6], J[36], B[6], R7[7];
#pragma unroll
  for (int i = 0; i < 6; i++) StateN[i] = S[TT_NR_STATE[i]];
  {
    T inf[12], tau[12];
    tt_gates<T, false>(V, S[10], inf, tau);
#pragma unroll
    for (int i = 0; i < 12; i++) {
      const int k = TT_GATE_STATE[i];
      const T b = T(-1.) / tau[i];
      const T a = inf[i] / tau[i];
      S[k] = (S[k] + dt * a) / (T(1) - dt * b);
    }
  }
  tt_currents<T, false, false>(S, istim_term, g, R7);
#pragma unroll
  for (int i = 0; i < 6; i++) oldR[i] = R7[1 + i];
  T resid = T(100.), prev = T(0);
  int it = 0;
  while (resid > T(1e-6) && it < max_newton) {
    // Column i of the Jacobian.  The loop stays rolled (two current
    // evaluations per pass); i is the same in every lane, and the state and
    // column it picks are reached through unrolled tests of i, keeping S and
    // J in registers.
    for (int i = 0; i < 6; i++) {
      T xi = T(0);
#pragma unroll
      for (int q = 0; q < 6; q++)
        if (q == i) xi = S[TT_NR_STATE[q]];
      // 1e-3 x (Tusscher.cc:158); at x = 0 (0/0 there) a step of 1e-12.
      const T eps = xi != T(0) ? T(1e-3) * xi : T(1e-12);
      const T twoeps = T(2.) * eps;
#pragma unroll
      for (int q = 0; q < 6; q++)
        if (q == i) S[TT_NR_STATE[q]] += eps;
      tt_currents<T, false, false>(S, istim_term, g, R7);
#pragma unroll
      for (int j = 0; j < 6; j++) fxph[j] = S[TT_NR_STATE[j]] - StateN[j] - dt * R7[1 + j];
#pragma unroll
      for (int q = 0; q < 6; q++)
        if (q == i) S[TT_NR_STATE[q]] -= T(2) * eps;
      tt_currents<T, false, false>(S, istim_term, g, R7);
#pragma unroll
      for (int j = 0; j < 6; j++) fxmh[j] = S[TT_NR_STATE[j]] - StateN[j] - dt * R7[1 + j];
#pragma unroll
      for (int q = 0; q < 6; q++)
        if (q == i) S[TT_NR_STATE[q]] += eps;
#pragma unroll
      for (int j = 0; j < 6; j++) {
        const T d = (fxph[j] - fxmh[j]) / twoeps;
#pragma unroll
        for (int q = 0; q < 6; q++)
          if (q == i) J[j * 6 + q] = d;
      }
    }
#pragma unroll
    for (int i = 0; i < 6; i++) B[i] = S[TT_NR_STATE[i]] - StateN[i] - dt * oldR[i];
    tt_solve6<T>(J, B, jt);
#pragma unroll
    for (int i = 0; i < 6; i++) S[TT_NR_STATE[i]] -= B[i];
    tt_currents<T, false, false>(S, istim_term, g, R7);
#pragma unroll
    for (int i = 0; i < 6; i++) oldR[i] = R7[1 + i];
    resid = T(0);
#pragma unroll
    for (int i = 0; i < 6; i++) resid += B[i] * B[i];
    resid = sqrt(resid);
    it++;
    if constexpr (sizeof(T) == 4) {
      if (it > 1 && !(resid < prev)) break;
      prev = resid;
    }
  }
  *iters = it;
  return R7[0];
}

// ======================= Corrias rabbit Purkinje ==============================
// State order of Purkinje.cc (22 states, V = 0); constants of
// SetUpPurkinjeParameters (SetUpIonicConstants.h:16-85), the last six derived.
constexpr int PK_NSTATE = 22;
constexpr double PK_INIT[PK_NSTATE] = {-88.34,   0.00001, 0.000032, 0.17,     6.7,      140,
                                       0.001337, 0.01,    0.000003, 0.1,      0.7,      0.0,
                                       0.7,      0.000007, 0.978861, 0.000012, 0.864489, 0.25,
                                       1.0,      0.0,     0.011099, 0.011099};
constexpr double PKL[57] = {
    310.0, 96485.3415, 8314.472, 2.0,    140,     5.4,    69.0,     13266.5, 0.6,
    0.2,   0.06,       700,      9000000, 500,    0.5,    -4320.0,  1.0,     0.9,
    0.7,   5.4,        5.0,      350.0,  10.0,    3.0,    0.005,    2.0,     1140.0,
    15.0,  2.0,        20.0,     1.5,    3.0,     3,      0.001,    0.001,   0.5,
    442.2, 0.01,       0.01,     0.0001, 5.0,     0.0001, 1.5,      0.188709677,
    0.045290323, 2500.0, 0.001, 120.0,  0.000246, 1.7,    1.6,      120.0,   0.000246,
    1.7,   1.6,        10.0,     5000.0};
constexpr double PK_RTF = (PKL[2] * PKL[0]) / PKL[1];            // Constants[58]
constexpr double PK_VOL_I = PKL[9] * PKL[7];                      // Constants[59]
constexpr double PK_VOL_SR = PKL[8] * PKL[7];                     // Constants[60]
constexpr double PK_VOL_MYO = (PKL[8] + PKL[9]) * PKL[7];         // Constants[61]
constexpr double PK_VOL_SUB = PKL[10] * PKL[7];                   // Constants[62]
constexpr double PK_XI = 2500.;        // Purkinje.h:53
constexpr double PK_AREA = 4.612E-5;   // Purkinje.h:55, cm^2

// All 22 rates of Purkinje::ComputeRates (Purkinje.cc:5-104).  stim is the
// stimulus current in pA (the reference's _Algebraic[0]).
template <class T, bool FAST>
__device__ inline void pk_rates(const T S[PK_NSTATE], T stim, T R[PK_NSTATE]) {
  using M = Mth<T, FAST>;
  const T V = S[0];
  auto sig = [](T x) { return T(1.0) / (T(1.0) + M::ex(x)); };
  // Gates with constant time constants.
  R[6] = (sig((V + T(47.8)) / T(-5.5)) - S[6]) / T(PKL[16]);
  R[8] = (sig((V + T(14.6)) / T(-5.5)) - S[8]) / T(PKL[18]);
  R[11] = (sig((V - T(7.0)) / T(-9.0)) - S[11]) / T(PKL[20]);
  R[12] = (sig((V + T(27.5)) / T(8.0)) - S[12]) / T(PKL[21]);
  R[13] = (sig((V + T(25.0)) / T(-5.0)) - S[13]) / T(PKL[24]);
  R[14] = (sig((V + T(69.0)) / T(3.96)) - S[14]) / T(PKL[25]);
  R[15] = (sig((V + T(30.0)) / T(-5.0)) - S[15]) / T(PKL[27]);
  // Gates with voltage- or calcium-dependent time constants.
  R[7] = (sig((V + T(67.9)) / T(3.87)) - S[7]) / (T(1.42271) * M::ex(T(-0.05119) * V));
  {
    const T q = T(0.028) * (V + T(14.5));
    R[9] = (sig((V + T(31.0)) / T(5.54)) - S[9]) / (T(25.1) / (T(0.04) + T(0.7) * M::ex(-(q * q))));
  }
  {
    const T c = S[1] / T(0.0001);
    const T den = T(1.0) + c * c;
    R[10] = ((T(0.4) + T(0.6) / den) - S[10]) / (T(2.0) + T(80.0) / den);
  }
  R[16] = ((T(0.1) + T(0.9) / (T(1.0) + M::ex((V + T(75.6)) / T(6.3)))) - S[16]) /
          (T(120.0) + M::ex((V + T(100.0)) / T(25.0)));
  const T xs_inf = sig((V - T(1.5)) / T(-16.7));
  T tau_xs;
  {
    const T u = V + T(30.0);
    tau_xs = (fabs(u) < T(0.0145))
                 ? T(417.946)
                 : T(1.0) / ((T(7.19e-05) * u) / (T(1.0) - M::ex(T(-0.148) * u)) +
                             (T(0.000131) * u) / (M::ex(T(0.0687) * u) - T(1.0)));
  }
  R[18] = (xs_inf - S[18]) / tau_xs;
  R[19] = (xs_inf - S[19]) / (T(4.0) * tau_xs);
  {
    const T inf = sig((V + T(109.0)) / T(10.0));
    const T tau = T(6000.0) / (M::ex(-(T(2.9) + T(0.04) * V)) + M::ex(T(3.6) + T(0.11) * V));
    R[20] = (inf - S[20]) / tau;
    R[21] = (inf - S[21]) / tau;
  }
  {
    const T u7 = V + T(7.0), u10 = V + T(10.0);
    const T a = (fabs(u7) > T(0.001)) ? (T(0.00138) * u7) / (T(1.0) - M::ex(T(-0.123) * u7))
                                     : T(0.00138) / T(0.123);
    const T b = (fabs(u10) > T(0.001)) ? (T(6.1e-05) * u10) / (M::ex(T(0.145) * u10) - T(1.0))
                                      : T(0.00061) / T(0.145);
    R[17] = (sig((V + T(50.0)) / T(-7.5)) - S[17]) * (a + b);
  }
  // Reversal potentials (mV).
  const T rtf = T(PK_RTF);
  const T Ko = T(PKL[5]), Nao = T(PKL[4]), Cao = T(PKL[3]);
  const T EK = rtf * M::lg(Ko / S[5]);
  const T ENa = rtf * M::lg(Nao / S[4]);
  const T ECa = T(0.5) * rtf * M::lg(Cao / S[1]);
  // Currents (pA).
  const T iK1 = T(PKL[29]) * M::pw(Ko / T(5.4), T(0.8)) * sig((T(92.0) + V) / T(10.0)) * (V - EK);
  const T ito = T(PKL[22]) * S[11] * S[12] * (V - EK);
  const T isus = T(PKL[23]) * sig((T(5.0) - V) / T(17.0)) * (V - EK);
  const T iKr = T(PKL[30]) * (Ko / T(5.4)) * sig((T(33.0) + V) / T(22.4)) * S[17] * (V - EK);
  const T iKs = T(PKL[31]) * S[18] * S[19] * (V - EK);
  const T ibK = T(PKL[37]) * (V - EK);
  const T iNaK = T(PKL[36]) * sig((V + T(80.0)) / T(-45.0)) * sig(V / T(125.0)) *
                 (T(1.0) / (T(1.0) + M::pw(T(1.9) / Ko, T(1.45)))) *
                 (T(1.0) / (T(1.0) + T(31.98) / S[4]));
  const T ifK = T(PKL[43]) * S[20] * (V - EK);
  const T iCaL = T(PKL[19]) * S[8] * S[9] * S[10] * (V - ECa);
  const T iCaT = T(PKL[17]) * S[6] * S[7] * (V - ECa);
  const T iNa = T(PKL[26]) * S[13] * S[14] * (V - ENa);
  const T iNaL = T(PKL[28]) * S[15] * S[16] * (V - ENa);
  const T ipCa = T(PKL[40]) * (T(1.0) / (T(1.0) + M::pw(T(PKL[41]) / S[1], T(PKL[42]))));
  const T ibCa = T(PKL[39]) * (V - ECa);
  const T ifNa = T(PKL[44]) * S[21] * (V - ENa);
  const T nH = T(PKL[32]), gam = T(PKL[35]);
  const T Nao_n = M::pw(Nao, nH), Nai_n = M::pw(S[4], nH);
  const T iNaCa = (T(512.009) * T(PKL[33]) *
                   (M::ex((gam * (nH - T(2.0)) * V) / rtf) * Nai_n * Cao -
                    S[1] * M::ex(((gam - T(1.0)) * (nH - T(2.0)) * V) / rtf) * Nao_n)) /
                  ((T(1.0) + T(PKL[34]) * (S[1] * Nao_n + Cao * Nai_n)) * (T(1.0) + S[1] / T(0.0069)));
  const T ibNa = T(PKL[38]) * (V - ENa);
  R[0] = (T(-1.0) / T(PKL[6])) * (iK1 + ito + isus + iKr + iKs + ibK + iNaK + iCaL + iCaT + iNa +
                                  iNaL + ipCa + ibCa + ifNa + ifK + iNaCa + ibNa + stim);
  const T F = T(PKL[1]);
  R[5] = ((-ito - isus - iKr - iKs - iK1 - ibK - ifK - stim + T(2.0) * iNaK) * T(1000.0)) /
         (F * T(PK_VOL_MYO));
  R[4] = ((-iNa - iNaL - T(3.0) * iNaK - T(3.0) * iNaCa - ifNa - ibNa) * T(1000.0)) / (F * T(PK_VOL_MYO));
  const T jrel = T(PKL[45]) * (S[3] / (T(1.0) + (T(PKL[46]) / S[1]) * (T(PKL[46]) / S[1])));
  const T nu = T(PKL[50]), nsr = T(PKL[54]);
  const T ui = M::pw(S[1] / T(PKL[48]), nu), us = M::pw(S[3] / T(PKL[49]), nu);
  const T jup = T(PKL[47]) * ((ui - us) / (T(1.0) + ui + us));
  const T vi = M::pw(S[2] / T(PKL[52]), nsr), vs = M::pw(S[3] / T(PKL[53]), nsr);
  const T jup2 = T(PKL[51]) * ((vi - vs) / (T(1.0) + vi + vs));
  const T jtr = T(PKL[55]) * (S[3] - S[2]);
  R[3] = ((-jrel - jtr + jup2 + jup) * T(1000.0)) / (T(2.0) * F * T(PK_VOL_SUB));
  const T jdiff = T(PKL[56]) * (S[1] - S[2]);
  R[1] = ((-iCaL - iCaT - ibCa + jrel - jup - jdiff + T(2.0) * iNaCa - ipCa) * T(1000.0)) /
         (T(2.0) * F * T(PK_VOL_I));
  R[2] = ((-jup2 + jdiff + jtr) * T(1000.0)) / (T(2.0) * F * T(PK_VOL_SR));
}

// Purkinje::Compute_Ion + UpdateStateVariables (Purkinje.cc:179-210):
// forward Euler on all 22 states, V included, in 10 sub-steps of dt/10 when
// |dV/dt| > 1.5 mV/ms; returns the last dV/dt evaluated.
// The fast FP32 form computes the rates in RcpF (reciprocal divisions).
template <class T, bool FAST>
__device__ VG_INLINE T pk_step(T S[PK_NSTATE], T V, T istim, T dt) {
  using U = typename RateT<T, FAST>::type;
  S[0] = V;
  U Su[PK_NSTATE];
#pragma unroll
  for (int j = 0; j < PK_NSTATE; j++) Su[j] = U(S[j]);
  const U stim = U(T(-1E6) * istim * T(PK_AREA) / T(PK_XI));
  U R[PK_NSTATE];
  pk_rates<U, FAST>(Su, stim, R);
  if (fabs(T(R[0])) > T(1.5)) {
    const U h = U(dt / T(10));
    for (int k = 0; k < 10; k++) {
      pk_rates<U, FAST>(Su, stim, R);
#pragma unroll
      for (int j = 0; j < PK_NSTATE; j++) Su[j] += R[j] * h;
    }
  } else {
#pragma unroll
    for (int j = 0; j < PK_NSTATE; j++) Su[j] += R[j] * U(dt);
  }
#pragma unroll
  for (int j = 0; j < PK_NSTATE; j++) S[j] = T(Su[j]);
  return T(R[0]);
}

// ======================= Mahajan 2008 (UCLA rabbit ventricle) ================
// IonicModel/Mahajan.cc, constants SetUpIonicConstants.h:159-251.  State
// order of getInternalParameters: 0 V, 1 m, 2 h, 3 j, 4 xr, 5 xs1, 6 xs2,
// 7 c_p, 8 c_s, 9 c_i, 10 c_j, 11 c_j', 12 I_r, 13 c1, 14 c2, 15 xi1ca,
// 16 xi1ba, 17 xi2ca, 18 xi2ba, 19 Na_i, 20 x_tof, 21 y_tof, 22 x_tos,
// 23 y_tos, 24 troponin-Ca (myoplasm), 25 troponin-Ca (submembrane).
// The Ca-cycling and Na_i states (MJ_ACC) are accumulated in FP64 in every
// precision; the gates and channel states (MJ_GATE) in T.
constexpr int MJ_NSTATE = 26;
constexpr int MJ_NGATE = 16;
constexpr int MJ_NACC = 9;
constexpr int MJ_GATE[MJ_NGATE] = {1, 2, 3, 4, 5, 6, 13, 14, 15, 16, 17, 18, 20, 21, 22, 23};
constexpr int MJ_ACC[MJ_NACC] = {7, 8, 9, 10, 11, 12, 19, 24, 25};
constexpr double MJ_XI = 1.55e-4 / 2.58e-8;  // Mahajan.cc:164
// G_tof and G_Ks by tissue type (epi/myo/endo x apex/center/base).
constexpr double MJ_GTOF[9] = {0.11, 0.11, 0.11, 0.11, 0.11, 0.11, 0.11 * 0.85, 0.11 * 0.85,
                               0.11 * 0.85};
constexpr double MJ_GKS[9] = {1.9 * 0.32,  1.4 * 0.32,  1.0 * 0.32, 0.74 * 0.32, 0.52 * 0.32,
                              0.35 * 0.32, 0.98 * 0.32, 0.7 * 0.32, 0.5 * 0.32};
// MAHAJAN_FAIL's per-type G_tof, G_Ks (SetUpMahajan_failParameters,
// SetUpIonicConstants.h:254-315: the case value times 0.67 and 0.6); row 9 is
// the set of types beyond 8 (its default branch), which differs from type 2.
constexpr double MJF_GTOF[10] = {0.12 * 0.67, 0.12 * 0.67, 0.12 * 0.67, 0.12 * 0.67,
                                 0.12 * 0.67, 0.12 * 0.67, 0.12 * 0.77 * 0.67, 0.12 * 0.77 * 0.67,
                                 0.12 * 0.77 * 0.67, 0.12 * 0.67};
constexpr double MJF_GKS[10] = {1.5 * 0.1386 * 0.6, 1.3 * 0.1386 * 0.6, 1.1 * 0.1386 * 0.6,
                                1.0 * 0.1386 * 0.6, 0.9 * 0.1386 * 0.6, 0.8 * 0.1386 * 0.6,
                                1.1 * 0.1386 * 0.6, 1.0 * 0.1386 * 0.6, 0.9 * 0.1386 * 0.6,
                                0.1386 * 0.6};
// The other constants that differ between the two sets.
template <bool FAIL>
struct MjK {
  static constexpr double gtos = 0.04, gk1 = 0.3, gnaca = 0.84, vup = 0.4, taus = 4.0,
                          taua = 100.0, av = 11.3, cstar = 90.0, sh = 50.0, leak = 0.0;
};
template <>
struct MjK<true> {
  static constexpr double gtos = 0.04 * 0.65, gk1 = 0.3 * 0.5, gnaca = 0.84 * 2.0, vup = 0.27,
                          taus = 6.0, taua = 200.0, av = 22.5, cstar = 50.0, sh = 10.0,
                          leak = 0.01;
};
// Initial states by tissue type (Mahajan.cc:27-158).
constexpr double MJ_INIT[9][MJ_NSTATE] = {
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

// Exponential relaxation over h toward x_inf (time constant tau) as a rate.
template <class T, bool FAST>
__device__ inline T mj_relax(T xinf, T tau, T x, T h) {
  return (xinf - (xinf - x) * Mth<T, FAST>::ex(-h / tau) - x) / h;
}

// 1 / (1 + sum of the instantaneous buffers' derivatives) at concentration c.
template <class T>
__device__ inline T mj_buffer(T c) {
  const T a = T(7.0) + c, b = T(0.6) + c, m = T(0.3) + c, s = T(13.0) + c;
  return T(1) / (T(1) + T(24.0 * 7.0) / (a * a) + T(47.0 * 0.6) / (b * b) +
                 T(15.0 * 0.3) / (m * m) + T(42.0 * 13.0) / (s * s));
}

// Mahajan::ucla_rhsfun with the comp_* currents (Mahajan.cc:168-603): every
// rate at gate step h, stimulus istim in uA/uF, tissue-type G_tof and G_Ks.
template <class T, bool FAST, bool FAIL = false>
__device__ VG_INLINE void mj_rates(const T S[MJ_NSTATE], T h, T istim, T gtof, T gks,
                                T R[MJ_NSTATE]) {
  using M = Mth<T, FAST>;
  using K = MjK<FAIL>;
  const T V = S[0];
  const T frt = T(96.485 / (8.314 * 308.0));
  const T nao = T(136.0), ki = T(140.0), ko = T(5.4), cao = T(1.8);
  const T nai = S[19], cs = S[8], ci = S[9], cj = S[10];
  const T sq_ko = T(1.0);  // sqrt(K_o / 5.4) with K_o = 5.4

  // I_K1 with a P_K : P_Na = 100 : 1 reversal
  const T ek1 = M::lg((T(100.0) * ko + nao) / (T(100.0) * ki + nai)) / frt;
  const T a_k1 = T(1.02) / (T(1) + M::ex(T(0.2385) * (V - ek1 - T(59.215))));
  const T b_k1 = (T(0.49124) * M::ex(T(0.08032) * (V - ek1 + T(5.476))) +
                  M::ex(T(0.061750) * (V - ek1 - T(594.31)))) /
                 (T(1) + M::ex(T(-0.5143) * (V - ek1 + T(4.753))));
  const T i_k1 = T(K::gk1) * sq_ko * a_k1 / (a_k1 + b_k1) * (V - ek1);

  // I_to slow and fast
  const T ek = M::lg(ko / ki) / frt;
  const T e3 = M::ex((V + T(3.0)) / T(15.0));
  const T x_inf = T(1) / (T(1) + T(1) / e3);
  const T y_inf = T(1) / (T(1) + M::ex((V + T(33.5)) / T(10.0)));
  const T tau_xs = T(9.0) / (T(1) + e3) + T(0.5);
  const T tau_ys = T(3000.0) / (T(1) + M::ex((V + T(60.0)) / T(10.0))) + T(30.0);
  const T i_tos = T(K::gtos) * S[22] * (S[23] + T(0.5) * y_inf) * (V - ek);
  R[22] = mj_relax<T, FAST>(x_inf, tau_xs, S[22], h);
  R[23] = mj_relax<T, FAST>(y_inf, tau_ys, S[23], h);
  const T v30 = V / T(30.0);
  const T tau_xf = T(3.5) * M::ex(-v30 * v30) + T(1.5);
  const T tau_yf = T(20.0) * y_inf + T(20.0);
  const T i_tof = gtof * S[20] * S[21] * (V - ek);
  R[20] = mj_relax<T, FAST>(x_inf, tau_xf, S[20], h);
  R[21] = mj_relax<T, FAST>(y_inf, tau_yf, S[21], h);

  // I_NaK
  const T sigma = T((exp(136.0 / 67.3) - 1.0) / 7.0);  // Na_o = 136 mM
  const T evf = M::ex(-V * frt);
  const T f_nak = T(1) / (T(1) + T(0.1245) * M::ex(T(-0.1) * V * frt) + T(0.0365) * sigma * evf);
  const T i_nak = T(1.5) * f_nak / (T(1) + T(12.0) / nai) * ko / (ko + T(1.5));

  // Na/Ca exchange flux (concentrations in mM)
  const T csm = cs / T(1000.0);
  const T nai3 = nai * nai * nai, nao3 = nao * nao * nao;
  const T e_in = M::ex(T(0.35) * V * frt), e_out = M::ex(T(0.35 - 1.0) * V * frt);
  const T zw3 = nai3 * cao * e_in - nao3 * csm * e_out;
  const T zw4 = T(1) + T(0.2) * e_out;
  const T rl = T(0.3) / cs;
  const T aloss = T(1) / (T(1) + rl * rl * rl);
  const T rn = nai / T(12.3);
  const T zw8 = T(1.3) * nai3 + T(87.5 * 87.5 * 87.5) * csm +
                T(12.3 * 12.3 * 12.3) * cao * (T(1) + csm / T(0.0036)) +
                T(0.0036) * nao3 * (T(1) + rn * rn * rn) + nai3 * cao + nao3 * csm;
  const T j_naca = T(K::gnaca) * aloss * zw3 / (zw4 * zw8);

  // Ca cycling
  const T j_d = (cs - ci) / T(K::taus);
  R[24] = T(0.0327) * ci * (T(70.0) - S[24]) - T(0.0196) * S[24];
  R[25] = T(0.0327) * cs * (T(70.0) - S[25]) - T(0.0196) * S[25];
  const T j_up = T(K::vup) * ci * ci / (ci * ci + T(0.25));
  const T j_leak = T(2.069e-5) * cj * cj / (cj * cj + T(2500.0)) * (T(16.667) * cj - ci);

  // L-type Ca channel Markov chain
  const T po_inf = T(1) / (T(1) + M::ex(-V / T(8.0)));
  const T alpha = po_inf, beta = T(1) - po_inf;
  const T rc = T(3.0) / S[7];
  const T fca = T(1) / (T(1) + rc * rc * rc);
  const T s1 = T(0.0182688) * fca, s1t = T(0.00195);
  const T k1 = T(0.024168) * fca, k2 = T(1.03615e-4), k1t = T(0.00413), k2t = T(0.00224);
  const T r1 = T(0.30), r2 = T(3.0);
  const T s2 = s1 * (r1 / r2) * (k2 / k1);
  const T s2t = T(0.00195 * (0.30 / 3.0) * (0.00224 / 0.00413));
  const T poi = T(1) / (T(1) + M::ex(-(V + T(40.0)) / T(3.0)));
  const T k3 = (T(1) - poi) / T(3.0), k3t = k3;
  const T Pr = T(1) - T(1) / (T(1) + M::ex(-(V + T(40.0)) / T(4.0)));
  const T recov = T(10.0) + T(4954.0) * M::ex(V / T(15.6));
  const T rcp = S[7] / T(6.09365);
  const T rcp2 = rcp * rcp;
  const T tau_ca0 = T(78.0329) / (T(1) + rcp2 * rcp2) + T(0.1);
  const T tau_ca = (recov - tau_ca0) * Pr + tau_ca0;
  const T tau_ba = (recov - T(450.0)) * Pr + T(450.0);
  const T Ps = T(1) / (T(1) + M::ex(-(V + T(40.0)) / T(11.32)));
  const T k6 = fca * Ps / tau_ca, k5 = (T(1) - Ps) / tau_ca;
  const T k6t = Ps / tau_ba, k5t = (T(1) - Ps) / tau_ba;
  const T k4 = k3 * (alpha / beta) * (k1 / k2) * (k5 / k6);
  const T k4t = k3t * (alpha / beta) * (k1t / k2t) * (k5t / k6t);
  const T c1 = S[13], c2 = S[14], i1c = S[15], i1b = S[16], i2c = S[17], i2b = S[18];
  const T po = T(1) - i1c - i2c - i1b - i2b - c1 - c2;
  R[14] = beta * c1 + k5 * i2c + k5t * i2b - (k6 + k6t + alpha) * c2;
  R[13] = alpha * c2 + k2 * i1c + k2t * i1b + r2 * po - (beta + r1 + k1t + k1) * c1;
  R[15] = k1 * c1 + k4 * i2c + s1 * po - (k3 + k2 + s2) * i1c;
  R[17] = k3 * i1c + k6 * c2 - (k5 + k4) * i2c;
  R[16] = k1t * c1 + k4t * i2b + s1t * po - (k3t + k2t + s2t) * i1b;
  R[18] = k3t * i1b + k6t * c2 - (k5t + k4t) * i2b;

  // Ca driving force
  const T za = T(2.0) * V * frt;
  const T f1 = T(4.0 * 0.00054 * 96.485 * 96.485 / (8.314 * 308.0));
  const T eza = M::ex(za);
  T rxa;
  if (fabs(za) < T(0.001)) rxa = f1 * (csm * eza - T(0.341) * cao) / (T(2.0) * frt);
  else rxa = V * f1 * (csm * eza - T(0.341) * cao) / (eza - T(1));
  const T j_ca = T(182.0) * po * rxa;

  R[8] = mj_buffer(cs) * (T(50.0) * (S[12] - j_d - j_ca + j_naca) - R[25]);
  R[9] = mj_buffer(ci) * (j_d - j_up + j_leak - R[24]);
  R[10] = -S[12] + j_up - j_leak;
  R[11] = (cj - S[11]) / T(K::taua);

  // SR release: load dependence, sparks, dyad
  T q = T(0);
  const T cjp = S[11];
  if (cjp > T(K::sh) && cjp < T(K::cstar)) q = cjp - T(K::sh);
  else if (cjp >= T(K::cstar)) q = T(K::av) * cjp + T((K::cstar - K::sh) - K::av * K::cstar);
  if (q >= T(153.0)) q = T(153.0);
  const T Qr = cj * q / T(K::cstar);
  const T arx = fabs(rxa);
  const T ey = M::ex(T(-0.05) * (V + T(30.0)));
  const T spark = T(2.58079) * po * arx * ey / (T(1) + ey);
  R[12] = spark * Qr - S[12] * (T(1) - T(30.0) * R[10] / cj) / T(30.0);
  const T ex = M::ex(T(-0.3576) * (V + T(30.0)));
  const T g_sr = T(26841.8) * ex / (T(1) + ex);
  R[7] = po * Qr * arx * g_sr + po * T(9000.0) * arx - (S[7] - cs) / T(0.5);

  // I_Na
  const T ena = M::lg(nao / nai) / frt;
  T am = T(3.2);
  if (fabs(V + T(47.13)) > T(0.001))
    am = T(0.32) * (V + T(47.13)) / (T(1) - M::ex(T(-0.1) * (V + T(47.13))));
  const T bm = T(0.08) * M::ex(-V / T(11.0));
  const T a = T(1) - T(1) / (T(1) + M::ex(-(V + T(40.0)) / T(0.24)));
  const T ah = a * T(0.135) * M::ex((T(80.0) + V) / T(-6.8));
  const T bh = (T(1) - a) / (T(0.13) * (T(1) + M::ex((V + T(10.66)) / T(-11.1)))) +
               a * (T(3.56) * M::ex(T(0.079) * V) + T(3.1e5) * M::ex(T(0.35) * V));
  const T aj = a * (T(-1.2714e5) * M::ex(T(0.2444) * V) - T(3.474e-5) * M::ex(T(-0.04391) * V)) *
               (V + T(37.78)) / (T(1) + M::ex(T(0.311) * (V + T(79.23))));
  const T bj = (T(1) - a) * (T(0.3) * M::ex(T(-2.535e-7) * V) / (T(1) + M::ex(T(-0.1) * (V + T(32.0))))) +
               a * (T(0.1212) * M::ex(T(-0.01052) * V) / (T(1) + M::ex(T(-0.1378) * (V + T(40.14)))));
  const T tau_h = T(1) / (ah + bh), tau_j = T(1) / (aj + bj), tau_m = T(1) / (am + bm);
  const T m = S[1];
  T i_na;
  if constexpr (FAIL) {
    // Leaky gates, (x + l) / (1 + l), in the reference's order (Mahajan.cc:269).
    const T l = T(K::leak), ol = T(1) + l;
    i_na = T(12.0) * (S[2] + l) / ol * (S[3] + l) / ol * (m + l) / ol * (m + l) / ol * (m + l) / ol *
           (V - ena);
  } else {
    i_na = T(12.0) * S[2] * S[3] * m * m * m * (V - ena);
  }
  R[2] = mj_relax<T, FAST>(ah * tau_h, tau_h, S[2], h);
  R[3] = mj_relax<T, FAST>(aj * tau_j, tau_j, S[3], h);
  R[1] = mj_relax<T, FAST>(am * tau_m, tau_m, S[1], h);

  // I_Kr
  T kr1 = T(0.00138 / 0.123), kr2 = T(0.00061 / 0.145);
  if (fabs(V + T(7.0)) > T(0.001)) kr1 = T(0.00138) * (V + T(7.0)) / (T(1) - M::ex(T(-0.123) * (V + T(7.0))));
  if (fabs(V + T(10.0)) > T(0.001)) kr2 = T(0.00061) * (V + T(10.0)) / (M::ex(T(0.145) * (V + T(10.0))) - T(1));
  const T xr_inf = T(1) / (T(1) + M::ex(-(V + T(50.0)) / T(7.5)));
  const T rg = T(1) / (T(1) + M::ex((V + T(33.0)) / T(22.4)));
  const T i_kr = T(0.0125) * sq_ko * S[4] * rg * (V - ek);
  R[4] = mj_relax<T, FAST>(xr_inf, T(1) / (kr1 + kr2), S[4], h);

  // I_Ks with Ca dependence
  const T eks = M::lg((ko + T(0.018330) * nao) / (ki + T(0.018330) * nai)) / frt;
  const T xs_inf = T(1) / (T(1) + M::ex(-(V - T(1.50)) / T(16.70)));
  T tau_xs1;
  if (fabs(V + T(30.0)) < T(0.001 / 0.0687))
    tau_xs1 = T(1.0 / (0.0000719 / 0.148 + 0.000131 / 0.0687));
  else
    tau_xs1 = T(1) / (T(0.0000719) * (V + T(30.0)) / (T(1) - M::ex(T(-0.148) * (V + T(30.0)))) +
                      T(0.000131) * (V + T(30.0)) / (M::ex(T(0.0687) * (V + T(30.0))) - T(1)));
  const T rks = T(0.5) / ci;
  const T g_ksx = T(1) + T(0.8) / (T(1) + rks * rks * rks);
  const T i_ks = gks * g_ksx * S[5] * S[6] * (V - eks);
  R[5] = mj_relax<T, FAST>(xs_inf, tau_xs1, S[5], h);
  R[6] = mj_relax<T, FAST>(xs_inf, T(4.0) * tau_xs1, S[6], h);

  // Na_i and V
  const T wca = T(8.0313);
  R[19] = -(T(1.0 / 8.0313 / 1000.0)) * (i_na + T(3.0) * i_nak + T(3.0) * wca * j_naca);
  R[0] = -(i_na + i_k1 + i_kr + i_ks + i_tos + i_tof + wca * j_naca + T(2.0) * wca * j_ca + i_nak) +
         istim;
}

// Mahajan::Compute_Ion + UpdateStateVariables (Mahajan.cc:607-644) for one
// node: S holds all 26 states in T (S[0] is set to V), acc the FP64 copies of
// the MJ_ACC states, which the updates accumulate.  Forward Euler over dt, or
// 10 sub-steps of dt/10 when |dV/dt| > 20 (int_abs: when |trunc(dV/dt)| > 20);
// returns the last dV/dt evaluated.
// The fast FP32 form computes the rates in RcpF (reciprocal divisions).
template <class T, bool FAST, bool FAIL = false>
__device__ VG_INLINE T mj_step(T S[MJ_NSTATE], double acc[MJ_NACC], T V, T istim, T dt, T gtof,
                            T gks, bool int_abs) {
  using U = typename RateT<T, FAST>::type;
  U Su[MJ_NSTATE], R[MJ_NSTATE];
  S[0] = V;
#pragma unroll
  for (int k = 0; k < MJ_NSTATE; k++) Su[k] = U(S[k]);
  mj_rates<U, FAST, FAIL>(Su, U(dt), U(istim), U(gtof), U(gks), R);
  const T r0 = T(R[0]);
  const bool adapt = int_abs ? fabs(trunc(r0)) > T(20) : fabs(r0) > T(20);
  const int nsub = adapt ? 10 : 1;
  const T h = adapt ? dt / T(10) : dt;
  for (int q = 0; q < nsub; q++) {
    if (adapt) mj_rates<U, FAST, FAIL>(Su, U(h), U(istim), U(gtof), U(gks), R);
    Su[0] += R[0] * U(h);
#pragma unroll
    for (int k = 0; k < MJ_NGATE; k++) Su[MJ_GATE[k]] += R[MJ_GATE[k]] * U(h);
#pragma unroll
    for (int k = 0; k < MJ_NACC; k++) {
      acc[k] += (double)T(R[MJ_ACC[k]]) * (double)h;
      Su[MJ_ACC[k]] = U((T)acc[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < MJ_NSTATE; k++) S[k] = T(Su[k]);
  return T(R[0]);
}

}  // namespace vg
