// vg_kernels.hpp -- gfx950 kernels of the monodomain time step.
//
// Layout in HBM (n_alloc = nodes rounded up to 256, in the solver's row order):
//   SELL-64 operator: col[int32], A_hat[T], K_rhs[T] per padded entry; one lane
//     per row, 64 rows per slice, lane-fastest so every load is coalesced.
//   vectors V_a, V_b, r, w, s (two copies each), p, x, sinv, diag: [n_alloc] T
//   ionic states: SoA [nstate][n_alloc] T (V excluded)
//
// Reductions never cross blocks inside a launch.  Each block of a launch sums
// its values in FP64 (wavefront butterfly, then LDS over the block's waves)
// and stores one partial per value; every block of the NEXT launch sums those
// partials itself, in one fixed order, so all blocks hold bit-identical totals
// and take the same decisions, and results repeat bit for bit run to run.  On
// a multi-XCD device a release/acquire handshake between blocks of one launch
// costs a write-back and an invalidation of an XCD's whole L2 per block; the
// launch boundary does that once.
#pragma once
#include <hip/hip_runtime.h>

#include "vg_ionic.hpp"
#include "vg_surrogate.hpp"

namespace vg {

constexpr int BLK = 256;
constexpr int WAVES = BLK / 64;
constexpr int MAX_SUB = 32;
constexpr int MAX_ECG = 16;
// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; grids
// of the SpMV-type launches are a multiple of this.
constexpr int XCDS = 8;

struct SubSteps {
  int n;
  double dt[MAX_SUB];
};

// CG scalars carried from one iteration launch to the next; launch j reads
// carry[j & 1] and block 0 writes carry[(j + 1) & 1].
struct CgCarry {
  double rho_ref;  // ||b||_D^2: the stopping test is ||r||_D <= tol ||b||_D
  double gamma;    // (r_j, r_j) of the launch that wrote it
  double alpha;
};

// Outcome of one solve, written by block 0 of the launch that ends it.
struct CgResult {
  int iter;     // iterations done
  int done;
  int notconv;  // ended on maxit, breakdown or a non-finite residual
  int stop;     // set when the solve has ended; later launches return at once
  int x_valid;  // x holds the increment (0: the increment is zero)
  double rho, rho_ref;
};

// A launch may solve several independent systems (instances: separate
// bodies batched into one handle, each slice-aligned).  The blocks of
// instance k are [inst_blk[k], inst_blk[k+1]) and its slices
// [inst_slice[k], inst_slice[k+1]); every block reduces only its instance's
// partials, so each instance runs its own CG to its own convergence.
struct CgBufs {
  double *part0, *part1;      // [3][grid] block partials of alternating launches
  CgCarry *carry;             // [2][n_inst]
  CgResult *res;              // [n_inst]
  const int *blk_inst;        // [grid]
  const int *inst_blk;        // [n_inst + 1]
  const int64_t *inst_slice;  // [n_inst + 1]
  int n_inst;
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sum of NV values, stored as part[k * gridDim.x + blockIdx.x].
template <int NV>
__device__ inline void put_partials(const double (&v)[NV], double *__restrict__ part) {
  __shared__ double sm[WAVES][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sm[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) s += sm[w][threadIdx.x];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// Totals over blocks [b0, b1) of the partials that a previous launch of nblk
// blocks stored; the same value in every thread of every block that asks.
template <int NV>
__device__ inline void sum_partials(const double *__restrict__ part, int nblk, int b0, int b1,
                                    double (&tot)[NV]) {
  __shared__ double sm[WAVES][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += BLK) s += part[(size_t)k * nblk + b];
    s = wave_sum(s);
    if (lane == 0) sm[wv][k] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) s += sm[w][k];
    tot[k] = s;
  }
}

// Instance of this block, its block range [b0, b1) and its slices [lo, hi):
// the instance's slices cut into b1 - b0 contiguous chunks.  With a single
// instance over a multiple of XCDS blocks the chunks of the blocks one XCD
// runs are adjacent, so the rows an XCD gathers from mostly sit in its own L2.
__device__ inline int block_work(const CgBufs &cb, int64_t &lo, int64_t &hi, int &b0, int &b1) {
  const int b = blockIdx.x, k = cb.blk_inst[b];
  b0 = cb.inst_blk[k];
  b1 = cb.inst_blk[k + 1];
  const int64_t g = b1 - b0, s0 = cb.inst_slice[k], ns = cb.inst_slice[k + 1] - s0;
  int64_t l = b - b0;
  if (cb.n_inst == 1 && g % XCDS == 0) l = (l % XCDS) * (g / XCDS) + l / XCDS;
  lo = s0 + ns * l / g;
  hi = s0 + ns * (l + 1) / g;
  return k;
}

// SELL-64 operator pattern (values are separate arrays on the same layout).
// Slice s: width[s] entries per lane, lane-fastest from off[s]; columns as
// 16-bit offsets from the row in col16 (same layout), or, for a slice with a
// column further than 32767 rows away (wide[s] >= 0), as int32 in col32 from
// wide[s].
struct Sell {
  const int64_t *off;
  const int *width;
  const int64_t *wide;
  const int16_t *col16;
  const int *col32;
};

// sum over j < wd of v(j) * gather(col(j)) in order j = 0, 1, ...; loads go in
// groups of 4 (then the 0-3 left over), all column and value loads of a
// group issued before its gathers, so each wave keeps a dozen independent
// loads in flight instead of one dependent chain.
template <class T, class CF, class VF, class G>
__device__ inline T sell_dot(int wd, CF colat, VF valat, G gather) {
  T acc = T(0);
  int j = 0;
  for (; j + 4 <= wd; j += 4) {
    int c[4];
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      c[q] = colat(j + q);
      v[q] = valat(j + q);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) acc += v[q] * gather(c[q]);
  }
  const int k = wd - j;
  if (k > 0) {
    int c[3];
    T v[3];
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q < k) {
        c[q] = colat(j + q);
        v[q] = valat(j + q);
      }
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q < k) acc += v[q] * gather(c[q]);
  }
  return acc;
}

// One lane's row of a slice: sum_j val[j] * gather(col[j]).
template <class T, class G>
__device__ inline T sell_row(const Sell &S, const T *__restrict__ val, int64_t slice, int lane,
                             G gather) {
  const int64_t o = S.off[slice] + lane;
  const int wd = S.width[slice];
  const int64_t wo = S.wide[slice];
  const T *__restrict__ vp = val + o;
  auto valat = [&](int j) { return vp[j * 64]; };
  if (wo < 0) {
    const int row = (int)(slice * 64 + lane);
    const int16_t *__restrict__ cp = S.col16 + o;
    return sell_dot<T>(wd, [&](int j) { return row + (int)cp[j * 64]; }, valat, gather);
  }
  const int *__restrict__ cp = S.col32 + wo + lane;
  return sell_dot<T>(wd, [&](int j) { return cp[j * 64]; }, valat, gather);
}

// Two operators on one pattern in one pass (columns loaded once):
// acc_a = sum va[j] ga(col[j]), acc_b = sum vb[j] gb(col[j]).
template <class T, class GA, class GB>
__device__ inline void sell_row2(const Sell &S, const T *__restrict__ va, const T *__restrict__ vb,
                                 int64_t slice, int lane, GA ga, GB gb, T &acc_a, T &acc_b) {
  const int64_t o = S.off[slice] + lane;
  const int wd = S.width[slice];
  const int64_t wo = S.wide[slice];
  const T *__restrict__ ap = va + o;
  const T *__restrict__ bp = vb + o;
  // Both sums share the column loads: gather a pair, accumulate both.
  struct P {
    T a, b;
  };
  auto pairdot = [&](auto colat) {
    T sa = T(0), sb = T(0);
    int j = 0;
    for (; j < wd; j += 4) {
      int c[4];
      T x[4], y[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (j + q < wd) {
          c[q] = colat(j + q);
          x[q] = ap[(j + q) * 64];
          y[q] = bp[(j + q) * 64];
        }
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (j + q < wd) {
          sa += x[q] * ga(c[q]);
          sb += y[q] * gb(c[q]);
        }
    }
    return P{sa, sb};
  };
  P r;
  if (wo < 0) {
    const int row = (int)(slice * 64 + lane);
    const int16_t *__restrict__ cp = S.col16 + o;
    r = pairdot([&](int j) { return row + (int)cp[j * 64]; });
  } else {
    const int *__restrict__ cp = S.col32 + wo + lane;
    r = pairdot([&](int j) { return cp[j * 64]; });
  }
  acc_a = r.a;
  acc_b = r.b;
}

__device__ inline bool cg_stopped(const CgResult *res) {
  return __hip_atomic_load(&res->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ inline void cg_set_stop(CgResult *res) {
  __hip_atomic_store(&res->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- diffusion right-hand side ---------------------------------------------
// b = K_rhs V and r_0 = b - A_hat x_0 (x_0 = 0 unless GUESS: then x holds the
// increment of the previous solve of this half-step); partials of ||b||_D^2
// and ||r_0||_D^2.  Opens a solve: clears the result record (no CG launch of
// the previous solve runs concurrently: launches on one stream are ordered).
template <class T, bool GUESS>
__global__ __launch_bounds__(BLK) void k_rhs(Sell S, const T *__restrict__ Krhs,
                                             const T *__restrict__ Ahat,
                                             const T *__restrict__ V, const T *__restrict__ x,
                                             T *__restrict__ r, const T *__restrict__ diag,
                                             CgBufs cb) {
  int64_t lo, hi;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, b0, b1);
  if (blockIdx.x == b0 && threadIdx.x == 0) {
    CgResult *res = cb.res + inst;
    res->iter = 0;
    res->done = 0;
    res->notconv = 0;
    res->x_valid = 0;
    res->rho = res->rho_ref = 0.0;
    __hip_atomic_store(&res->stop, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const int lane = threadIdx.x & 63;
  double v[2] = {0.0, 0.0};
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += WAVES) {
    const int64_t row = sl * 64 + lane;
    const double d = (double)diag[row];
    if constexpr (GUESS) {
      T bk, ax;
      sell_row2<T>(S, Krhs, Ahat, sl, lane, [&](int c) { return V[c]; },
                   [&](int c) { return x[c]; }, bk, ax);
      const T rr = bk - ax;
      r[row] = rr;
      v[0] += d * (double)bk * (double)bk;
      v[1] += d * (double)rr * (double)rr;
    } else {
      const T bk = sell_row<T>(S, Krhs, sl, lane, [&](int c) { return V[c]; });
      r[row] = bk;
      const double e = d * (double)bk * (double)bk;
      v[0] += e;
      v[1] += e;
    }
  }
  put_partials<2>(v, cb.part1);
}

// ---- CG (Chronopoulos-Gear form, Jacobi scaling folded into A_hat) ---------
// Start: ||b||_D^2 from the right-hand side's partials (b = 0: the increment
// is zero); w_0 = A_hat r_0 and partials of (r_0, r_0), (r_0, w_0), ||r_0||_D^2.
template <class T>
__global__ __launch_bounds__(BLK) void k_cg_start(Sell S, const T *__restrict__ val,
                                                  const T *__restrict__ r, T *__restrict__ w,
                                                  const T *__restrict__ diag, CgBufs cb) {
  int64_t lo, hi;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, b0, b1);
  CgResult *res = cb.res + inst;
  double t0[2];
  sum_partials<2>(cb.part1, gridDim.x, b0, b1, t0);
  const double rho_ref = t0[0];
  const bool first = blockIdx.x == b0 && threadIdx.x == 0;
  if (!(rho_ref > 0.0) || !isfinite(rho_ref)) {
    if (first) {
      res->rho_ref = res->rho = rho_ref;
      res->done = 1;
      res->notconv = rho_ref == 0.0 ? 0 : 1;
      cg_set_stop(res);
    }
    return;
  }
  const int lane = threadIdx.x & 63;
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += WAVES) {
    const T acc = sell_row<T>(S, val, sl, lane, [&](int c) { return r[c]; });
    const int64_t row = sl * 64 + lane;
    w[row] = acc;
    const double rr = (double)r[row];
    v[0] += rr * rr;
    v[1] += rr * (double)acc;
    v[2] += (double)diag[row] * rr * rr;
  }
  put_partials<3>(v, cb.part0);
  if (first) {
    cb.carry[inst] = CgCarry{rho_ref, 0.0, 0.0};
    res->rho_ref = rho_ref;
    res->rho = t0[1];
  }
}

// Iteration j in one launch.  From the partials of launch j-1 every block
// forms gamma_j = (r_j, r_j), delta_j = (r_j, w_j), rho_j = ||r_j||_D^2, tests
// rho_j <= tol^2 ||b||_D^2 and forms alpha_j, beta_j; then each row updates
//   p_j = r_j + beta_j p_{j-1},  s_j = w_j + beta_j s_{j-1},
//   x_{j+1} = x_j + alpha_j p_j,  r_{j+1} = r_j - alpha_j s_j,
// and w_{j+1} = A_hat r_{j+1}, rebuilding the neighbours' r_{j+1} from their
// r_j, w_j, s_{j-1} (ping-pong buffers by parity of j), so no other block's
// writes of this launch are read.
template <class T, bool FIRST>
__device__ inline T cg_iter_row(const Sell &S, const T *__restrict__ val,
                                const T *__restrict__ r0, const T *__restrict__ w0,
                                const T *__restrict__ s0, int64_t slice, int lane, T alpha,
                                T beta) {
  return sell_row<T>(S, val, slice, lane, [&](int c) {
    const T sc = FIRST ? w0[c] : w0[c] + beta * s0[c];
    return r0[c] - alpha * sc;
  });
}

template <class T>
__global__ __launch_bounds__(BLK) void k_cg_iter(Sell S, const T *__restrict__ val, T *rA, T *rB, T *wA,
                                                 T *wB, T *sA, T *sB, T *__restrict__ p,
                                                 T *__restrict__ x, const T *__restrict__ diag,
                                                 CgBufs cb, int j, double tol2, int maxit,
                                                 int guess) {
  CgResult *res = cb.res + cb.blk_inst[blockIdx.x];
  if (cg_stopped(res)) return;
  int64_t lo, hi;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, b0, b1);
  const int par = j & 1;
  double t[3];
  sum_partials<3>(par ? cb.part1 : cb.part0, gridDim.x, b0, b1, t);
  const CgCarry c = cb.carry[par * cb.n_inst + inst];
  const double gamma = t[0], delta = t[1], rho = t[2];
  int end = 0;  // 1: converged, 2: not converged
  double alpha = 0.0, beta = 0.0;
  if (rho <= tol2 * c.rho_ref) {
    end = 1;
  } else if (j >= maxit || !isfinite(rho)) {
    end = 2;
  } else {
    if (j == 0) {
      alpha = gamma / delta;
    } else {
      beta = gamma / c.gamma;
      alpha = gamma / (delta - beta * gamma / c.alpha);
    }
    if (!(alpha > 0.0) || !isfinite(alpha)) end = 2;  // breakdown
  }
  if (end) {
    if (blockIdx.x == b0 && threadIdx.x == 0) {
      res->iter = j;
      res->done = 1;
      res->notconv = end == 2 ? 1 : 0;
      res->x_valid = (j > 0 || guess) ? 1 : 0;
      res->rho = rho;
      cg_set_stop(res);
    }
    return;
  }
  const T *r0 = par ? rB : rA, *w0 = par ? wB : wA, *s0 = par ? sB : sA;
  T *r1 = par ? rA : rB, *w1 = par ? wA : wB, *s1 = par ? sA : sB;
  const T a = (T)alpha, bt = (T)beta;
  const int lane = threadIdx.x & 63;
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += WAVES) {
    const int64_t row = sl * 64 + lane;
    const T ri = r0[row], wi = w0[row];
    T pi, si, xi, acc;
    if (j == 0) {
      pi = ri;
      si = wi;
      xi = guess ? x[row] + a * pi : a * pi;
      acc = cg_iter_row<T, true>(S, val, r0, w0, s0, sl, lane, a, bt);
    } else {
      pi = ri + bt * p[row];
      si = wi + bt * s0[row];
      xi = x[row] + a * pi;
      acc = cg_iter_row<T, false>(S, val, r0, w0, s0, sl, lane, a, bt);
    }
    const T rn = ri - a * si;
    r1[row] = rn;
    w1[row] = acc;
    s1[row] = si;
    p[row] = pi;
    x[row] = xi;
    const double rd = (double)rn;
    v[0] += rd * rd;
    v[1] += rd * (double)acc;
    v[2] += (double)diag[row] * rd * rd;
  }
  put_partials<3>(v, par ? cb.part0 : cb.part1);
  if (blockIdx.x == b0 && threadIdx.x == 0) {
    cb.carry[(par ^ 1) * cb.n_inst + inst] = CgCarry{c.rho_ref, gamma, alpha};
    res->iter = j + 1;
  }
}

// ---- ionic update (fused with V = V_a + D^-1/2 x of the first half-step) ---
struct IonicArgs {
  int64_t n, n_alloc;
  int nsets;
  double Cm;
  int max_newton;
  int newton_jt;
  int mj_int_abs;
  // ten Tusscher: record the update of every rec_stride-th row into samples
  // rec_base + row / rec_stride of rec_x[.][25], rec_y[.][19] (off when
  // rec_x is null; samples at or past rec_cap are dropped).
  float *rec_x, *rec_y;
  int64_t rec_base, rec_cap;
  int rec_stride;
  float log_dt;
  // Update only the rows flagged here (the learned step's fallback), or all.
  const unsigned char *only;
};

template <class T, int MODEL, int INTEG>
__global__ __launch_bounds__(BLK) void k_ionic(const T *__restrict__ Vin, const T *__restrict__ x,
                                               const T *__restrict__ sinv, T *__restrict__ Vout,
                                               T *__restrict__ st, double *__restrict__ acc,
                                               const T *__restrict__ istim,
                                               const uint16_t *__restrict__ pidx,
                                               const T *__restrict__ scales, SubSteps sub,
                                               IonicArgs a, unsigned long long *newton_sum,
                                               const CgResult *cg,
                                               const int *__restrict__ slice_inst) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int newton = 0;
  if (i < a.n && (!a.only || a.only[i])) {
    // x holds the half-step increment unless that increment is zero.
    const bool use_x = x && cg[slice_inst ? slice_inst[i >> 6] : 0].x_valid;
    T V = Vin[i];
    if (use_x) V += sinv[i] * x[i];
    const T stim = istim ? istim[i] : T(0);
    const int set = pidx ? (int)pidx[i] : 0;
    const T *sc = scales + set * VG_TT_NSCALE;
    const T Cm = (T)a.Cm;
    if constexpr (MODEL == VG_MODEL_LR) {
      T s[LR_NSTATE];
#pragma unroll
      for (int k = 0; k < LR_NSTATE - 1; k++) s[k] = st[k * a.n_alloc + i];
      for (int q = 0; q < sub.n; q++) {
        const T dt = (T)sub.dt[q];
        const T dv = lr_compute_ion<T, sizeof(T) == 4>(s, Cm, dt, V, stim, sc[0], sc[3]);
        V += dt * dv;
      }
#pragma unroll
      for (int k = 0; k < LR_NSTATE - 1; k++) st[k * a.n_alloc + i] = s[k];
    } else if constexpr (MODEL == VG_MODEL_MJ) {
      // Gates and channel states in T (rows of st), Ca-cycling and Na_i in
      // FP64 (rows of acc); sc[0], sc[1] = G_tof, G_Ks of the node's type.
      T s[MJ_NSTATE];
      double ac[MJ_NACC];
#pragma unroll
      for (int k = 0; k < MJ_NGATE; k++) s[MJ_GATE[k]] = st[k * a.n_alloc + i];
#pragma unroll
      for (int k = 0; k < MJ_NACC; k++) {
        ac[k] = acc[k * a.n_alloc + i];
        s[MJ_ACC[k]] = (T)ac[k];
      }
      const T it = stim / (Cm * T(MJ_XI));
      for (int q = 0; q < sub.n; q++) {
        const T dt = (T)sub.dt[q];
        V += dt * mj_step<T, sizeof(T) == 4>(s, ac, V, it, dt, sc[0], sc[1], a.mj_int_abs != 0);
      }
#pragma unroll
      for (int k = 0; k < MJ_NGATE; k++) st[k * a.n_alloc + i] = s[MJ_GATE[k]];
#pragma unroll
      for (int k = 0; k < MJ_NACC; k++) acc[k * a.n_alloc + i] = ac[k];
    } else {
      T s[TT_NSTATE];
#pragma unroll
      for (int k = 1; k < TT_NSTATE; k++) s[k] = st[(k - 1) * a.n_alloc + i];
      const TTG<T> g = tt_conductances<T>(sc);
      const T it = -stim / (Cm * T(TT_XI));
      float *rx = nullptr, *ry = nullptr;
      if (a.rec_x && i % a.rec_stride == 0) {
        const int64_t k = a.rec_base + i / a.rec_stride;
        if (k < a.rec_cap) {
          rx = a.rec_x + k * SG_NIN;
          ry = a.rec_y + k * SG_NOUT;
          rx[0] = (float)V;
#pragma unroll
          for (int q = 1; q < TT_NSTATE; q++) rx[q] = (float)s[q];
          rx[19] = a.log_dt;
#pragma unroll
          for (int q = 0; q < VG_TT_NSCALE; q++) rx[20 + q] = (float)sc[q];
        }
      }
      for (int q = 0; q < sub.n; q++) {
        const T dt = (T)sub.dt[q];
        T dv;
        if constexpr (INTEG == VG_INTEG_REFERENCE) {
          int nit = 0;
          dv = tt_step_reference<T>(s, V, it, dt, g, a.max_newton, a.newton_jt != 0, &nit);
          newton += nit;
        } else {
          dv = tt_step_rl<T, sizeof(T) == 4>(s, V, it, dt, g);
        }
        V += dt * dv;
      }
      if (ry) {
        // Increments from the unrounded start values (still in memory).
        T V0 = Vin[i];
        if (use_x) V0 += sinv[i] * x[i];
        ry[0] = (float)(V - V0);
#pragma unroll
        for (int q = 1; q < TT_NSTATE; q++) ry[q] = (float)(s[q] - st[(q - 1) * a.n_alloc + i]);
      }
#pragma unroll
      for (int k = 1; k < TT_NSTATE; k++) st[(k - 1) * a.n_alloc + i] = s[k];
    }
    Vout[i] = V;
  }
  if constexpr (INTEG == VG_INTEG_REFERENCE) {
    if (newton_sum) {
      double v[1] = {(double)newton};
      const int lane = threadIdx.x & 63;
      const double s = wave_sum(v[0]);
      if (lane == 0 && s != 0.0) atomicAdd(newton_sum, (unsigned long long)s);
    }
  }
}

// Purkinje segment: rows [row0, row0 + n) of the solver order, states SoA with
// n_alloc lanes per state (all 22, the model keeps its own V copy).
template <class T>
__global__ __launch_bounds__(BLK) void k_ionic_pk(const T *__restrict__ Vin, const T *__restrict__ x,
                                                  const T *__restrict__ sinv, T *__restrict__ Vout,
                                                  T *__restrict__ st, const T *__restrict__ istim,
                                                  int64_t row0, int64_t n, int64_t n_alloc,
                                                  SubSteps sub, const CgResult *cg,
                                                  const int *__restrict__ slice_inst) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int64_t row = row0 + i;
  T V = Vin[row];
  if (x && cg[slice_inst ? slice_inst[row >> 6] : 0].x_valid) V += sinv[row] * x[row];
  const T stim = istim ? istim[row] : T(0);
  T s[PK_NSTATE];
#pragma unroll
  for (int k = 0; k < PK_NSTATE; k++) s[k] = st[k * n_alloc + i];
  for (int q = 0; q < sub.n; q++) {
    const T dt = (T)sub.dt[q];
    V += dt * pk_step<T, sizeof(T) == 4>(s, V, stim, dt);
  }
#pragma unroll
  for (int k = 0; k < PK_NSTATE; k++) st[k * n_alloc + i] = s[k];
  Vout[row] = V;
}

// ---- learned ten Tusscher ionic update ---------------------------------------
// Lanes of one wave exchange activations through the wave's own LDS buffer.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One hidden or output layer for the wave's 32 rows: A (the wave's activation
// rows, BF16, SG_ACT_LD apart) times the layer's transposed weights in LDS,
// NT output tiles of 32 columns accumulated in FP32.
template <int KS, int NT>
__device__ inline void surr_layer(const __bf16 *A, const __bf16 *W, int ldw, int r, int kq,
                                  f32x16 (&acc)[4]) {
  bf16x8 a[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ks++) a[ks] = *reinterpret_cast<const bf16x8 *>(A + r * SG_ACT_LD + ks * 16 + kq);
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    acc[nt] = f32x16{};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const bf16x8 b = *reinterpret_cast<const bf16x8 *>(W + (nt * 32 + r) * ldw + ks * 16 + kq);
      acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], b, acc[nt], 0, 0, 0);
    }
  }
}

// Bias + SiLU of a hidden layer's outputs back into the activation rows.
__device__ inline void surr_hidden_out(__bf16 *A, const f32x16 (&acc)[4], const float *b, int l,
                                       int r) {
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const int col = nt * 32 + r;
    const float bc = b[col];
#pragma unroll
    for (int j = 0; j < 16; j++) A[mfma_row(l, j) * SG_ACT_LD + col] = (__bf16)sg_silu(acc[nt][j] + bc);
  }
}

// ten Tusscher ionic update through the learned step for rows [0, n): a row
// whose 25 inputs and 19 outputs lie inside the training envelopes and that
// has no stimulus gets V and its 18 states advanced by the network's
// increments (gates and R' clamped to [0, 1]); the others are flagged in
// `fallback` for the plain kernel.  Persistent blocks keep all weights in LDS; each wave takes tiles
// of 32 rows, lane r and r + 32 sharing row r (the MFMA operand layout).
// usage[2 * (block * 4 + wave) + {0, 1}] counts network / fallback rows.
template <class T>
__global__ __launch_bounds__(BLK) void k_ionic_surr(
    const T *__restrict__ Vin, const T *__restrict__ x, const T *__restrict__ sinv,
    T *__restrict__ Vout, T *__restrict__ st, const T *__restrict__ istim,
    const uint16_t *__restrict__ pidx, const T *__restrict__ scales, int64_t n, int64_t n_alloc,
    float log_dt, SurrDevice sd, unsigned char *__restrict__ fallback,
    unsigned long long *__restrict__ usage, const CgResult *cg, const int *__restrict__ slice_inst) {
  __shared__ __attribute__((aligned(16))) __bf16 wl[SG_WT_TOTAL];
  __shared__ __attribute__((aligned(16))) __bf16 actb[WAVES][32 * SG_ACT_LD];
  for (int q = threadIdx.x; q < SG_WT_TOTAL / 8; q += BLK)
    reinterpret_cast<bf16x8 *>(wl)[q] = reinterpret_cast<const bf16x8 *>(sd.wt)[q];
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 31, kq = (l >> 5) * 8;
  __bf16 *A = actb[wave];
  float *Y = reinterpret_cast<float *>(A);  // output staging [32][33]
  unsigned long long n_net = 0, n_fb = 0;
  const int64_t ntiles = (n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * WAVES + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * WAVES) {
    const int64_t i = tile * 32 + r;
    const bool mine = l < 32 && i < n;
    T s[TT_NSTATE];
    bool ok = false;
    if (mine) {
      T V = Vin[i];
      if (x && cg[slice_inst ? slice_inst[i >> 6] : 0].x_valid) V += sinv[i] * x[i];
      s[0] = V;
#pragma unroll
      for (int k = 1; k < TT_NSTATE; k++) s[k] = st[(k - 1) * n_alloc + i];
      const T *sc = scales + (pidx ? (int)pidx[i] : 0) * VG_TT_NSCALE;
      float f[SG_NIN];
#pragma unroll
      for (int k = 0; k < TT_NSTATE; k++) f[k] = (float)s[k];
      f[19] = log_dt;
#pragma unroll
      for (int k = 0; k < VG_TT_NSCALE; k++) f[20 + k] = (float)sc[k];
      ok = !(istim && istim[i] != T(0));
#pragma unroll
      for (int k = 0; k < SG_NIN; k++) ok = ok && f[k] >= sd.lo[k] && f[k] <= sd.hi[k];
#pragma unroll
      for (int k = 0; k < SG_KIN; k++)
        A[r * SG_ACT_LD + k] = (__bf16)(ok && k < SG_NIN ? (f[k] - sd.mu_x[k]) / sd.sd_x[k] : 0.0f);
      fallback[i] = ok ? 0 : 1;
    } else if (l < 32) {
#pragma unroll
      for (int k = 0; k < SG_KIN; k++) A[r * SG_ACT_LD + k] = (__bf16)0.0f;
    }
    wave_lds_sync();
    f32x16 acc[4];
    surr_layer<2, 4>(A, wl + sg_wt_off(0), sg_wt_ld(0), r, kq, acc);
    wave_lds_sync();
    surr_hidden_out(A, acc, sd.b[0], l, r);
    wave_lds_sync();
#pragma unroll 1
    for (int L = 1; L < SG_LAYERS - 1; L++) {
      surr_layer<8, 4>(A, wl + sg_wt_off(L), sg_wt_ld(L), r, kq, acc);
      wave_lds_sync();
      surr_hidden_out(A, acc, sd.b[L], l, r);
      wave_lds_sync();
    }
    surr_layer<8, 1>(A, wl + sg_wt_off(SG_LAYERS - 1), sg_wt_ld(SG_LAYERS - 1), r, kq, acc);
    wave_lds_sync();
    if (r < SG_NOUT) {
      const float bc = sd.b[SG_LAYERS - 1][r];
#pragma unroll
      for (int j = 0; j < 16; j++) Y[mfma_row(l, j) * 33 + r] = acc[0][j] + bc;
    }
    wave_lds_sync();
    if (mine && ok) {
      // Outputs outside their envelope (or not finite) leave the node to
      // the plain model too.
      float y[SG_NOUT];
#pragma unroll
      for (int k = 0; k < SG_NOUT; k++) {
        const float t = Y[r * 33 + k];
        ok = ok && t >= sd.t_lo[k] && t <= sd.t_hi[k];
        y[k] = sg_from_target(t, sd.mu_y[k], sd.sd_y[k], sd.c_y[k]);
      }
      if (ok) {
        Vout[i] = s[0] + (T)y[0];
#pragma unroll
        for (int k = 1; k < TT_NSTATE; k++) {
          T v = s[k] + (T)y[k];
          if ((k >= 4 && k <= 9) || (k >= 11 && k <= 16) || k == 18) v = fmin(fmax(v, T(0)), T(1));
          st[(k - 1) * n_alloc + i] = v;
        }
        n_net++;
      } else {
        fallback[i] = 1;
      }
    }
    if (mine && !ok) n_fb++;
    wave_lds_sync();
  }
  for (int o = 32; o > 0; o >>= 1) {
    n_net += __shfl_xor(n_net, o, 64);
    n_fb += __shfl_xor(n_fb, o, 64);
  }
  if (l == 0) {
    usage[2 * (blockIdx.x * WAVES + wave)] += n_net;
    usage[2 * (blockIdx.x * WAVES + wave) + 1] += n_fb;
  }
}

// ---- second half-step apply: V_a = V_b + D^-1/2 x, activation, ECG ---------
// Grid-stride over nodes; with ECG each block stores MAX_ECG partials of
// sum_n H[e][n] V_n, finished by k_ecg_finish.
// The step's start time is step_ctr[0] * dt (the counter k_fold advances).
template <class T, bool AT, bool ECG>
__global__ __launch_bounds__(BLK) void k_apply(const T *__restrict__ Vb, const T *__restrict__ x,
                                               const T *__restrict__ sinv, T *__restrict__ Va,
                                               int64_t n, T *__restrict__ at, T thr,
                                               const int64_t *__restrict__ step_ctr,
                                               double dt, const T *__restrict__ H, int ne,
                                               int64_t n_alloc, double *__restrict__ part,
                                               const CgResult *cg,
                                               const int *__restrict__ slice_inst) {
  const double t0 = (double)step_ctr[0] * dt;
  double v[MAX_ECG];
#pragma unroll
  for (int e = 0; e < MAX_ECG; e++) v[e] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const bool use_x = cg[slice_inst ? slice_inst[i >> 6] : 0].x_valid != 0;
    const T Vn = use_x ? Vb[i] + sinv[i] * x[i] : Vb[i];
    if constexpr (AT) {
      const T Vo = Va[i];
      if (at[i] < T(0) && Vo < thr && Vn >= thr)
        at[i] = (T)(t0 + dt * (double)((thr - Vo) / (Vn - Vo)));
    }
    Va[i] = Vn;
    if constexpr (ECG) {
#pragma unroll
      for (int e = 0; e < MAX_ECG; e++)
        if (e < ne) v[e] += (double)H[(int64_t)e * n_alloc + i] * (double)Vn;
    }
  }
  if constexpr (ECG) put_partials<MAX_ECG>(v, part);
}

// Running totals of the solves of a captured step (read by the host lazily).
struct CgTotals {
  unsigned long long iters, solves, notconv, nonfinite;
  int max_iter[2];
  double max_relres;
};

// One thread: folds the result records of the solve of half-step `which`
// into the totals (fold != 0) and, with tick != 0, advances the step counter.
__global__ void k_fold(const CgResult *__restrict__ res, int n_inst, CgTotals *tot, int which,
                       int fold, int64_t *step_ctr, int tick) {
  if (threadIdx.x != 0) return;
  if (fold) {
    for (int k = 0; k < n_inst; k++) {
      const CgResult c = res[k];
      tot->iters += (unsigned long long)c.iter;
      tot->solves += 1;
      tot->notconv += (unsigned long long)(c.done == 0 || c.notconv != 0);
      if (!isfinite(c.rho) || !isfinite(c.rho_ref)) tot->nonfinite += 1;
      if (c.iter > tot->max_iter[which]) tot->max_iter[which] = c.iter;
      const double rel = c.rho_ref > 0.0 ? sqrt(c.rho / c.rho_ref) : 0.0;
      if (rel > tot->max_relres) tot->max_relres = rel;
    }
  }
  if (tick) step_ctr[0] += 1;
}

// ECG of V alone (partials as k_apply's).
template <class T>
__global__ __launch_bounds__(BLK) void k_ecg(const T *__restrict__ V, int64_t n,
                                             const T *__restrict__ H, int ne, int64_t n_alloc,
                                             double *__restrict__ part) {
  double v[MAX_ECG];
#pragma unroll
  for (int e = 0; e < MAX_ECG; e++) v[e] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const double Vi = (double)V[i];
#pragma unroll
    for (int e = 0; e < MAX_ECG; e++)
      if (e < ne) v[e] += (double)H[(int64_t)e * n_alloc + i] * Vi;
  }
  put_partials<MAX_ECG>(v, part);
}

// One block: out[e] = sum of the nblk partials of lead e.
__global__ __launch_bounds__(BLK) void k_ecg_finish(const double *__restrict__ part, int nblk,
                                                    int ne, double *__restrict__ out) {
  double tot[MAX_ECG];
  sum_partials<MAX_ECG>(part, nblk, 0, nblk, tot);
  if (threadIdx.x == 0)
    for (int e = 0; e < ne; e++) out[e] = tot[e];
}

// ---- independent cells (IonicModel adapter) --------------------------------
template <class T, int MODEL, int INTEG>
__global__ __launch_bounds__(BLK) void k_cells(const double *__restrict__ Vd,
                                               const double *__restrict__ stim_d,
                                               double *__restrict__ dvdt, T *__restrict__ st,
                                               double *__restrict__ acc,
                                               const uint16_t *__restrict__ pidx,
                                               const T *__restrict__ scales, int64_t n,
                                               int64_t n_alloc, double Cm_d, double dt_d,
                                               int max_newton, int newton_jt, int mj_int_abs) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const T V = (T)Vd[i], stim = stim_d ? (T)stim_d[i] : T(0), Cm = (T)Cm_d, dt = (T)dt_d;
  const int set = pidx ? (int)pidx[i] : 0;
  const T *sc = scales + set * VG_TT_NSCALE;
  if constexpr (MODEL == VG_MODEL_LR) {
    T s[LR_NSTATE];
#pragma unroll
    for (int k = 0; k < LR_NSTATE - 1; k++) s[k] = st[k * n_alloc + i];
    dvdt[i] = (double)lr_compute_ion<T, sizeof(T) == 4>(s, Cm, dt, V, stim, sc[0], sc[3]);
#pragma unroll
    for (int k = 0; k < LR_NSTATE - 1; k++) st[k * n_alloc + i] = s[k];
  } else if constexpr (MODEL == VG_MODEL_MJ) {
    T s[MJ_NSTATE];
    double ac[MJ_NACC];
#pragma unroll
    for (int k = 0; k < MJ_NGATE; k++) s[MJ_GATE[k]] = st[k * n_alloc + i];
#pragma unroll
    for (int k = 0; k < MJ_NACC; k++) {
      ac[k] = acc[k * n_alloc + i];
      s[MJ_ACC[k]] = (T)ac[k];
    }
    dvdt[i] = (double)mj_step<T, sizeof(T) == 4>(s, ac, V, stim / (Cm * T(MJ_XI)), dt, sc[0],
                                                 sc[1], mj_int_abs != 0);
#pragma unroll
    for (int k = 0; k < MJ_NGATE; k++) st[k * n_alloc + i] = s[MJ_GATE[k]];
#pragma unroll
    for (int k = 0; k < MJ_NACC; k++) acc[k * n_alloc + i] = ac[k];
  } else if constexpr (MODEL == VG_MODEL_PURKINJE) {
    // All 22 states are kept: the model advances its own V copy internally.
    T s[PK_NSTATE];
#pragma unroll
    for (int k = 0; k < PK_NSTATE; k++) s[k] = st[k * n_alloc + i];
    dvdt[i] = (double)pk_step<T, sizeof(T) == 4>(s, V, stim, dt);
#pragma unroll
    for (int k = 0; k < PK_NSTATE; k++) st[k * n_alloc + i] = s[k];
  } else {
    T s[TT_NSTATE];
#pragma unroll
    for (int k = 1; k < TT_NSTATE; k++) s[k] = st[(k - 1) * n_alloc + i];
    const TTG<T> g = tt_conductances<T>(sc);
    const T it = -stim / (Cm * T(TT_XI));
    T dv;
    if constexpr (INTEG == VG_INTEG_REFERENCE) {
      int nit = 0;
      dv = tt_step_reference<T>(s, V, it, dt, g, max_newton, newton_jt != 0, &nit);
    } else {
      dv = tt_step_rl<T, sizeof(T) == 4>(s, V, it, dt, g);
    }
    dvdt[i] = (double)dv;
#pragma unroll
    for (int k = 1; k < TT_NSTATE; k++) st[(k - 1) * n_alloc + i] = s[k];
  }
}

}  // namespace vg
