// vg_kernels.hpp -- gfx950 kernels of the monodomain time step.
//
// Layout in HBM (n_alloc = nodes rounded up to 256, in the solver's row order):
//   SELL-64 operator: col[int32], A_hat[T], K_rhs[T] per padded entry; one lane
//     per row, 64 rows per slice, lane-fastest so every load is coalesced.
//   vectors V_a, V_b, r, w, p, s, x, sinv, diag: [n_alloc] T
//   ionic states: SoA [nstate][n_alloc] T (V excluded)
// Reductions are FP64: wavefront butterflies, LDS across the 4 waves of a
// block, then one FP64 partial per block and a fixed-order sum in the last
// block to finish (bit-reproducible run to run).
#pragma once
#include <hip/hip_runtime.h>

#include "vg_ionic.hpp"

namespace vg {

constexpr int BLK = 256;
constexpr int WAVES = BLK / 64;
constexpr int MAX_SUB = 32;
constexpr int MAX_ECG = 16;

struct CgState {
  double alpha, beta, gamma_old, alpha_old;
  double rho0, rho, tol2;
  int iter, maxit, done, first, upd_first, notconv;
};

struct SubSteps {
  int n;
  double dt[MAX_SUB];
};

struct Reduce {
  double *part;       // [gridDim.x * NV]
  unsigned *counter;  // zero between launches
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sums NV values over the block, writes the block partial, and returns true
// in every thread of the block that finishes last; in that block tot[] holds
// the grid totals, summed in block order.
template <int NV>
__device__ inline bool grid_reduce(double (&v)[NV], Reduce red, double (&tot)[NV]) {
  __shared__ double sm[WAVES][NV];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sm[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
      for (int w = 0; w < WAVES; w++) s += sm[w][k];
      red.part[(size_t)blockIdx.x * NV + k] = s;
    }
    __threadfence();
    const unsigned prev = atomicAdd(red.counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  // Fixed-order sum: thread t takes blocks t, t+BLK, ...; then a tree.
  __shared__ double st[BLK];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += BLK)
      s += __builtin_nontemporal_load(&red.part[(size_t)b * NV + k]);
    st[threadIdx.x] = s;
    __syncthreads();
    for (int o = BLK / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) st[threadIdx.x] += st[threadIdx.x + o];
      __syncthreads();
    }
    tot[k] = st[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) *red.counter = 0u;
  return true;
}

__device__ inline int load_done(const CgState *cg) {
  return __hip_atomic_load(&cg->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- diffusion right-hand side: r = K_rhs V, rho0 = ||r||_D^2 --------------
template <class T>
__global__ __launch_bounds__(BLK) void k_rhs(const int64_t *__restrict__ off,
                                             const int *__restrict__ width,
                                             const int *__restrict__ col,
                                             const T *__restrict__ val,
                                             const T *__restrict__ V, T *__restrict__ r,
                                             const T *__restrict__ diag, int64_t n_slices,
                                             CgState *cg, Reduce red, double tol2, int maxit) {
  const int64_t slice = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double v[1] = {0.0};
  if (slice < n_slices) {
    const int64_t o = off[slice];
    const int w = width[slice];
    T acc = T(0);
    for (int j = 0; j < w; j++) {
      const int64_t k = o + (int64_t)j * 64 + lane;
      acc += val[k] * V[col[k]];
    }
    const int64_t row = slice * 64 + lane;
    r[row] = acc;
    v[0] = (double)diag[row] * (double)acc * (double)acc;
  }
  double tot[1];
  if (grid_reduce<1>(v, red, tot) && threadIdx.x == 0) {
    cg->rho0 = tot[0];
    cg->rho = tot[0];
    cg->tol2 = tol2;
    cg->maxit = maxit;
    cg->iter = 0;
    cg->first = 1;
    cg->upd_first = 1;
    cg->notconv = 0;
    cg->alpha = cg->beta = cg->gamma_old = cg->alpha_old = 0.0;
    cg->done = (tot[0] == 0.0) ? 1 : 0;
  }
}

// ---- CG (Chronopoulos-Gear form, Jacobi scaling folded into A_hat) ---------
// w = A_hat r, gamma = (r,r), delta = (r,w); the last block forms alpha, beta.
template <class T>
__global__ __launch_bounds__(BLK) void k_cg_spmv_dots(const int64_t *__restrict__ off,
                                                      const int *__restrict__ width,
                                                      const int *__restrict__ col,
                                                      const T *__restrict__ val,
                                                      const T *__restrict__ r, T *__restrict__ w,
                                                      int64_t n_slices, CgState *cg,
                                                      Reduce red) {
  if (load_done(cg)) return;
  const int64_t slice = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double v[2] = {0.0, 0.0};
  if (slice < n_slices) {
    const int64_t o = off[slice];
    const int wd = width[slice];
    T acc = T(0);
    for (int j = 0; j < wd; j++) {
      const int64_t k = o + (int64_t)j * 64 + lane;
      acc += val[k] * r[col[k]];
    }
    const int64_t row = slice * 64 + lane;
    w[row] = acc;
    const double rr = (double)r[row];
    v[0] = rr * rr;
    v[1] = rr * (double)acc;
  }
  double tot[2];
  if (grid_reduce<2>(v, red, tot) && threadIdx.x == 0) {
    const double gamma = tot[0], delta = tot[1];
    double alpha, beta;
    if (cg->first) {
      beta = 0.0;
      alpha = gamma / delta;
    } else {
      beta = gamma / cg->gamma_old;
      alpha = gamma / (delta - beta * gamma / cg->alpha_old);
    }
    if (!(alpha > 0.0) || !isfinite(alpha)) {  // breakdown: stop, report
      cg->done = 1;
      cg->notconv = 1;
    }
    cg->alpha = alpha;
    cg->beta = beta;
    cg->gamma_old = gamma;
    cg->alpha_old = alpha;
    cg->upd_first = cg->first;
    cg->first = 0;
  }
}

// p = r + beta p, s = w + beta s, x += alpha p, r -= alpha s; rho = ||r||_D^2.
template <class T>
__global__ __launch_bounds__(BLK) void k_cg_update(T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ p, T *__restrict__ s,
                                                   const T *__restrict__ w,
                                                   const T *__restrict__ diag, int64_t n_alloc,
                                                   CgState *cg, Reduce red) {
  if (load_done(cg)) return;
  const T alpha = (T)cg->alpha, beta = (T)cg->beta;
  const bool first = cg->upd_first != 0;
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  double v[1] = {0.0};
  if (i < n_alloc) {
    T pi, si, xi;
    const T ri = r[i], wi = w[i];
    if (first) {
      pi = ri;
      si = wi;
      xi = alpha * pi;
    } else {
      pi = ri + beta * p[i];
      si = wi + beta * s[i];
      xi = x[i] + alpha * pi;
    }
    const T rn = ri - alpha * si;
    p[i] = pi;
    s[i] = si;
    x[i] = xi;
    r[i] = rn;
    v[0] = (double)diag[i] * (double)rn * (double)rn;
  }
  double tot[1];
  if (grid_reduce<1>(v, red, tot) && threadIdx.x == 0) {
    cg->iter += 1;
    cg->rho = tot[0];
    if (tot[0] <= cg->tol2 * cg->rho0) {
      cg->done = 1;
    } else if (cg->iter >= cg->maxit || !isfinite(tot[0])) {
      cg->done = 1;
      cg->notconv = 1;
    }
  }
}

// ---- ionic update (fused with V = V_a + D^-1/2 x of the first half-step) ---
struct IonicArgs {
  int64_t n, n_alloc;
  int nsets;
  double Cm;
  int max_newton;
  int newton_jt;
};

template <class T, int MODEL, int INTEG>
__global__ __launch_bounds__(BLK) void k_ionic(const T *__restrict__ Vin, const T *__restrict__ x,
                                               const T *__restrict__ sinv, T *__restrict__ Vout,
                                               T *__restrict__ st, const T *__restrict__ istim,
                                               const uint16_t *__restrict__ pidx,
                                               const T *__restrict__ scales, SubSteps sub,
                                               IonicArgs a, unsigned long long *newton_sum,
                                               const CgState *cg) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int newton = 0;
  // x holds the half-step increment only if the solve did an iteration.
  const bool use_x = x && cg->iter > 0;
  if (i < a.n) {
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
    } else {
      T s[TT_NSTATE];
#pragma unroll
      for (int k = 1; k < TT_NSTATE; k++) s[k] = st[(k - 1) * a.n_alloc + i];
      const TTG<T> g = tt_conductances<T>(sc);
      const T it = -stim / (Cm * T(TT_XI));
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

// ---- second half-step apply: V_a = V_b + D^-1/2 x, activation, ECG ---------
template <class T, bool AT, bool ECG>
__global__ __launch_bounds__(BLK) void k_apply(const T *__restrict__ Vb, const T *__restrict__ x,
                                               const T *__restrict__ sinv, T *__restrict__ Va,
                                               int64_t n, T *__restrict__ at, T thr, double t0,
                                               double dt, const T *__restrict__ H, int ne,
                                               int64_t n_alloc, Reduce red,
                                               double *__restrict__ ecg_out, const CgState *cg) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  const bool use_x = cg->iter > 0;
  double v[MAX_ECG];
#pragma unroll
  for (int e = 0; e < MAX_ECG; e++) v[e] = 0.0;
  if (i < n) {
    const T Vn = use_x ? Vb[i] + sinv[i] * x[i] : Vb[i];
    if constexpr (AT) {
      const T Vo = Va[i];
      if (at[i] < T(0) && Vo < thr && Vn >= thr)
        at[i] = (T)(t0 + dt * (double)((thr - Vo) / (Vn - Vo)));
    }
    Va[i] = Vn;
    if constexpr (ECG) {
      for (int e = 0; e < ne; e++) v[e] = (double)H[(int64_t)e * n_alloc + i] * (double)Vn;
    }
  }
  if constexpr (ECG) {
    double tot[MAX_ECG];
    if (grid_reduce<MAX_ECG>(v, red, tot) && threadIdx.x == 0)
      for (int e = 0; e < ne; e++) ecg_out[e] = tot[e];
  }
}

template <class T>
__global__ __launch_bounds__(BLK) void k_ecg(const T *__restrict__ V, int64_t n,
                                             const T *__restrict__ H, int ne, int64_t n_alloc,
                                             Reduce red, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  double v[MAX_ECG];
#pragma unroll
  for (int e = 0; e < MAX_ECG; e++) v[e] = 0.0;
  if (i < n) {
    const double Vi = (double)V[i];
    for (int e = 0; e < ne; e++) v[e] = (double)H[(int64_t)e * n_alloc + i] * Vi;
  }
  double tot[MAX_ECG];
  if (grid_reduce<MAX_ECG>(v, red, tot) && threadIdx.x == 0)
    for (int e = 0; e < ne; e++) out[e] = tot[e];
}

// ---- independent cells (IonicModel adapter) --------------------------------
template <class T, int MODEL, int INTEG>
__global__ __launch_bounds__(BLK) void k_cells(const double *__restrict__ Vd,
                                               const double *__restrict__ stim_d,
                                               double *__restrict__ dvdt, T *__restrict__ st,
                                               const uint16_t *__restrict__ pidx,
                                               const T *__restrict__ scales, int64_t n,
                                               int64_t n_alloc, double Cm_d, double dt_d,
                                               int max_newton, int newton_jt) {
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
