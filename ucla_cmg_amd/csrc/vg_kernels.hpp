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
  int par;  // which ping-pong buffer holds the current r, w, s (fused CG)
  // Running totals kept by the persistent solver across solves.
  unsigned long long iters_total, solves_total, notconv_total;
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
    cg->par = 0;
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

// One whole CG iteration in one launch.  Given alpha_k, beta_k (from the
// previous launch's last block) and the previous r, w, s (read-only here,
// ping-pong buffers selected by cg->par), each row forms p, s, x, r of
// iteration k+1 and w_{k+1} = A_hat r_{k+1}; the neighbours' r_{k+1} =
// r - alpha (w + beta s) is rebuilt from their previous values, so no other
// block's writes are needed.  The last block tests ||r_{k+1}||_D against
// tol ||r_0||_D and forms alpha_{k+1}, beta_{k+1}.
template <class T, bool FIRST>
__device__ inline T cg_fused_row(const int64_t *__restrict__ off, const int *__restrict__ width,
                                 const int *__restrict__ col, const T *__restrict__ val,
                                 const T *__restrict__ r0, const T *__restrict__ w0,
                                 const T *__restrict__ s0, int64_t slice, int lane, T alpha,
                                 T beta) {
  const int64_t o = off[slice];
  const int wd = width[slice];
  T acc = T(0);
  for (int j = 0; j < wd; j++) {
    const int64_t k = o + (int64_t)j * 64 + lane;
    const int c = col[k];
    const T sc = FIRST ? w0[c] : w0[c] + beta * s0[c];
    acc += val[k] * (r0[c] - alpha * sc);
  }
  return acc;
}

template <class T>
__global__ __launch_bounds__(BLK) void k_cg_fused(const int64_t *__restrict__ off,
                                                  const int *__restrict__ width,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ val, T *rA, T *rB, T *wA,
                                                  T *wB, T *sA, T *sB, T *__restrict__ p,
                                                  T *__restrict__ x, const T *__restrict__ diag,
                                                  int64_t n_slices, CgState *cg, Reduce red) {
  if (load_done(cg)) return;
  const bool first = cg->upd_first != 0;
  const int par = cg->par;
  const T *r0 = par ? rB : rA, *w0 = par ? wB : wA, *s0 = par ? sB : sA;
  T *r1 = par ? rA : rB, *w1 = par ? wA : wB, *s1 = par ? sA : sB;
  const T alpha = (T)cg->alpha, beta = first ? T(0) : (T)cg->beta;
  const int64_t slice = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double v[3] = {0.0, 0.0, 0.0};
  if (slice < n_slices) {
    const int64_t row = slice * 64 + lane;
    const T ri = r0[row], wi = w0[row];
    T pi, si, xi, acc;
    if (first) {
      pi = ri;
      si = wi;
      xi = alpha * pi;
      acc = cg_fused_row<T, true>(off, width, col, val, r0, w0, s0, slice, lane, alpha, beta);
    } else {
      pi = ri + beta * p[row];
      si = wi + beta * s0[row];
      xi = x[row] + alpha * pi;
      acc = cg_fused_row<T, false>(off, width, col, val, r0, w0, s0, slice, lane, alpha, beta);
    }
    const T rn = ri - alpha * si;
    r1[row] = rn;
    w1[row] = acc;
    s1[row] = si;
    p[row] = pi;
    x[row] = xi;
    const double rd = (double)rn;
    v[0] = rd * rd;
    v[1] = rd * (double)acc;
    v[2] = (double)diag[row] * rd * rd;
  }
  double tot[3];
  if (grid_reduce<3>(v, red, tot) && threadIdx.x == 0) {
    cg->iter += 1;
    cg->rho = tot[2];
    cg->par = par ^ 1;
    if (tot[2] <= cg->tol2 * cg->rho0) {
      cg->done = 1;
    } else if (cg->iter >= cg->maxit || !isfinite(tot[2])) {
      cg->done = 1;
      cg->notconv = 1;
    } else {
      const double gamma = tot[0], delta = tot[1];
      const double beta_n = gamma / cg->gamma_old;
      const double alpha_n = gamma / (delta - beta_n * gamma / cg->alpha_old);
      if (!(alpha_n > 0.0) || !isfinite(alpha_n)) {
        cg->done = 1;
        cg->notconv = 1;
      }
      cg->alpha = alpha_n;
      cg->beta = beta_n;
      cg->gamma_old = gamma;
      cg->alpha_old = alpha_n;
      cg->upd_first = 0;
    }
  }
}

// ---- persistent half-step solve ---------------------------------------------
// One cooperative launch does a whole diffusion half-step: r0 = K_rhs V,
// w0 = A_hat r0, then CG iterations until ||r||_D <= tol ||r0||_D or maxit.
// Every block loops over slices grid-stride; phases are separated by a grid
// barrier whose last arriving block also finishes the FP64 reductions (in
// block order) and updates the CG scalars before releasing the others, so
// every block leaves the loop on the same iteration.
struct Barrier {
  unsigned *count;  // arrivals in the current phase (reset by the last block)
  unsigned *gen;    // phase generation, bumped by the last block
};

// Block-sums NV values into red.part[block]; in the last block to arrive,
// `finish(tot)` runs (thread 0) before the barrier opens.
template <int NV, class F>
__device__ inline void barrier_reduce(double (&v)[NV], Reduce red, Barrier bar, F finish) {
  __shared__ double sm[WAVES][NV];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sm[wv][k] = s;
  }
  __syncthreads();
  unsigned g = 0;
  if (threadIdx.x == 0) {
    g = __hip_atomic_load(bar.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = 0.0;
      for (int w = 0; w < WAVES; w++) s += sm[w][k];
      red.part[(size_t)blockIdx.x * NV + k] = s;
    }
    __threadfence();
    last = atomicAdd(bar.count, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last) {
    __threadfence();
    __shared__ double st[BLK];
    double tot[NV];
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
    if (threadIdx.x == 0) {
      finish(tot);
      __hip_atomic_store(bar.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      __hip_atomic_fetch_add(bar.gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (threadIdx.x == 0) {
    while (__hip_atomic_load(bar.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g)
      __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
  __threadfence();
}

template <class T>
__global__ __launch_bounds__(BLK) void k_cg_solve(const int64_t *__restrict__ off,
                                                  const int *__restrict__ width,
                                                  const int *__restrict__ col,
                                                  const T *__restrict__ Ahat,
                                                  const T *__restrict__ Krhs,
                                                  const T *__restrict__ V, T *rA, T *rB, T *wA,
                                                  T *wB, T *sA, T *sB, T *__restrict__ p,
                                                  T *__restrict__ x, const T *__restrict__ diag,
                                                  int64_t n_slices, CgState *cg, Reduce red,
                                                  Barrier bar, double tol2, int maxit) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * WAVES;
  const int64_t s0 = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6);
  // Phase 0: r0 = K_rhs V and ||r0||_D^2.
  {
    double v[1] = {0.0};
    for (int64_t sl = s0; sl < n_slices; sl += stride) {
      const int64_t o = off[sl];
      const int wd = width[sl];
      T acc = T(0);
      for (int j = 0; j < wd; j++) {
        const int64_t k = o + (int64_t)j * 64 + lane;
        acc += Krhs[k] * V[col[k]];
      }
      const int64_t row = sl * 64 + lane;
      rA[row] = acc;
      v[0] += (double)diag[row] * (double)acc * (double)acc;
    }
    barrier_reduce<1>(v, red, bar, [&](const double *t) {
      cg->solves_total += 1;
      cg->rho0 = t[0];
      cg->rho = t[0];
      cg->tol2 = tol2;
      cg->maxit = maxit;
      cg->iter = 0;
      cg->notconv = 0;
      cg->par = 0;
      cg->done = (t[0] == 0.0) ? 1 : 0;
    });
  }
  if (__hip_atomic_load(&cg->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  // Phase 1: w0 = A_hat r0, gamma0 = (r0,r0), delta0 = (r0,w0).
  {
    double v[2] = {0.0, 0.0};
    for (int64_t sl = s0; sl < n_slices; sl += stride) {
      const int64_t o = off[sl];
      const int wd = width[sl];
      T acc = T(0);
      for (int j = 0; j < wd; j++) {
        const int64_t k = o + (int64_t)j * 64 + lane;
        acc += Ahat[k] * rA[col[k]];
      }
      const int64_t row = sl * 64 + lane;
      wA[row] = acc;
      const double rr = (double)rA[row];
      v[0] += rr * rr;
      v[1] += rr * (double)acc;
    }
    barrier_reduce<2>(v, red, bar, [&](const double *t) {
      const double alpha = t[0] / t[1];
      cg->alpha = alpha;
      cg->beta = 0.0;
      cg->gamma_old = t[0];
      cg->alpha_old = alpha;
      cg->upd_first = 1;
      if (!(alpha > 0.0) || !isfinite(alpha)) {
        cg->done = 1;
        cg->notconv = 1;
        cg->notconv_total += 1;
      }
    });
  }
  // Iterations: the fused update + SpMV of k_cg_fused, one barrier each.
  for (;;) {
    if (__hip_atomic_load(&cg->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const bool first = cg->upd_first != 0;
    const int par = cg->par;
    const T *r0 = par ? rB : rA, *w0 = par ? wB : wA, *sp0 = par ? sB : sA;
    T *r1 = par ? rA : rB, *w1 = par ? wA : wB, *s1 = par ? sA : sB;
    const T alpha = (T)cg->alpha, beta = first ? T(0) : (T)cg->beta;
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t sl = s0; sl < n_slices; sl += stride) {
      const int64_t row = sl * 64 + lane;
      const T ri = r0[row], wi = w0[row];
      T pi, si, xi, acc;
      if (first) {
        pi = ri;
        si = wi;
        xi = alpha * pi;
        acc = cg_fused_row<T, true>(off, width, col, Ahat, r0, w0, sp0, sl, lane, alpha, beta);
      } else {
        pi = ri + beta * p[row];
        si = wi + beta * sp0[row];
        xi = x[row] + alpha * pi;
        acc = cg_fused_row<T, false>(off, width, col, Ahat, r0, w0, sp0, sl, lane, alpha, beta);
      }
      const T rn = ri - alpha * si;
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
    barrier_reduce<3>(v, red, bar, [&](const double *t) {
      cg->iter += 1;
      cg->iters_total += 1;
      cg->rho = t[2];
      cg->par = par ^ 1;
      cg->upd_first = 0;
      if (t[2] <= cg->tol2 * cg->rho0) {
        cg->done = 1;
      } else if (cg->iter >= cg->maxit || !isfinite(t[2])) {
        cg->done = 1;
        cg->notconv = 1;
        cg->notconv_total += 1;
      } else {
        const double beta_n = t[0] / cg->gamma_old;
        const double alpha_n = t[0] / (t[1] - beta_n * t[0] / cg->alpha_old);
        if (!(alpha_n > 0.0) || !isfinite(alpha_n)) {
          cg->done = 1;
          cg->notconv = 1;
          cg->notconv_total += 1;
        }
        cg->alpha = alpha_n;
        cg->beta = beta_n;
        cg->gamma_old = t[0];
        cg->alpha_old = alpha_n;
      }
    });
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

// Purkinje segment: rows [row0, row0 + n) of the solver order, states SoA with
// n_alloc lanes per state (all 22, the model keeps its own V copy).
template <class T>
__global__ __launch_bounds__(BLK) void k_ionic_pk(const T *__restrict__ Vin, const T *__restrict__ x,
                                                  const T *__restrict__ sinv, T *__restrict__ Vout,
                                                  T *__restrict__ st, const T *__restrict__ istim,
                                                  int64_t row0, int64_t n, int64_t n_alloc,
                                                  SubSteps sub, const CgState *cg) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const int64_t row = row0 + i;
  T V = Vin[row];
  if (x && cg->iter > 0) V += sinv[row] * x[row];
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
  } else if constexpr (MODEL == VG_MODEL_PURKINJE) {
    // All 22 states are kept: the model advances its own V copy internally.
    T s[PK_NSTATE];
#pragma unroll
    for (int k = 1; k < PK_NSTATE; k++) s[k] = st[k * n_alloc + i];
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
