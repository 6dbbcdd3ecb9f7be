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
  // Per-node sub-steps of CardiacBody::getVoltageFromIonicCurrent
  // (CardiacBody.cc:819-902) instead of the list: dt_ion, then strides of
  // dt_ion (5 dt_ion when the first dV/dt truncates to int 0) up to dT.
  int adaptive;
  double dT, dt_ion;
};

// Runs step(dt) (which advances a node by dt and returns its dV/dt) over the
// sub-steps of one ionic update: the fixed list, or (ADAPT, for sub.adaptive)
// the per-node strides.  The adaptive form's zero-length final call (strides
// landing exactly on dT) changes nothing and is skipped.  The two forms are
// separate instantiations: one kernel holding both inlines the model's step
// twice and needs half again the registers.
template <bool ADAPT, class F>
__device__ inline void ionic_substeps(const SubSteps &sub, F &&step) {
  if constexpr (!ADAPT) {
    for (int q = 0; q < sub.n; q++) step(sub.dt[q]);
    return;
  }
  double dt = sub.dt_ion;
  const double dT = sub.dT;
  if (!(dT / dt > 1.0)) {
    step(dt);
    return;
  }
  const double dv = step(dt);
  double tot = dt;
  if (fabs(dv) < 1.0) dt *= 5.0;  // int slope = iion; if (!slope) dt *= 5
  for (;;) {
    const bool last = tot + dt > dT;
    const double del = last ? dT - tot : dt;
    if (del > 0.0) step(del);
    tot += del;
    if (last) break;
  }
}

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
  int64_t n_slices;           // all slices (the block split needs no table with one instance)
  CgResult *pub;              // [n_inst] host-visible copies of the results, or null
};

// NV wave sums, the shuffles of the NV values interleaved so their latencies
// overlap; every lane ends with the totals.
template <int NV>
__device__ inline void wave_sum_n(double (&v)[NV]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) t[k] = __shfl_xor(v[k], o, 64);
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += t[k];
  }
}

__device__ inline double wave_sum(double v) {
  double u[1] = {v};
  wave_sum_n<1>(u);
  return u[0];
}

// Block sum of NV values, stored as part[k * gridDim.x + blockIdx.x].
template <int NV>
__device__ inline void put_partials(const double (&v)[NV], double *__restrict__ part) {
  __shared__ double sm[WAVES][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double u[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) u[k] = v[k];
  wave_sum_n<NV>(u);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sm[wv][k] = u[k];
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
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) s[k] = 0.0;
  // Block-outer, so the loads of several blocks' partials are in flight at once.
#pragma unroll 4
  for (int b = b0 + threadIdx.x; b < b1; b += BLK) {
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] += part[(size_t)k * nblk + b];
  }
  wave_sum_n<NV>(s);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) sm[wv][k] = s[k];
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

// Instance of this block, its block range [b0, b1) and its slices: wave w
// takes lo + w, lo + w + step, ... below hi.  A single instance over a
// multiple of XCDS blocks is cut into one contiguous range per XCD
// (workgroups go round-robin over the XCDs) whose slices the XCD's blocks
// take in turn, WAVES at a time: the waves an XCD runs at any moment work on
// one window of consecutive slices, so the rows they gather from (the
// operator's band around that window) were loaded into that XCD's L2 moments
// before.  Otherwise (instances, other grids) each block takes a contiguous
// chunk of its instance's slices.
__device__ inline int block_work(const CgBufs &cb, int64_t &lo, int64_t &hi, int64_t &step,
                                 int &b0, int &b1) {
  if (cb.n_inst == 1) {  // from kernel arguments alone: no load before the first stream
    const int b = blockIdx.x, g = gridDim.x;
    b0 = 0;
    b1 = g;
    if (g % XCDS == 0) {
      const int x = b % XCDS;
      lo = cb.n_slices * x / XCDS + (int64_t)(b / XCDS) * WAVES;
      hi = cb.n_slices * (x + 1) / XCDS;
      step = (int64_t)(g / XCDS) * WAVES;
    } else {
      lo = cb.n_slices * b / g;
      hi = cb.n_slices * (b + 1) / g;
      step = WAVES;
    }
    return 0;
  }
  const int b = blockIdx.x, k = cb.blk_inst[b];
  b0 = cb.inst_blk[k];
  b1 = cb.inst_blk[k + 1];
  const int64_t g = b1 - b0, s0 = cb.inst_slice[k], ns = cb.inst_slice[k + 1] - s0;
  const int64_t l = b - b0;
  lo = s0 + ns * l / g;
  hi = s0 + ns * (l + 1) / g;
  step = WAVES;
  return k;
}

// SELL-64 operator pattern (values are separate arrays on the same layout).
// Slice s: width[s] entries per lane from off[s]; the first 4*(width/4) in
// groups of four consecutive entries per lane ([group][lane][4]: one 16-byte
// value load and one 8-byte column load per lane fetch a group), the 0-3 left
// over lane-fastest ([entry][lane]).  Columns are 16-bit offsets from the row
// in col16 (same layout), or, for a slice with a column further than 32767
// rows away (wide[s] >= 0), int32 in col32 from wide[s].  A slice whose rows
// all draw from one sorted set of relative offsets (ustart[s] >= 0) stores
// no per-lane columns: entry j of every lane gathers row + uoff[ustart[s] + j]
// (a row lacking that neighbour holds a zero value there), the offsets being
// the same for the whole wave (scalar loads).
struct Sell {
  const int64_t *off;
  const int *width;
  const int64_t *wide;
  const int16_t *col16;
  const int *col32;
  const int64_t *ustart;
  const int *uoff;
};

template <class T>
struct Vec4Of;
template <>
struct Vec4Of<float> {
  using type = float4;
};
template <>
struct Vec4Of<double> {
  using type = double4;
};
template <class C>
struct Col4Of;
template <>
struct Col4Of<int16_t> {
  using type = short4;
};
template <>
struct Col4Of<int> {
  using type = int4;
};

// sum over j < wd of v(j) * gather(col(j)) in order j = 0, 1, ... for one
// lane of a slice whose values start at vs and columns at cs; colmap turns a
// stored column into a row index.  Two groups (8 entries) are loaded before
// their gathers, so each wave keeps two wide loads and 8 gathers in flight.
// A wide load of the operator's values.  NT: a nontemporal load, for
// operators larger than the Infinity Cache, whose values are read once per
// launch: the hint keeps them from displacing the row records the gathers
// reuse (on the 4.6M-row slab 4.28 -> 4.04 ms/step; on a cache-resident
// operator the same hint costs 19 %, so it is chosen per operator).
template <bool NT, class V4>
__device__ inline V4 ld_vals(const V4 *p) {
  if constexpr (NT) {
    typedef decltype(p->x) E;
    typedef E ev4 __attribute__((ext_vector_type(4)));
    const ev4 t = __builtin_nontemporal_load(reinterpret_cast<const ev4 *>(p));
    return V4{t.x, t.y, t.z, t.w};
  } else {
    return *p;
  }
}
template <bool NT, class E>
__device__ inline E ld_val(const E *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <class T, bool NT = false, class C, class CM, class G>
__device__ inline T sell_dot(int wd, const T *__restrict__ vs, const C *__restrict__ cs, int lane,
                             CM colmap, G gather) {
  using V4 = typename Vec4Of<T>::type;
  using C4 = typename Col4Of<C>::type;
  const int n4 = wd >> 2;
  const V4 *__restrict__ vp = reinterpret_cast<const V4 *>(vs) + lane;
  const C4 *__restrict__ cp = reinterpret_cast<const C4 *>(cs) + lane;
  T acc = T(0);
  int g = 0;
  for (; g + 2 <= n4; g += 2) {
    V4 v[2];
    C4 c[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      v[q] = ld_vals<NT>(vp + (g + q) * 64);
      c[q] = cp[(g + q) * 64];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      acc += v[q].x * gather(colmap(c[q].x));
      acc += v[q].y * gather(colmap(c[q].y));
      acc += v[q].z * gather(colmap(c[q].z));
      acc += v[q].w * gather(colmap(c[q].w));
    }
  }
  if (g < n4) {
    const V4 v = ld_vals<NT>(vp + g * 64);
    const C4 c = cp[g * 64];
    acc += v.x * gather(colmap(c.x));
    acc += v.y * gather(colmap(c.y));
    acc += v.z * gather(colmap(c.z));
    acc += v.w * gather(colmap(c.w));
  }
  const int rem = wd & 3;
  const T *__restrict__ vt = vs + (int64_t)n4 * 256 + lane;
  const C *__restrict__ ct = cs + (int64_t)n4 * 256 + lane;
  int cc[3];
  T vv[3];
#pragma unroll
  for (int q = 0; q < 3; q++)
    if (q < rem) {
      cc[q] = colmap(ct[q * 64]);
      vv[q] = ld_val<NT>(vt + q * 64);
    }
#pragma unroll
  for (int q = 0; q < 3; q++)
    if (q < rem) acc += vv[q] * gather(cc[q]);
  return acc;
}

// The same value in every lane of the wave, held in scalar registers: a
// slice index is (all lanes of a wave work on one slice), and loads from a
// scalar address go through the scalar cache instead of vector memory.
__device__ inline int64_t wave_uniform(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// sell_dot for a shared-offset slice: entry j gathers row + uo[j], uo the
// slice's (wave-uniform) offsets, padded to whole groups of four.  The gather
// addresses need no per-lane column, so the groups are taken three at a time:
// a 14-entry row -- the common width on P1 tets -- has all three of its
// groups' values and gathers in flight at once (LV, CG iteration: 24.8 us
// with two at a time, 22.8 us with three; a separate two-group tail, or the
// chunk as a function template, measured 23.8 us).
template <class T, bool NT = false, class G>
__device__ inline T sell_dot_u(int wd, const T *__restrict__ vs, const int *__restrict__ uo,
                               int lane, int row, G gather) {
  using V4 = typename Vec4Of<T>::type;
  constexpr int NG = 3;
  const int n4 = wd >> 2;
  const V4 *__restrict__ vp = reinterpret_cast<const V4 *>(vs) + lane;
  const int4 *__restrict__ up = reinterpret_cast<const int4 *>(uo);
  T acc = T(0);
  int g = 0;
  for (; g + NG <= n4; g += NG) {
    V4 v[NG];
    int4 c[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) {
      v[q] = ld_vals<NT>(vp + (g + q) * 64);
      c[q] = up[g + q];
    }
#pragma unroll
    for (int q = 0; q < NG; q++) {
      acc += v[q].x * gather(row + c[q].x);
      acc += v[q].y * gather(row + c[q].y);
      acc += v[q].z * gather(row + c[q].z);
      acc += v[q].w * gather(row + c[q].w);
    }
  }
  for (; g < n4; g++) {
    const V4 v = ld_vals<NT>(vp + g * 64);
    const int4 c = up[g];
    acc += v.x * gather(row + c.x);
    acc += v.y * gather(row + c.y);
    acc += v.z * gather(row + c.z);
    acc += v.w * gather(row + c.w);
  }
  const int rem = wd & 3;
  if (rem) {
    const T *__restrict__ vt = vs + (int64_t)n4 * 256 + lane;
    const int4 c = up[n4];
    const int cc[3] = {c.x, c.y, c.z};
    T vv[3];
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q < rem) vv[q] = ld_val<NT>(vt + q * 64);
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q < rem) acc += vv[q] * gather(row + cc[q]);
  }
  return acc;
}

// One lane's row of a slice: sum_j val[j] * gather(col[j]).
template <class T, bool NT = false, class G>
__device__ inline T sell_row(const Sell &S, const T *__restrict__ val, int64_t slice, int lane,
                             G gather) {
  slice = wave_uniform(slice);
  const int64_t o = S.off[slice];
  const int wd = S.width[slice];
  const int64_t us = S.ustart[slice];
  if (us >= 0) return sell_dot_u<T, NT>(wd, val + o, S.uoff + us, lane, (int)(slice * 64 + lane), gather);
  const int64_t wo = S.wide[slice];
  if (wo < 0) {
    const int row = (int)(slice * 64 + lane);
    return sell_dot<T, NT>(wd, val + o, S.col16 + o, lane, [&](int c) { return row + c; }, gather);
  }
  return sell_dot<T, NT>(wd, val + o, S.col32 + wo, lane, [](int c) { return c; }, gather);
}

// Two operators on one pattern in one pass (columns loaded once):
// acc_a = sum va[j] ga(col[j]), acc_b = sum vb[j] gb(col[j]).
template <class T, bool NT = false, class C, class CM, class GA, class GB>
__device__ inline void sell_dot2(int wd, const T *__restrict__ as, const T *__restrict__ bs,
                                 const C *__restrict__ cs, int lane, CM colmap, GA ga, GB gb,
                                 T &acc_a, T &acc_b) {
  using V4 = typename Vec4Of<T>::type;
  using C4 = typename Col4Of<C>::type;
  const int n4 = wd >> 2;
  const V4 *__restrict__ ap = reinterpret_cast<const V4 *>(as) + lane;
  const V4 *__restrict__ bp = reinterpret_cast<const V4 *>(bs) + lane;
  const C4 *__restrict__ cp = reinterpret_cast<const C4 *>(cs) + lane;
  T sa = T(0), sb = T(0);
  for (int g = 0; g < n4; g++) {
    const V4 x = ld_vals<NT>(ap + g * 64), y = ld_vals<NT>(bp + g * 64);
    const C4 c4 = cp[g * 64];
    const int c[4] = {colmap(c4.x), colmap(c4.y), colmap(c4.z), colmap(c4.w)};
    const T xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      sa += xs[q] * ga(c[q]);
      sb += ys[q] * gb(c[q]);
    }
  }
  const int rem = wd & 3;
  const int64_t t = (int64_t)n4 * 256 + lane;
  int c[3];
  T x[3], y[3];
#pragma unroll
  for (int q = 0; q < 3; q++)
    if (q < rem) {
      c[q] = colmap(cs[t + q * 64]);
      x[q] = ld_val<NT>(as + t + q * 64);
      y[q] = ld_val<NT>(bs + t + q * 64);
    }
#pragma unroll
  for (int q = 0; q < 3; q++)
    if (q < rem) {
      sa += x[q] * ga(c[q]);
      sb += y[q] * gb(c[q]);
    }
  acc_a = sa;
  acc_b = sb;
}

// sell_dot2 for a shared-offset slice (see sell_dot_u).
template <class T, bool NT = false, class GA, class GB>
__device__ inline void sell_dot2_u(int wd, const T *__restrict__ as, const T *__restrict__ bs,
                                   const int *__restrict__ uo, int lane, int row, GA ga, GB gb,
                                   T &acc_a, T &acc_b) {
  using V4 = typename Vec4Of<T>::type;
  const int n4 = wd >> 2;
  const V4 *__restrict__ ap = reinterpret_cast<const V4 *>(as) + lane;
  const V4 *__restrict__ bp = reinterpret_cast<const V4 *>(bs) + lane;
  const int4 *__restrict__ up = reinterpret_cast<const int4 *>(uo);
  T sa = T(0), sb = T(0);
  for (int g = 0; g < n4; g++) {
    const V4 x = ld_vals<NT>(ap + g * 64), y = ld_vals<NT>(bp + g * 64);
    const int4 c4 = up[g];
    const int c[4] = {row + c4.x, row + c4.y, row + c4.z, row + c4.w};
    const T xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      sa += xs[q] * ga(c[q]);
      sb += ys[q] * gb(c[q]);
    }
  }
  const int rem = wd & 3;
  if (rem) {
    const int64_t t = (int64_t)n4 * 256 + lane;
    const int4 c4 = up[n4];
    const int c[3] = {row + c4.x, row + c4.y, row + c4.z};
    T x[3], y[3];
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q < rem) {
        x[q] = ld_val<NT>(as + t + q * 64);
        y[q] = ld_val<NT>(bs + t + q * 64);
      }
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q < rem) {
        sa += x[q] * ga(c[q]);
        sb += y[q] * gb(c[q]);
      }
  }
  acc_a = sa;
  acc_b = sb;
}

template <class T, bool NT = false, class GA, class GB>
__device__ inline void sell_row2(const Sell &S, const T *__restrict__ va, const T *__restrict__ vb,
                                 int64_t slice, int lane, GA ga, GB gb, T &acc_a, T &acc_b) {
  slice = wave_uniform(slice);
  const int64_t o = S.off[slice];
  const int wd = S.width[slice];
  const int64_t us = S.ustart[slice];
  if (us >= 0) {
    sell_dot2_u<T, NT>(wd, va + o, vb + o, S.uoff + us, lane, (int)(slice * 64 + lane), ga, gb, acc_a,
                   acc_b);
    return;
  }
  const int64_t wo = S.wide[slice];
  if (wo < 0) {
    const int row = (int)(slice * 64 + lane);
    sell_dot2<T, NT>(wd, va + o, vb + o, S.col16 + o, lane, [&](int c) { return row + c; }, ga, gb,
                 acc_a, acc_b);
  } else {
    sell_dot2<T, NT>(wd, va + o, vb + o, S.col32 + wo, lane, [](int c) { return c; }, ga, gb, acc_a,
                 acc_b);
  }
}

__device__ inline bool cg_stopped(const CgResult *res) {
  return __hip_atomic_load(&res->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ inline void cg_set_stop(CgResult *res) {
  __hip_atomic_store(&res->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Copies a finished solve's result into host-coherent memory, `done` last
// with a system-scope release, so the host sees the outcome while the
// stream's remaining launches still run.
__device__ inline void cg_publish(CgResult *pub, int iter, int notconv, int x_valid, double rho,
                                  double rho_ref) {
  if (!pub) return;
  pub->iter = iter;
  pub->notconv = notconv;
  pub->stop = 1;
  pub->x_valid = x_valid;
  pub->rho = rho;
  pub->rho_ref = rho_ref;
  __hip_atomic_store(&pub->done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- explicit diffusion stride ------------------------------------------------
// Vout = Vin + h (-C_L^-1 K) Vin, one row per lane (EP_ExplicitSolver.cc:
// 52-85; Kexp holds the off-diagonal entries of -K / C_L in the SELL layout,
// whose rows sum to zero, so the product is sum_j Kexp_ij (Vin_j - Vin_i)).
template <class T>
__global__ __launch_bounds__(BLK) void k_explicit(Sell S, const T *__restrict__ Kexp,
                                                  const T *__restrict__ Vin, T *__restrict__ Vout,
                                                  T h, int64_t n_slices) {
  const int lane = threadIdx.x & 63;
  for (int64_t sl = (int64_t)blockIdx.x * WAVES + (threadIdx.x >> 6); sl < n_slices;
       sl += (int64_t)gridDim.x * WAVES) {
    const int64_t row = sl * 64 + lane;
    const T vi = Vin[row];
    const T acc = sell_row<T>(S, Kexp, sl, lane, [&](int c) { return Vin[c] - vi; });
    Vout[row] = vi + h * acc;
  }
}

// ---- Purkinje rows of the initial guess ----------------------------------------
// After the right-hand side of a solve that starts from a guess, one
// workgroup improves the guess on the Purkinje rows [r0, r0 + npk) with the
// myocardium rows held: PK_ITERS CG iterations on A_pp e = r0_p, A_pp the
// Purkinje rows' block of A_hat, then x_p += e and r0 -= A_hat e (on the
// Purkinje rows that is the local residual; on the myocardium rows coupled to
// the tree, the PMJ targets, it is their coupling entries times e).  The
// tree's fronts are sharp in time, so any guess carried from earlier solves
// is poor there, and while the tree conducts those ~0.4 % of the rows set the
// initial residual; A_pp is well conditioned (Jacobi-scaled mass-dominated
// cable, eigenvalues ~[0.5, 1.5] on the LV tree), so a few local iterations
// remove that part.  The solve that follows stops on the same test as
// before: only its starting point moves.
// The block (blk_v/blk_j: at most PK_MAXE entries per row, local column
// indices, entry-major) and the coupling rows (cpl_*: a myocardium row, its
// CPL_MAXE entries in Purkinje columns) are built on the host; unused slots
// hold value 0 and index 0.  The search direction lives in LDS; each thread
// owns rows tid + k PK_BLK.
constexpr int PK_BLK = 1024;
// Rows per thread (registers hold each row's block entries): at most
// PK_BLK * pk_rows<T>() Purkinje rows.
template <class T>
constexpr int pk_rows() {
  return sizeof(T) == 8 ? 4 : 5;
}
constexpr int PK_MAXE = 4;  // block entries per row (a tree node: itself + 3 neighbours)
constexpr int CPL_MAXE = 4;
constexpr int PK_ITERS = 6;  // LV beat: 4 0.915, 6 0.918, 8 0.930, 12 0.950 ms/step

// Block sums of two values; sm holds two slots per wave and alternates
// between calls (buf), so one barrier suffices: a slot is written again only
// two calls later, after another barrier.
__device__ inline void pk_block_sum2(double &a, double &b, double (*sm)[PK_BLK / 64][2], int buf) {
  double v[2] = {a, b};
  wave_sum_n<2>(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sm[buf][wv][0] = v[0];
    sm[buf][wv][1] = v[1];
  }
  __syncthreads();
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int w = 0; w < PK_BLK / 64; w++) {
    s0 += sm[buf][w][0];
    s1 += sm[buf][w][1];
  }
  a = s0;
  b = s1;
}

// Local solve in the Chronopoulos-Gear form (one fused reduction and one
// product per iteration): the residual is the vector other rows read, in LDS.
template <class T>
__global__ __launch_bounds__(PK_BLK) void k_pk_presolve(T *__restrict__ r, T *__restrict__ x,
                                                        const T *__restrict__ blk_v,
                                                        const int *__restrict__ blk_j, int64_t r0,
                                                        int npk, const int64_t *__restrict__ cpl_row,
                                                        const T *__restrict__ cpl_v,
                                                        const int *__restrict__ cpl_j, int ncpl) {
  extern __shared__ double pk_lds[];
  T *rl = reinterpret_cast<T *>(pk_lds);
  __shared__ double sm[2][PK_BLK / 64][2];
  const int tid = threadIdx.x;
  constexpr int PK_ROWS = pk_rows<T>();
  T e[PK_ROWS], rr[PK_ROWS], p[PK_ROWS], sv[PK_ROWS], ev[PK_ROWS][PK_MAXE];
  int ec[PK_ROWS][PK_MAXE];
#pragma unroll
  for (int k = 0; k < PK_ROWS; k++) {
    const int i = tid + k * PK_BLK;
    e[k] = rr[k] = p[k] = sv[k] = T(0);
    const bool on = i < npk;
#pragma unroll
    for (int m = 0; m < PK_MAXE; m++) {
      ev[k][m] = on ? blk_v[(int64_t)m * npk + i] : T(0);
      ec[k][m] = on ? blk_j[(int64_t)m * npk + i] : 0;
    }
    if (on) {
      rr[k] = r[r0 + i];
      rl[i] = rr[k];
    }
  }
  __syncthreads();
  double gamma_old = 0.0, alpha_old = 0.0;
  for (int it = 0; it < PK_ITERS; it++) {
    T q[PK_ROWS];
    double gamma = 0.0, delta = 0.0;
#pragma unroll
    for (int k = 0; k < PK_ROWS; k++) {
      T a = T(0);
#pragma unroll
      for (int m = 0; m < PK_MAXE; m++) a += ev[k][m] * rl[ec[k][m]];
      q[k] = a;
      gamma += (double)rr[k] * (double)rr[k];
      delta += (double)rr[k] * (double)a;
    }
    pk_block_sum2(gamma, delta, sm, it & 1);  // also: every read of rl is done
    if (!(gamma > 0.0) || !isfinite(gamma) || !isfinite(delta)) break;
    double beta = 0.0, alpha;
    if (it == 0) {
      alpha = gamma / delta;
    } else {
      beta = gamma / gamma_old;
      alpha = gamma / (delta - beta * gamma / alpha_old);
    }
    if (!(alpha > 0.0) || !isfinite(alpha)) break;
    gamma_old = gamma;
    alpha_old = alpha;
    const T a = (T)alpha, b = (T)beta;
#pragma unroll
    for (int k = 0; k < PK_ROWS; k++) {
      const int i = tid + k * PK_BLK;
      p[k] = rr[k] + b * p[k];
      sv[k] = q[k] + b * sv[k];
      e[k] += a * p[k];
      rr[k] -= a * sv[k];
      if (i < npk) rl[i] = rr[k];
    }
    __syncthreads();  // the new residual before the next product
  }
  // Every read of rl is behind a barrier (the one ending an iteration, or
  // the reduction's before a break), so rl can take e for the coupling rows.
#pragma unroll
  for (int k = 0; k < PK_ROWS; k++) {
    const int i = tid + k * PK_BLK;
    if (i < npk) {
      x[r0 + i] += e[k];
      r[r0 + i] = rr[k];
      rl[i] = e[k];
    }
  }
  __syncthreads();
  for (int c = tid; c < ncpl; c += PK_BLK) {
    T a = T(0);
#pragma unroll
    for (int m = 0; m < CPL_MAXE; m++) a += cpl_v[(int64_t)m * ncpl + c] * rl[cpl_j[(int64_t)m * ncpl + c]];
    r[cpl_row[c]] -= a;
  }
}

// The same local solve as a Chebyshev iteration on the block's eigenvalue
// interval [lmin, lmax] (Gershgorin bounds of A_pp, computed on the host;
// Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1): no dot
// products, so no block reductions -- each iteration is one product with the
// direction d (neighbours' d from LDS, double-buffered) and one barrier.
// With [0.51, 1.49] (the LV tree) PK_ITERS = 6 iterations bound the error
// reduction at 2 / T_6(theta / delta) ~ 1e-3.
// The recurrence's coefficients depend on the interval only, so the host forms
// them once per operator (pk_cheb_coef, in FP64 as the kernel would: IEEE
// products and quotients round alike on both sides) and the kernel's
// iterations carry no FP64 division, whose dependent chain per iteration had
// cost more than the iteration's product and barrier.
struct PkCheb {
  double theta;                     // centre of the interval
  double c0[PK_ITERS], c1[PK_ITERS];  // d <- c0[it] d + c1[it] r after iteration it
};
inline PkCheb pk_cheb_coef(double lmin, double lmax) {
  PkCheb c{};
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  c.theta = theta;
  double rho = 1.0 / sigma;
  for (int it = 0; it < PK_ITERS; it++) {
    const double rho_n = 1.0 / (2.0 * sigma - rho);
    c.c0[it] = rho_n * rho;
    c.c1[it] = 2.0 * rho_n / delta;
    rho = rho_n;
  }
  return c;
}

template <class T>
__global__ __launch_bounds__(PK_BLK) void k_pk_cheb(T *__restrict__ r, T *__restrict__ x,
                                                    const T *__restrict__ blk_v,
                                                    const int *__restrict__ blk_j, int64_t r0,
                                                    int npk, const int64_t *__restrict__ cpl_row,
                                                    const T *__restrict__ cpl_v,
                                                    const int *__restrict__ cpl_j, int ncpl,
                                                    PkCheb cf) {
  extern __shared__ double pk_lds[];
  T *dl0 = reinterpret_cast<T *>(pk_lds), *dl1 = dl0 + npk;
  const int tid = threadIdx.x;
  constexpr int PK_ROWS = pk_rows<T>();
  const double theta = cf.theta;
  T e[PK_ROWS], rr[PK_ROWS], d[PK_ROWS], ev[PK_ROWS][PK_MAXE];
  int ec[PK_ROWS][PK_MAXE];
#pragma unroll
  for (int k = 0; k < PK_ROWS; k++) {
    const int i = tid + k * PK_BLK;
    e[k] = rr[k] = d[k] = T(0);
    const bool on = i < npk;
#pragma unroll
    for (int m = 0; m < PK_MAXE; m++) {
      ev[k][m] = on ? blk_v[(int64_t)m * npk + i] : T(0);
      ec[k][m] = on ? blk_j[(int64_t)m * npk + i] : 0;
    }
    if (on) {
      rr[k] = r[r0 + i];
      d[k] = (T)((double)rr[k] / theta);
      dl0[i] = d[k];
    }
  }
  __syncthreads();
  // Unrolled in FP32; rolled in FP64, whose unrolled form spills registers.
  constexpr int PK_UNROLL = sizeof(T) == 8 ? 1 : PK_ITERS;
#pragma unroll PK_UNROLL
  for (int it = 0; it < PK_ITERS; it++) {
    const T *dc = (it & 1) ? dl1 : dl0;
    T *dn = (it & 1) ? dl0 : dl1;
#pragma unroll
    for (int k = 0; k < PK_ROWS; k++) {
      T a = T(0);
#pragma unroll
      for (int m = 0; m < PK_MAXE; m++) a += ev[k][m] * dc[ec[k][m]];
      e[k] += d[k];
      rr[k] -= a;
    }
    if (it + 1 == PK_ITERS) break;
    const T c0 = (T)cf.c0[it], c1 = (T)cf.c1[it];
#pragma unroll
    for (int k = 0; k < PK_ROWS; k++) {
      const int i = tid + k * PK_BLK;
      d[k] = c0 * d[k] + c1 * rr[k];
      if (i < npk) dn[i] = d[k];
    }
    // The new direction before the next product; the buffer written next
    // (dc) is read no more after this barrier.
    __syncthreads();
  }
  __syncthreads();  // every read of the direction buffers is done
#pragma unroll
  for (int k = 0; k < PK_ROWS; k++) {
    const int i = tid + k * PK_BLK;
    if (i < npk) {
      x[r0 + i] += e[k];
      r[r0 + i] = rr[k];
      dl0[i] = e[k];
    }
  }
  __syncthreads();
  for (int c = tid; c < ncpl; c += PK_BLK) {
    T a = T(0);
#pragma unroll
    for (int m = 0; m < CPL_MAXE; m++) a += cpl_v[(int64_t)m * ncpl + c] * dl0[cpl_j[(int64_t)m * ncpl + c]];
    r[cpl_row[c]] -= a;
  }
}

// ---- diffusion right-hand side ---------------------------------------------
// b = K_rhs V and r_0 = b - A_hat x_0 (x_0 = 0 unless GUESS: then x holds the
// guess of this solve); partials of ||b||_D^2 and ||r_0||_D^2.  Opens a solve:
// clears the result record (no CG launch of the previous solve runs
// concurrently: launches on one stream are ordered).
// Every row of the stiffness sums to zero (K 1 = 0), so b_i is formed as
// sum_j K_ij (V_j - V_i): the same value, without summing terms of ~|K_ij| 85
// mV that cancel to a small b_i.  In FP32 that cancellation leaves rounding
// noise of ~1e-7 |K_ii| 85 mV in every row, and in a tissue mostly at rest the
// noise of 1M rows outweighs the wavefront's b, so a solve to 1e-7 ||b|| had
// to resolve the noise.
template <class T, bool GUESS, bool NT = false>
__global__ __launch_bounds__(BLK) void k_rhs(Sell S, const T *__restrict__ Krhs,
                                             const T *__restrict__ Ahat,
                                             const T *__restrict__ V, const T *__restrict__ x,
                                             T *__restrict__ r, const T *__restrict__ diag,
                                             CgBufs cb) {
  int64_t lo, hi, step;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, step, b0, b1);
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
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += step) {
    const int64_t row = sl * 64 + lane;
    const double d = (double)diag[row];
    const T vi = V[row];
    if constexpr (GUESS) {
      T bk, ax;
      const T xi = x[row];
      sell_row2<T, NT>(S, Krhs, Ahat, sl, lane, [&](int c) { return V[c] - vi; },
                   [&](int c) { return x[c]; }, bk, ax);
      const T rr = bk - (xi + ax);  // A_hat's unit diagonal
      r[row] = rr;
      v[0] += d * (double)bk * (double)bk;
      v[1] += d * (double)rr * (double)rr;
    } else {
      const T bk = sell_row<T, NT>(S, Krhs, sl, lane, [&](int c) { return V[c] - vi; });
      r[row] = bk;
      const double e = d * (double)bk * (double)bk;
      v[0] += e;
      v[1] += e;
    }
  }
  put_partials<2>(v, cb.part1);
}

// Increment history of one half-step for the extrapolated initial guess: x
// holds that half-step's last increment d_n (its CG unknown), y = d_{n-1},
// z = d_{n-2}.  order = how many of them are valid (0: no history kept).
template <class T>
struct XHist {
  T *x, *y, *z;
  int order;
};

// Turns a half-step's history into the guess of its next solve, row by row:
// x <- d_n (order 1), 2 d_n - d_{n-1} (2) or d_{n-2} + 3 (d_n - d_{n-1}) (3),
// and shifts the history.  Runs inside the other half-step's solve, after the
// consumers of d_n (the ionic update or the apply) and before that
// half-step's next right-hand side, which gathers the guess.
template <class T>
__device__ inline void rotate_history(const XHist<T> &h, int64_t row) {
  const T xn = h.x[row], y = h.y[row], z = h.z[row];
  const T g = h.order == 1 ? xn : h.order == 2 ? xn + (xn - y) : z + T(3) * (xn - y);
  h.z[row] = y;
  h.y[row] = xn;
  h.x[row] = g;
}

// ---- CG (Chronopoulos-Gear form, Jacobi scaling folded into A_hat) ---------
// Start: ||b||_D^2 from the right-hand side's partials (b = 0: the increment
// is zero); w_0 = A_hat r_0, the records {r_0, w_0, 0, 0} of parity 0, and
// partials of (r_0, r_0), (r_0, w_0), ||r_0||_D^2.  With hist.order > 0 it
// also rotates the other half-step's increment history (row-local).
template <class T, bool NT = false>
__global__ __launch_bounds__(BLK) void k_cg_start(Sell S, const T *__restrict__ val,
                                                  const T *__restrict__ r,
                                                  typename Vec4Of<T>::type *__restrict__ q,
                                                  const T *__restrict__ diag, CgBufs cb,
                                                  XHist<T> hist) {
  int64_t lo, hi, step;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, step, b0, b1);
  CgResult *res = cb.res + inst;
  if (hist.order > 0)
    for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += step)
      rotate_history(hist, sl * 64 + (threadIdx.x & 63));
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
      cg_publish(cb.pub ? cb.pub + inst : nullptr, 0, res->notconv, 0, rho_ref, rho_ref);
    }
    return;
  }
  const int lane = threadIdx.x & 63;
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += step) {
    const int64_t row = sl * 64 + lane;
    const T ri = r[row];
    const T acc = ri + sell_row<T, NT>(S, val, sl, lane, [&](int c) { return r[c]; });
    q[row] = {ri, acc, T(0), T(0)};
    const double rr = (double)ri;
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
// r_j, w_j, s_{j-1}.  Rows are {r, w, s, p} in one 16-byte (FP64: 32-byte)
// record, so a neighbour costs one wide gather; launch j reads the records
// {r_j, w_j, s_{j-1}, p_{j-1}} of parity j and writes {r_{j+1}, w_{j+1}, s_j,
// p_j} to the other array, so no other block's writes of this launch are read.
template <class T, bool FIRST, bool NT>
__device__ inline T cg_iter_row(const Sell &S, const T *__restrict__ val,
                                const typename Vec4Of<T>::type *__restrict__ q0, int64_t slice,
                                int lane, T alpha, T beta) {
  return sell_row<T, NT>(S, val, slice, lane, [&](int c) {
    const typename Vec4Of<T>::type g = q0[c];
    const T sc = FIRST ? g.y : g.y + beta * g.z;
    return g.x - alpha * sc;
  });
}

template <class T, bool NT = false>
__global__ __launch_bounds__(BLK) void k_cg_iter(Sell S, const T *__restrict__ val,
                                                 typename Vec4Of<T>::type *qA,
                                                 typename Vec4Of<T>::type *qB,
                                                 T *__restrict__ x, const T *__restrict__ diag,
                                                 CgBufs cb, int j, double tol2, int maxit,
                                                 int guess) {
  using V4 = typename Vec4Of<T>::type;
  int64_t lo, hi, step;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, step, b0, b1);
  CgResult *res = cb.res + inst;
  // The stop flag, the carried scalars and the partials are loaded together;
  // a launch queued past the end of its solve reads partials nobody wrote
  // for it and discards them.
  const bool stopped = cg_stopped(res);
  const int par = j & 1;
  const CgCarry c = cb.carry[par * cb.n_inst + inst];
  double t[3];
  sum_partials<3>(par ? cb.part1 : cb.part0, gridDim.x, b0, b1, t);
  if (stopped) return;
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
      cg_publish(cb.pub ? cb.pub + inst : nullptr, j, end == 2 ? 1 : 0, res->x_valid, rho,
                 c.rho_ref);
    }
    return;
  }
  const V4 *__restrict__ q0 = par ? qB : qA;
  V4 *__restrict__ q1 = par ? qA : qB;
  const T a = (T)alpha, bt = (T)beta;
  const int lane = threadIdx.x & 63;
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += step) {
    const int64_t row = sl * 64 + lane;
    const V4 own = q0[row];
    const T ri = own.x, wi = own.y;
    T pi, si, xi, acc;
    if (j == 0) {
      pi = ri;
      si = wi;
      xi = guess ? x[row] + a * pi : a * pi;
      acc = cg_iter_row<T, true, NT>(S, val, q0, sl, lane, a, bt);
    } else {
      pi = ri + bt * own.w;
      si = wi + bt * own.z;
      xi = x[row] + a * pi;
      acc = cg_iter_row<T, false, NT>(S, val, q0, sl, lane, a, bt);
    }
    const T rn = ri - a * si;
    acc += rn;  // A_hat's unit diagonal
    q1[row] = {rn, acc, si, pi};
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
    // x holds an iterate from here on: a solve cut short by the end of a
    // captured launch sequence still passes its last iterate on.
    res->x_valid = 1;
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
  // Luo-Rudy: LuoRudyGateTable rows in T (LUORUDYWITHGATE), or null.
  const void *lr_table;
};

template <class T, int MODEL, int INTEG, bool ADAPT>
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
      ionic_substeps<ADAPT>(sub, [&](double dtd) {
        const T dt = (T)dtd;
        const T dv = lr_compute_ion<T, sizeof(T) == 4>(s, Cm, dt, V, stim, sc[0], sc[3],
                                                       (const T *)a.lr_table);
        V += dt * dv;
        return (double)dv;
      });
#pragma unroll
      for (int k = 0; k < LR_NSTATE - 1; k++) st[k * a.n_alloc + i] = s[k];
    } else if constexpr (MODEL == VG_MODEL_MJ || MODEL == VG_MODEL_MJ_FAIL) {
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
      ionic_substeps<ADAPT>(sub, [&](double dtd) {
        const T dt = (T)dtd;
        const T dv = mj_step<T, sizeof(T) == 4, MODEL == VG_MODEL_MJ_FAIL>(s, ac, V, it, dt, sc[0], sc[1], a.mj_int_abs != 0);
        V += dt * dv;
        return (double)dv;
      });
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
      ionic_substeps<ADAPT>(sub, [&](double dtd) {
        const T dt = (T)dtd;
        T dv;
        if constexpr (INTEG == VG_INTEG_REFERENCE) {
          int nit = 0;
          dv = tt_step_reference<T>(s, V, it, dt, g, a.max_newton, a.newton_jt != 0, &nit);
          newton += nit;
        } else {
          dv = tt_step_rl<T, sizeof(T) == 4>(s, V, it, dt, g);
        }
        V += dt * dv;
        return (double)dv;
      });
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
template <class T, bool ADAPT>
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
  ionic_substeps<ADAPT>(sub, [&](double dtd) {
    const T dt = (T)dtd;
    const T dv = pk_step<T, sizeof(T) == 4>(s, V, stim, dt);
    V += dt * dv;
    return (double)dv;
  });
#pragma unroll
  for (int k = 0; k < PK_NSTATE; k++) st[k * n_alloc + i] = s[k];
  Vout[row] = V;
}

// ---- learned ten Tusscher ionic update ---------------------------------------
// The network runs transposed, Z^T = W^T A^T: the weights are the MFMA A
// operand (read from LDS, row = output unit), the activations the B operand
// (column = node), so every layer's outputs stay in the registers of the
// lanes that own their node.  Lane l holds node l % 32 and half h = l / 32;
// for unit tile nt its 16 results j are units 32 nt + 8 g + 4 h + r4
// (g = j / 4, r4 = j % 4).  The next layer's operand for k step
// ks = 2 nt + s needs units 16 ks + 8 h + q (q < 8): four the lane holds
// (g = 2 s + h), four its partner l ^ 32 holds (same g), so a layer hands
// over with one exchange of four BF16 per k step and no LDS traffic.

// 12 waves, three per SIMD: the weights' LDS image (~132 KB) admits one block
// per CU, so the block's size sets the waves resident (166 VGPRs fit three).
constexpr int SBLK = 768;
constexpr int SWAVES = SBLK / 64;
constexpr int SG_BIAS_TOTAL = (SG_LAYERS - 1) * SG_HID + SG_NOUTP;

// Two floats to BF16 (round to nearest even), a in the low half: one
// two-operand v_cvt_pk_bf16_f32 (converting each float on its own costs a
// conversion each plus a shift and an or).
__device__ inline uint32_t sg_pack2(float a, float b) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

__device__ inline bf16x8 sg_frag(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  return __builtin_bit_cast(bf16x8, u32x4{d0, d1, d2, d3});
}

// acc[nt] = bias + W^T (tiles of 32 units) times the KS operand fragments b:
// the accumulators start at the bias of the units lane half h holds (unit
// 32 nt + 8 (j / 4) + 4 h + j % 4 for result j), so no add follows.
template <int KS, int NT>
__device__ inline void surr_mm(const __bf16 *W, int ldw, int r, int kq, const bf16x8 (&b)[8],
                               const float *bias, int h, f32x16 (&acc)[4]) {
#pragma unroll
  for (int nt = 0; nt < NT; nt++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[nt][j] = bias[nt * 32 + 8 * (j >> 2) + 4 * h + (j & 3)];
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(W + (nt * 32 + r) * ldw + ks * 16 + kq);
      acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[ks], acc[nt], 0, 0, 0);
    }
    // Keep the weight fragments of one k step in flight at a time (the
    // registers go to more waves per SIMD instead).
    __builtin_amdgcn_sched_barrier(0);
  }
}

// SiLU of a hidden layer (bias already in acc), handed over as the next
// layer's fragments.
// For k step 2 nt + s, lane half 0 needs the units of g = 2 s (its own four,
// then its partner's), half 1 those of g = 2 s + 1 (its partner's, then its
// own).  v_permlane32_swap exchanges the upper 32 lanes of its first operand
// with the lower 32 of its second: with P (g = 2 s) first and Q (g = 2 s + 1)
// second, every lane then holds its fragment as {P', Q'}, with no select on
// the lane half and no LDS-routed shuffle.
__device__ inline void surr_hand_over(const f32x16 (&acc)[4], bf16x8 (&b)[8]) {
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    float a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = sg_silu(acc[nt][j]);
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int gp = 8 * s, gq = 8 * s + 4;
      const auto w0 = __builtin_amdgcn_permlane32_swap(sg_pack2(a[gp], a[gp + 1]),
                                                       sg_pack2(a[gq], a[gq + 1]), false, false);
      const auto w1 = __builtin_amdgcn_permlane32_swap(sg_pack2(a[gp + 2], a[gp + 3]),
                                                       sg_pack2(a[gq + 2], a[gq + 3]), false, false);
      b[2 * nt + s] = sg_frag(w0[0], w1[0], w0[1], w1[1]);
    }
  }
}

// The last layer's results j = 4 g + r4 of lane half h are output units
// 8 g + 4 h + r4; slots j < SG_OUT_SLOTS (g < 3) hold every real unit.
constexpr int SG_OUT_SLOTS = 12;
// States clamped to [0, 1] after the increment: the gates and R'.
constexpr uint32_t SG_UNIT_STATES = (0x3Fu << 4) | (0x3Fu << 11) | (1u << 18);

// Checks the real units of lane half h (acc holds their bias) against their
// envelope (oc rows:
// t_lo, t_hi, mu_y, sd_y, c_y) and turns every slot into an increment
// y[j] (padding units: c_y = 0, so 0).  The half is a runtime index, not a
// template argument: both halves of a wave then run one instruction stream
// instead of each half's copy under its own mask.
__device__ inline bool surr_outputs(const f32x16 &acc, const float *oc, int h,
                                    float (&y)[SG_OUT_SLOTS]) {
  // Bitwise, not short-circuit: one instruction stream, no branch per slot
  // (u < 24 stays inside the 32-wide oc rows).
  bool ok = true;
#pragma unroll
  for (int j = 0; j < SG_OUT_SLOTS; j++) {
    const int u = 8 * (j >> 2) + 4 * h + (j & 3);
    const float t = acc[j];
    ok &= (u >= SG_NOUT) | ((t >= oc[u]) & (t <= oc[32 + u]));
    y[j] = sg_from_target(t, oc[64 + u], oc[96 + u], oc[128 + u]);
  }
  return ok;
}

// New V and states of lane half h's units; the start values are read again
// (cache hits) rather than held across the network.
template <class T>
__device__ inline void surr_store(const float (&y)[SG_OUT_SLOTS], T V, int64_t i, int64_t n_alloc,
                                  int h, T *__restrict__ Vout, T *__restrict__ st) {
#pragma unroll
  for (int j = 0; j < SG_OUT_SLOTS; j++) {
    const int k = 8 * (j >> 2) + 4 * h + (j & 3);
    if (k == 0) {
      Vout[i] = V + (T)y[j];
    } else if (k < SG_NOUT) {
      T v = st[(k - 1) * n_alloc + i] + (T)y[j];
      if ((SG_UNIT_STATES >> k) & 1u) v = fmin(fmax(v, T(0)), T(1));
      st[(k - 1) * n_alloc + i] = v;
    }
  }
}

// ten Tusscher ionic update through the learned step for rows [0, n): a row
// whose 25 inputs and 19 outputs lie inside the training envelopes and that
// has no stimulus gets V and its 18 states advanced by the network's
// increments (gates and R' clamped to [0, 1]); the others are flagged in
// `fallback` for the plain kernel.  Persistent blocks keep weights, biases
// and output constants in LDS; each wave takes tiles of 32 rows.
// usage[2 * (block * SWAVES + wave) + {0, 1}] counts network / fallback rows.
template <class T>
__global__ __launch_bounds__(SBLK) void k_ionic_surr(
    const T *__restrict__ Vin, const T *__restrict__ x, const T *__restrict__ sinv,
    T *__restrict__ Vout, T *__restrict__ st, const T *__restrict__ istim,
    const uint16_t *__restrict__ pidx, const T *__restrict__ scales, int64_t n, int64_t n_alloc,
    float log_dt, SurrDevice sd, unsigned char *__restrict__ fallback,
    unsigned long long *__restrict__ usage, const CgResult *cg, const int *__restrict__ slice_inst) {
  __shared__ __attribute__((aligned(16))) __bf16 wl[SG_WT_TOTAL];
  __shared__ float bl[SG_BIAS_TOTAL];
  __shared__ float oc[5 * 32];
  __shared__ float ic[4 * 32];  // lo, hi, mu_x, 1 / sd_x
  for (int q = threadIdx.x; q < SG_WT_TOTAL / 8; q += SBLK)
    reinterpret_cast<bf16x8 *>(wl)[q] = reinterpret_cast<const bf16x8 *>(sd.wt)[q];
  for (int q = threadIdx.x; q < SG_BIAS_TOTAL; q += SBLK) {
    const int L = q / SG_HID < SG_LAYERS - 1 ? q / SG_HID : SG_LAYERS - 1;
    bl[q] = sd.b[L][q - L * SG_HID];
  }
  for (int q = threadIdx.x; q < 5 * 32; q += SBLK) {
    const int row = q / 32, u = q % 32;
    const float *src[5] = {sd.t_lo, sd.t_hi, sd.mu_y, sd.sd_y, sd.c_y};
    oc[q] = u < SG_NOUT ? src[row][u] : 0.0f;
  }
  for (int q = threadIdx.x; q < 4 * 32; q += SBLK) {
    const int row = q / 32, u = q % 32;
    const float *src[4] = {sd.lo, sd.hi, sd.mu_x, sd.sd_x};
    const float v = u < SG_NIN ? src[row][u] : 1.0f;
    ic[q] = row == 3 ? 1.0f / v : v;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 31, h = l >> 5, kq = h * 8;
  unsigned long long n_net = 0, n_fb = 0;
  const int64_t ntiles = (n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * SWAVES + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * SWAVES) {
    const int64_t i = tile * 32 + r;
    const bool mine = i < n;
    T V = T(0);
    float xn[SG_NIN];
    bool ok = false;
    if (mine) {
      V = Vin[i];
      if (x && cg[slice_inst ? slice_inst[i >> 6] : 0].x_valid) V += sinv[i] * x[i];
      const T *sc = scales + (pidx ? (int)pidx[i] : 0) * VG_TT_NSCALE;
      float f[SG_NIN];
      f[0] = (float)V;
#pragma unroll
      for (int k = 1; k < TT_NSTATE; k++) f[k] = (float)st[(k - 1) * n_alloc + i];
      f[19] = log_dt;
#pragma unroll
      for (int k = 0; k < VG_TT_NSCALE; k++) f[20 + k] = (float)sc[k];
      ok = !(istim && istim[i] != T(0)) && !(f[0] > sd.up_v && f[SG_STATE_H] > sd.up_h) &&
           !(f[0] > sd.band_lo && f[0] < sd.band_hi);
#pragma unroll
      for (int k = 0; k < SG_NIN; k++) ok &= (f[k] >= ic[k]) & (f[k] <= ic[32 + k]);
#pragma unroll
      for (int k = 0; k < SG_NIN; k++) xn[k] = ok ? sg_norm_in(f[k], ic[64 + k], ic[96 + k]) : 0.0f;
    } else {
#pragma unroll
      for (int k = 0; k < SG_NIN; k++) xn[k] = 0.0f;
    }
    // Layer 0 operand: input values 16 ks + 8 h + q (his, los, zeros).
    bf16x8 b[8];
#pragma unroll
    for (int ks = 0; ks < SG_KIN / 16; ks++) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; q++)
        v[q] = h ? sg_in_value(xn, 16 * ks + 8 + q) : sg_in_value(xn, 16 * ks + q);
      b[ks] = sg_frag(sg_pack2(v[0], v[1]), sg_pack2(v[2], v[3]), sg_pack2(v[4], v[5]),
                      sg_pack2(v[6], v[7]));
    }
    f32x16 acc[4];
    surr_mm<SG_KIN / 16, 4>(wl + sg_wt_off(0), sg_wt_ld(0), r, kq, b, bl, h, acc);
    surr_hand_over(acc, b);
#pragma unroll 1
    for (int L = 1; L < SG_LAYERS - 1; L++) {
      surr_mm<8, 4>(wl + sg_wt_off(L), sg_wt_ld(L), r, kq, b, bl + L * SG_HID, h, acc);
      surr_hand_over(acc, b);
    }
    surr_mm<8, 1>(wl + sg_wt_off(SG_LAYERS - 1), sg_wt_ld(SG_LAYERS - 1), r, kq, b,
                  bl + (SG_LAYERS - 1) * SG_HID, h, acc);
    float y[SG_OUT_SLOTS];
    const bool ok_out = surr_outputs(acc[0], oc, h, y);
    // The partner half's flag: the swap hands lanes 0-31 the upper half's
    // value in its second result and lanes 32-63 the lower half's in its first.
    const auto sw = __builtin_amdgcn_permlane32_swap((unsigned)ok_out, (unsigned)ok_out, false, false);
    const int both = (int)(h ? sw[0] : sw[1]) & (int)ok_out;
    ok = ok && both;
    if (mine && ok) surr_store<T>(y, V, i, n_alloc, h, Vout, st);
    if (mine && h == 0) {
      fallback[i] = ok ? 0 : 1;
      if (ok) n_net++;
      else n_fb++;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    n_net += __shfl_xor(n_net, o, 64);
    n_fb += __shfl_xor(n_fb, o, 64);
  }
  if (l == 0) {
    usage[2 * (blockIdx.x * SWAVES + wave)] += n_net;
    usage[2 * (blockIdx.x * SWAVES + wave) + 1] += n_fb;
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
  // One instance: whether x holds an increment is the same for every row, so
  // it is read once, and four rows per pass keep their loads in flight
  // together (the grid is capped, so each thread takes many rows).  Not with
  // the ECG sums, whose accumulators the unrolled form slowed down.
  const bool x_one = x && !slice_inst && cg[0].x_valid != 0;
  constexpr int APPLY_UNROLL = ECG ? 1 : 4;
#pragma unroll APPLY_UNROLL
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const bool use_x = slice_inst ? (x && cg[slice_inst[i >> 6]].x_valid != 0) : x_one;
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
// truncated: solves still running when the captured launches ran out (their
// last iterate was used; counted in notconv too).
struct CgTotals {
  unsigned long long iters, solves, notconv, nonfinite, truncated;
  int max_iter[2];
  double max_relres;
};

// One thread: folds the result records of the solve of half-step `which`
// into the totals (fold != 0) and, with tick != 0, advances the step counter.
// A solve that has not ended (done == 0) ran out of captured launches: the
// running count of those goes to trunc_pub, host-coherent memory the host
// reads after every replay without synchronising.
__global__ void k_fold(const CgResult *__restrict__ res, int n_inst, CgTotals *tot, int which,
                       int fold, int64_t *step_ctr, int tick, unsigned long long *trunc_pub) {
  if (threadIdx.x != 0) return;
  if (fold) {
    bool cut = false;
    for (int k = 0; k < n_inst; k++) {
      const CgResult c = res[k];
      tot->iters += (unsigned long long)c.iter;
      tot->solves += 1;
      tot->notconv += (unsigned long long)(c.done == 0 || c.notconv != 0);
      if (c.done == 0) {
        tot->truncated += 1;
        cut = true;
      }
      if (!isfinite(c.rho) || !isfinite(c.rho_ref)) tot->nonfinite += 1;
      if (c.iter > tot->max_iter[which]) tot->max_iter[which] = c.iter;
      const double rel = c.rho_ref > 0.0 ? sqrt(c.rho / c.rho_ref) : 0.0;
      if (rel > tot->max_relres) tot->max_relres = rel;
    }
    if (cut && trunc_pub)
      __hip_atomic_store(trunc_pub, tot->truncated, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if (tick) step_ctr[0] += 1;
}

// Position of entry j of a lane of SELL slice `slice` (the layout sell_dot
// reads) and, in col, the row index it gathers from.
__device__ inline int64_t sell_pos(const Sell &S, int64_t slice, int lane, int j, int64_t &col) {
  const int64_t o = S.off[slice];
  const int w4 = S.width[slice] & ~3;
  const int64_t rel = j < w4 ? (int64_t)(j >> 2) * 256 + 4 * lane + (j & 3)
                             : (int64_t)w4 * 64 + (int64_t)(j - w4) * 64 + lane;
  const int64_t us = S.ustart[slice];
  const int64_t wo = S.wide[slice];
  col = us >= 0   ? slice * 64 + lane + S.uoff[us + j]
        : wo < 0  ? slice * 64 + lane + S.col16[o + rel]
                  : (int64_t)S.col32[wo + rel];
  return o + rel;
}

// True residual of the solve just ended, in FP64 arithmetic on the stored
// operator: b = K_rhs V and r = b - A_hat x with every product and sum in
// FP64 (the CG's own stop test uses the recurrence residual, which in FP32
// drifts from b - A_hat x).  Partials of ||r||_D^2 and ||b||_D^2 per block;
// x counts as zero when the solve left no increment (x_valid = 0).
// The reference reports the same quantity after its solve
// (EP_CrankNicholson.cc:104-109, TrueResidual()).
template <class T>
__global__ __launch_bounds__(BLK) void k_true_res(Sell S, const T *__restrict__ Krhs,
                                                  const T *__restrict__ Ahat,
                                                  const T *__restrict__ V,
                                                  const T *__restrict__ x,
                                                  const T *__restrict__ diag, CgBufs cb,
                                                  double *__restrict__ part) {
  int64_t lo, hi, step;
  int b0, b1;
  const int inst = block_work(cb, lo, hi, step, b0, b1);
  const bool use_x = x && cb.res[inst].x_valid;
  const int lane = threadIdx.x & 63;
  double v[2] = {0.0, 0.0};
  for (int64_t sl = lo + (threadIdx.x >> 6); sl < hi; sl += step) {
    const int64_t row = sl * 64 + lane;
    const int w = S.width[sl];
    const double vi = (double)V[row];
    double b = 0.0, ax = use_x ? (double)x[row] : 0.0;  // A_hat's unit diagonal
    for (int j = 0; j < w; j++) {
      int64_t c;
      const int64_t e = sell_pos(S, sl, lane, j, c);
      b += (double)Krhs[e] * ((double)V[c] - vi);
      if (use_x) ax += (double)Ahat[e] * (double)x[c];
    }
    const double r = b - ax, d = (double)diag[row];
    v[0] += d * r * r;
    v[1] += d * b * b;
  }
  put_partials<2>(v, part);
}

// One block: per instance sqrt(||r||_D^2 / ||b||_D^2) from k_true_res's
// partials (instances with b = 0 skipped); out[0] = the largest seen so far,
// out[1] = checks so far, out[2] = this check's largest.
__global__ __launch_bounds__(BLK) void k_true_finish(const double *__restrict__ part, int nblk,
                                                     CgBufs cb, double *__restrict__ out) {
  double worst = 0.0;
  for (int k = 0; k < cb.n_inst; k++) {
    const int b0 = cb.n_inst == 1 ? 0 : cb.inst_blk[k];
    const int b1 = cb.n_inst == 1 ? nblk : cb.inst_blk[k + 1];
    double t[2];
    sum_partials<2>(part, nblk, b0, b1, t);
    if (t[1] > 0.0) worst = fmax(worst, sqrt(t[0] / t[1]));
    __syncthreads();  // the next sum reuses sum_partials' shared slots
  }
  if (threadIdx.x == 0) {
    out[0] = fmax(out[0], worst);
    out[1] += 1.0;
    out[2] = worst;
  }
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
                                               int max_newton, int newton_jt, int mj_int_abs,
                                               const T *__restrict__ lr_table) {
  const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const T V = (T)Vd[i], stim = stim_d ? (T)stim_d[i] : T(0), Cm = (T)Cm_d, dt = (T)dt_d;
  const int set = pidx ? (int)pidx[i] : 0;
  const T *sc = scales + set * VG_TT_NSCALE;
  if constexpr (MODEL == VG_MODEL_LR) {
    T s[LR_NSTATE];
#pragma unroll
    for (int k = 0; k < LR_NSTATE - 1; k++) s[k] = st[k * n_alloc + i];
    dvdt[i] = (double)lr_compute_ion<T, sizeof(T) == 4>(s, Cm, dt, V, stim, sc[0], sc[3], lr_table);
#pragma unroll
    for (int k = 0; k < LR_NSTATE - 1; k++) st[k * n_alloc + i] = s[k];
  } else if constexpr (MODEL == VG_MODEL_MJ || MODEL == VG_MODEL_MJ_FAIL) {
    T s[MJ_NSTATE];
    double ac[MJ_NACC];
#pragma unroll
    for (int k = 0; k < MJ_NGATE; k++) s[MJ_GATE[k]] = st[k * n_alloc + i];
#pragma unroll
    for (int k = 0; k < MJ_NACC; k++) {
      ac[k] = acc[k * n_alloc + i];
      s[MJ_ACC[k]] = (T)ac[k];
    }
    dvdt[i] = (double)mj_step<T, sizeof(T) == 4, MODEL == VG_MODEL_MJ_FAIL>(s, ac, V, stim / (Cm * T(MJ_XI)), dt, sc[0],
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
