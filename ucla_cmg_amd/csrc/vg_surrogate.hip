// vg_surrogate.hip -- training of the learned ten Tusscher ionic step with
// hand-written BF16 MFMA kernels (forward, backward, Adam) and its C ABI.
//
// Parameters are kept in FP32 (master copy, padded layout: per layer W[kp][np]
// row-major then b[np], padding held at zero) and converted to BF16 for the
// GEMMs after every update.  Activations between layers are BF16, all sums
// FP32; loss and bias gradients are summed in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/voomgpu.h"
#include "vg_surrogate.hpp"

using namespace vg;

namespace {

constexpr int GB = 256;  // threads per block of the elementwise kernels

// Offset of layer l's W in the padded parameter array (W[kp][np], b[np] per layer).
__host__ __device__ constexpr int64_t LAY_W_OFF(int l) {
  return l == 0 ? 0 : LAY_W_OFF(l - 1) + (int64_t)sg_kp(l - 1) * sg_np(l - 1) + sg_np(l - 1);
}

// Offsets of each layer's W and b in the padded parameter array.
struct Layout {
  int64_t w[SG_LAYERS], b[SG_LAYERS], total;
  Layout() {
    int64_t o = 0;
    for (int l = 0; l < SG_LAYERS; l++) {
      w[l] = o;
      o += (int64_t)sg_kp(l) * sg_np(l);
      b[l] = o;
      o += sg_np(l);
    }
    total = o;
  }
};
const Layout LAY;

// Unpadded parameter count: per layer W[k][n] then b[n].
int64_t n_params_unpadded() {
  int64_t o = 0;
  for (int l = 0; l < SG_LAYERS; l++) o += (int64_t)sg_k(l) * sg_n(l) + sg_n(l);
  return o;
}

// What a GEMM does with its result tile (row-major, ldc columns):
//   EPI_STORE    C[z] = the product;
//   EPI_BIAS_ACT Z = product + bias on the real columns (0 on padding), and
//                when act, A = SiLU(Z) in BF16 (a layer's forward);
//   EPI_ACT_BACK dZ = product * SiLU'(Zin) on the real columns, FP32 and
//                BF16 (the backward through a hidden layer's activation).
constexpr int EPI_STORE = 0, EPI_BIAS_ACT = 1, EPI_ACT_BACK = 2;
struct Epi {
  const float *bias = nullptr;
  int nreal = 0, act = 0;
  float *Z = nullptr;
  __bf16 *A = nullptr;
  const float *Zin = nullptr;
  float *dZf = nullptr;
  __bf16 *dZ = nullptr;
};

// ---- GEMM: C[z] = op(A) op(B) on MFMA -----------------------------------------
// op(A) is M x K (stored [M][lda], or [K][lda] when AT), op(B) is K x N
// (stored [K][ldb], or [N][ldb] when BT).  One wave per 32 x 32 tile of C,
// four waves per block stacked along M; blockIdx.z takes the k range
// [z*kchunk, (z+1)*kchunk) and writes C + z*zstride (split K).  M, N are
// multiples of 32, K and kchunk of 16, lda/ldb of 8.  The epilogue EPI
// (above) is applied in registers, with C's layout.
template <bool AT, bool BT, int EPI = EPI_STORE>
__global__ __launch_bounds__(256) void k_gemm(const __bf16 *__restrict__ A, int lda,
                                              const __bf16 *__restrict__ B, int ldb,
                                              float *__restrict__ C, int ldc, int M, int N, int K,
                                              int kchunk, int64_t zstride, Epi ep = Epi{}) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + wave) * 32, col0 = blockIdx.y * 32;
  if (row0 >= M) return;
  const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
  const int r = l & 31, kq = (l >> 5) * 8;
  f32x16 acc = {};
  for (int k = k0; k < k1; k += 16) {
    bf16x8 a, b;
    if constexpr (!AT) {
      a = *reinterpret_cast<const bf16x8 *>(A + (size_t)(row0 + r) * lda + k + kq);
    } else {
#pragma unroll
      for (int q = 0; q < 8; q++) a[q] = A[(size_t)(k + kq + q) * lda + row0 + r];
    }
    if constexpr (BT) {
      b = *reinterpret_cast<const bf16x8 *>(B + (size_t)(col0 + r) * ldb + k + kq);
    } else {
#pragma unroll
      for (int q = 0; q < 8; q++) b[q] = B[(size_t)(k + kq + q) * ldb + col0 + r];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int n = col0 + r;
  if constexpr (EPI == EPI_STORE) {
    float *Cz = C + (int64_t)blockIdx.z * zstride;
#pragma unroll
    for (int j = 0; j < 16; j++) Cz[(size_t)(row0 + mfma_row(l, j)) * ldc + n] = acc[j];
  } else if constexpr (EPI == EPI_BIAS_ACT) {
    const bool real = n < ep.nreal;
    const float bn = real ? ep.bias[n] : 0.0f;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const size_t i = (size_t)(row0 + mfma_row(l, j)) * ldc + n;
      const float z = real ? acc[j] + bn : 0.0f;
      ep.Z[i] = z;
      if (ep.act) ep.A[i] = (__bf16)(real ? sg_silu(z) : 0.0f);
    }
  } else {
    const bool real = n < ep.nreal;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const size_t i = (size_t)(row0 + mfma_row(l, j)) * ldc + n;
      float g = 0.0f;
      if (real) {
        const float z = ep.Zin[i];
        const float sg = sg_sigmoid(z);
        g = acc[j] * sg * (1.0f + z * (1.0f - sg));
      }
      ep.dZf[i] = g;
      ep.dZ[i] = (__bf16)g;
    }
  }
}

// ---- weight gradient: P[z] = A^T D over a batch chunk, staged through LDS ------
// A is [rows][M] and D [rows][N] (BF16, row-major, M and N multiples of 32 up
// to 128); block z takes the batch rows [z*kchunk, (z+1)*kchunk) (rows past
// the chunk's end read as zero) and writes
// the M x N partial P + z*M*N (row-major).  Both operands need their batch
// index along the MFMA k dimension, i.e. transposed: the block copies 64
// batch rows at a time from HBM with 16-byte row reads and stores them
// transposed into LDS (rows padded by 8), from where each MFMA fragment is one
// 16-byte read.  Four waves share the (M/32) x (N/32) output tiles.
constexpr int TN_ROWS = 64, TN_LD = TN_ROWS + 8;
constexpr int TN_CHUNKS = 256;  // batch chunks (blocks) of a weight-gradient GEMM
__global__ __launch_bounds__(256) void k_gemm_tn(const __bf16 *__restrict__ A, int M,
                                                 const __bf16 *__restrict__ D, int N, int kchunk,
                                                 float *__restrict__ P) {
  __shared__ __attribute__((aligned(16))) __bf16 At[SG_HID * TN_LD];
  __shared__ __attribute__((aligned(16))) __bf16 Dt[SG_HID * TN_LD];
  const int t = threadIdx.x, wave = t >> 6, l = t & 63, r = l & 31, kq = (l >> 5) * 8;
  const int tm = M / 32, ntiles = tm * (N / 32);
  f32x16 acc[4];
#pragma unroll
  for (int q = 0; q < 4; q++) acc[q] = f32x16{};
  const int64_t k0 = (int64_t)blockIdx.x * kchunk;
  for (int kk = 0; kk < kchunk; kk += TN_ROWS) {
    __syncthreads();  // the previous sub-chunk's fragments have been read
    // Copy TN_ROWS rows of A (M columns) and D (N columns), 8 columns per
    // thread and pass, transposed into At[m][row] and Dt[n][row].
    for (int e = t; e < TN_ROWS * (M / 8); e += 256) {
      const int row = e % TN_ROWS, g = e / TN_ROWS;
      bf16x8 v = {};
      if (kk + row < kchunk) v = *reinterpret_cast<const bf16x8 *>(A + (k0 + kk + row) * M + g * 8);
#pragma unroll
      for (int q = 0; q < 8; q++) At[(g * 8 + q) * TN_LD + row] = v[q];
    }
    for (int e = t; e < TN_ROWS * (N / 8); e += 256) {
      const int row = e % TN_ROWS, g = e / TN_ROWS;
      bf16x8 v = {};
      if (kk + row < kchunk) v = *reinterpret_cast<const bf16x8 *>(D + (k0 + kk + row) * N + g * 8);
#pragma unroll
      for (int q = 0; q < 8; q++) Dt[(g * 8 + q) * TN_LD + row] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tile = wave + 4 * q;
      if (tile < ntiles) {
        const int mt = tile % tm, nt = tile / tm;
#pragma unroll
        for (int ks = 0; ks < TN_ROWS; ks += 16) {
          const bf16x8 a = *reinterpret_cast<const bf16x8 *>(At + (mt * 32 + r) * TN_LD + ks + kq);
          const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Dt + (nt * 32 + r) * TN_LD + ks + kq);
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
        }
      }
    }
  }
  float *Pz = P + (int64_t)blockIdx.x * M * N;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int tile = wave + 4 * q;
    if (tile < ntiles) {
      const int mt = tile % tm, nt = tile / tm;
#pragma unroll
      for (int j = 0; j < 16; j++)
        Pz[(size_t)(mt * 32 + mfma_row(l, j)) * N + nt * 32 + r] = acc[q][j];
    }
  }
}

// ---- elementwise kernels ---------------------------------------------------------
__device__ inline uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Batch of B samples (hash-drawn from Ns, or rows first.. when idx_seed is 0
// and first >= 0), normalised: X0[B][64] BF16 (layer 0's values: his, los, zeros),
// Yn[B][19].
__global__ void k_sg_gather(const float *__restrict__ X, const float *__restrict__ Y, int64_t Ns,
                            uint64_t seed, const int64_t *__restrict__ step, int64_t first, int B,
                            const float *__restrict__ mu_x,
                            const float *__restrict__ sd_x, const float *__restrict__ mu_y,
                            const float *__restrict__ sd_y, const float *__restrict__ c_y,
                            __bf16 *__restrict__ X0,
                            float *__restrict__ Yn) {
  const int i = blockIdx.x * GB + threadIdx.x;
  if (i >= B) return;
  // Training batches: the seed of Adam step t + 1 (step holds t).
  if (step) seed = seed * 0x100000001B3ull + (uint64_t)(*step + 1) * 0x9E3779B97F4A7C15ull;
  const int64_t s = first >= 0 ? first + i : (int64_t)(splitmix(seed + (uint64_t)i) % (uint64_t)Ns);
  float xn[SG_NIN];
  for (int f = 0; f < SG_NIN; f++) xn[f] = sg_norm_in(X[s * SG_NIN + f], mu_x[f], 1.0f / sd_x[f]);
  for (int f = 0; f < SG_KIN; f++) X0[(size_t)i * SG_KIN + f] = (__bf16)sg_in_value(xn, f);
  for (int f = 0; f < SG_NOUT; f++)
    Yn[(size_t)i * SG_NOUT + f] = sg_to_target(Y[s * SG_NOUT + f], mu_y[f], sd_y[f], c_y[f]);
}

// Mean-square loss over the 19 outputs: dZ = 2 (pred - y) / (B*19) (FP32 and
// BF16), per-block partial of sum (pred - y)^2 (fixed order).
__global__ void k_sg_loss(const float *__restrict__ Z, const float *__restrict__ Yn, int B,
                          float *__restrict__ dZf, __bf16 *__restrict__ dZ, double *__restrict__ part) {
  __shared__ double sm[GB];
  const int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
  double s = 0.0;
  if (i < (int64_t)B * SG_NOUTP) {
    const int n = (int)(i % SG_NOUTP);
    const int64_t row = i / SG_NOUTP;
    float g = 0.0f;
    if (n < SG_NOUT) {
      const float d = Z[i] - Yn[row * SG_NOUT + n];
      s = (double)d * d;
      g = 2.0f * d / (float)((int64_t)B * SG_NOUT);
    }
    dZf[i] = g;
    dZ[i] = (__bf16)g;
  }
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = GB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Column sums of M[rows][np] in two fixed-order stages.  Stage 1: block b
// takes rows [64 b, 64 b + 64); thread t reads column t % np of every
// (GB/np)-th of them (all loads issued, then summed in order), and the block
// adds its GB/np lanes per column into part[b][np].
__global__ void k_sg_colsum_part(const float *__restrict__ Mx, int rows, int np,
                                 float *__restrict__ part) {
  __shared__ float sm[GB];
  const int t = threadIdx.x, lanes = GB / np, c = t % np, q = t / np;
  const int r0 = blockIdx.x * 64, per = 64 / lanes;
  float v[32];
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int r = r0 + q + j * lanes;
    v[j] = (j < per && r < rows) ? Mx[(size_t)r * np + c] : 0.0f;
  }
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < 32; j++) s += v[j];
  sm[t] = s;
  __syncthreads();
  if (t < np) {
    float a = 0.0f;
    for (int k = 0; k < lanes; k++) a += sm[k * np + t];
    part[(size_t)blockIdx.x * np + t] = a;
  }
}

// Stage 2, one block per column n: out[n] = the sum of the nparts partials of
// column n by a fixed tree over the block.
__global__ void k_sg_colsum_fin(const float *__restrict__ part, int nparts, int np,
                                float *__restrict__ out) {
  __shared__ float sm[GB];
  const int n = blockIdx.x;
  float s = 0.0f;
  for (int z = threadIdx.x; z < nparts; z += GB) s += part[(size_t)z * np + n];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int o = GB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = sm[0];
}

// out[i] = sum_z P[z][i] on the real (k < kreal, n < nreal) entries, 0 else;
// the loads are issued 16 at a time, the sum taken in z order.
__global__ void k_sg_reduce_z(const float *__restrict__ P, int nz, int np, int kreal, int nreal,
                              int64_t mn, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
  if (i >= mn) return;
  const int n = (int)(i % np), k = (int)(i / np);
  float s = 0.0f;
  if (k < kreal && n < nreal) {
    int z = 0;
    for (; z + 16 <= nz; z += 16) {
      float v[16];
#pragma unroll
      for (int q = 0; q < 16; q++) v[q] = P[(int64_t)(z + q) * mn + i];
#pragma unroll
      for (int q = 0; q < 16; q++) s += v[q];
    }
    for (; z < nz; z++) s += P[(int64_t)z * mn + i];
  }
  out[i] = s;
}

// Adam's step count and learning rate, in device memory.
struct TrainState {
  int64_t t;
  float lr;
};

// Advances the Adam step count (after the step's kernels have read it).
__global__ void k_sg_tick(TrainState *ts) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ts->t += 1;
}

// The BF16 copies of the parameters are pbf[0, n) in the padded layout and,
// at pbf[n + ...], every layer's W transposed (W^T[np][kp] at W's offset), so
// the forward GEMMs read whole 16-byte weight fragments.  The transposed
// index of padded parameter i, or -1 for a bias.
__device__ inline int64_t sg_transposed(int64_t i) {
#pragma unroll
  for (int l = 0; l < SG_LAYERS; l++) {
    const int64_t w = LAY_W_OFF(l), nw = (int64_t)sg_kp(l) * sg_np(l);
    if (i >= w && i < w + nw) {
      const int64_t o = i - w;
      return w + (o % sg_np(l)) * sg_kp(l) + o / sg_np(l);
    }
  }
  return -1;
}

__device__ inline void sg_store_bf16(__bf16 *pbf, int64_t n, int64_t i, float v) {
  const __bf16 q = (__bf16)v;
  pbf[i] = q;
  const int64_t t = sg_transposed(i);
  if (t >= 0) pbf[n + t] = q;
}

// Adam on the padded parameter array; padding entries stay zero.  Then the
// BF16 copies of the weights.
// Step t + 1 of Adam, t and the learning rate read from the training state
// (k_sg_tick advances t afterwards), so a captured iteration replays as is.
__global__ void k_sg_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, const unsigned char *__restrict__ mask, int64_t n,
                          const TrainState *__restrict__ ts, float b1, float b2, float eps,
                          __bf16 *__restrict__ pbf) {
  const int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
  if (i >= n) return;
  const float tt = (float)(ts->t + 1), lr = ts->lr;
  const float c1 = 1.0f - powf(b1, tt), c2 = 1.0f - powf(b2, tt);
  if (!mask[i]) {
    p[i] = 0.0f;
    sg_store_bf16(pbf, n, i, 0.0f);
    return;
  }
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  sg_store_bf16(pbf, n, i, pi);
}

// The step kernel's LDS image: wt[off(l) + j*ld(l) + k] = W_l[k][j] (BF16),
// zero on the row padding.
__global__ void k_sg_wt_image(const float *__restrict__ p, __bf16 *__restrict__ wt) {
  const int i = blockIdx.x * GB + threadIdx.x;
  if (i >= SG_WT_TOTAL) return;
  int l = 0;
  while (l + 1 < SG_LAYERS && i >= sg_wt_off(l + 1)) l++;
  const int o = i - sg_wt_off(l), j = o / sg_wt_ld(l), k = o % sg_wt_ld(l);
  wt[i] = (__bf16)(k < sg_kp(l) ? p[LAY_W_OFF(l) + (int64_t)k * sg_np(l) + j] : 0.0f);
}

__global__ void k_sg_to_bf16(const float *__restrict__ p, int64_t n, __bf16 *__restrict__ pbf) {
  const int64_t i = (int64_t)blockIdx.x * GB + threadIdx.x;
  if (i < n) sg_store_bf16(pbf, n, i, p[i]);
}

// Per-block partial sums, sums of squares, minima, maxima and sums of |x| of
// F features over the count samples of X[.][F] (of asinh(x / c[f]) when c is
// given); thread f < F owns feature f.
__global__ void k_sg_stats(const float *__restrict__ X, int F, int64_t count,
                           const float *__restrict__ c, double *__restrict__ part) {
  const int f = threadIdx.x;
  const int64_t per = (count + gridDim.x - 1) / gridDim.x;
  const int64_t s0 = (int64_t)blockIdx.x * per, s1 = min(count, s0 + per);
  double s = 0.0, q = 0.0, lo = 1e300, hi = -1e300, a = 0.0;
  if (f < F)
    for (int64_t i = s0; i < s1; i++) {
      const double x = c ? (double)asinhf(X[i * F + f] / c[f]) : (double)X[i * F + f];
      s += x;
      q += x * x;
      lo = fmin(lo, x);
      hi = fmax(hi, x);
      a += fabs(x);
    }
  if (f < F) {
    double *o = part + ((size_t)blockIdx.x * F + f) * 5;
    o[0] = s;
    o[1] = q;
    o[2] = lo;
    o[3] = hi;
    o[4] = a;
  }
}

}  // namespace

struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t alloc(size_t b) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = b;
    return hipMalloc(&p, b ? b : 1);
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <class X> X *as() const { return static_cast<X *>(p); }
};

struct vg_surr {
  std::string err;
  hipStream_t stream = nullptr;
  uint64_t seed = 1;
  // Parameters (padded), Adam moments, BF16 copy, mask of real entries.
  DBuf par, grad, m1, m2, parbf, mask;
  // What a time step reads, copied at attach: the transposed BF16 image of
  // the weights, and the biases, scaling and envelope (FP32).
  DBuf wt, dep;
  int64_t adam_t = 0;
  DBuf ts;  // TrainState mirror of adam_t and the learning rate
  // A captured training iteration (gather, forward, backward, Adam, tick),
  // valid while the batch size, the dataset (buffers and count) and the work
  // buffers are those it was captured with.
  hipGraphExec_t graph = nullptr;
  int graph_batch = 0;
  const void *graph_x = nullptr;
  int64_t graph_count = 0;
  int work_gen = 0, graph_gen = -1;
  // Scaling and envelope [SG_NIN | SG_NOUT] and whether fitted.
  DBuf scl;  // one vg_surr_scaling
  bool scaled = false;
  // Upstroke fallback thresholds (V in mV, the h gate); +inf: off.
  float up_v = INFINITY, up_h = INFINITY;
  // Voltage-band fallback (mV); empty: off.
  float band_lo = INFINITY, band_hi = -INFINITY;
  // Dataset (FP32 raw features / increments).
  DBuf X, Y;
  int64_t count = 0, cap = 0;
  // Training work space for batches up to wb samples.
  int wb = 0;
  DBuf X0, Yn, Zs, As, dZf, dZ, dZ2, P, cpart, lpart;
  int fail(int c, const std::string &m) {
    err = m;
    return c;
  }
  ~vg_surr() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Device address of a field of the scaling.
#define SCL(s, f) ((float *)((char *)(s)->scl.p + offsetof(vg_surr_scaling, f)))

#define SCHK(s, call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess) return (s)->fail(VG_ERR_HIP, std::string(#call) + ": " + \
                                                           hipGetErrorString(_e)); \
  } while (0)

namespace vg {

bool surr_dataset(vg_surr *s, float **x, float **y, int64_t *count, int64_t *cap) {
  if (!s || !s->X.p) return false;
  *x = s->X.as<float>();
  *y = s->Y.as<float>();
  *count = s->count;
  *cap = s->cap;
  return true;
}

void surr_set_count(vg_surr *s, int64_t count) { s->count = std::min(count, s->cap); }

bool surr_device(vg_surr *s, SurrDevice *d) {
  if (!s || !s->scaled) return false;
  if (!s->wt.p && s->wt.alloc((size_t)SG_WT_TOTAL * 2) != hipSuccess) return false;
  hipLaunchKernelGGL(k_sg_wt_image, dim3((SG_WT_TOTAL + GB - 1) / GB), dim3(GB), 0, s->stream,
                     s->par.as<float>(), s->wt.as<__bf16>());
  if (hipStreamSynchronize(s->stream) != hipSuccess) return false;
  d->wt = s->wt.as<__bf16>();
  int64_t nf = (int64_t)(sizeof(vg_surr_scaling) / 4);
  for (int l = 0; l < SG_LAYERS; l++) nf += sg_np(l);
  if (!s->dep.p && s->dep.alloc((size_t)nf * 4) != hipSuccess) return false;
  float *o = s->dep.as<float>();
  auto take = [&](const void *src, int64_t count) {
    const float *p = o;
    o += count;
    return hipMemcpyAsync((void *)p, src, count * 4, hipMemcpyDeviceToDevice, s->stream) ==
                   hipSuccess ? p : nullptr;
  };
  for (int l = 0; l < SG_LAYERS; l++) d->b[l] = take(s->par.as<float>() + LAY.b[l], sg_np(l));
  const float *sc = take(s->scl.p, (int64_t)(sizeof(vg_surr_scaling) / 4));
  if (hipStreamSynchronize(s->stream) != hipSuccess || !sc) return false;
  for (int l = 0; l < SG_LAYERS; l++)
    if (!d->b[l]) return false;
  auto field = [&](size_t off) { return sc + off / 4; };
  d->mu_x = field(offsetof(vg_surr_scaling, mu_x));
  d->sd_x = field(offsetof(vg_surr_scaling, sd_x));
  d->lo = field(offsetof(vg_surr_scaling, lo));
  d->hi = field(offsetof(vg_surr_scaling, hi));
  d->mu_y = field(offsetof(vg_surr_scaling, mu_y));
  d->sd_y = field(offsetof(vg_surr_scaling, sd_y));
  d->c_y = field(offsetof(vg_surr_scaling, c_y));
  d->t_lo = field(offsetof(vg_surr_scaling, t_lo));
  d->t_hi = field(offsetof(vg_surr_scaling, t_hi));
  d->up_v = s->up_v;
  d->up_h = s->up_h;
  d->band_lo = s->band_lo;
  d->band_hi = s->band_hi;
  return true;
}

}  // namespace vg

namespace {

int ensure_work(vg_surr *s, int B) {
  if (B <= s->wb) return VG_OK;
  s->work_gen++;  // a captured training iteration holds the old buffers
  const size_t b = (size_t)B;
  SCHK(s, s->X0.alloc(b * SG_KIN * 2));
  SCHK(s, s->Yn.alloc(b * SG_NOUT * 4));
  SCHK(s, s->Zs.alloc(b * SG_HID * 4 * SG_LAYERS));
  SCHK(s, s->As.alloc(b * SG_HID * 2 * SG_LAYERS));
  SCHK(s, s->dZ2.alloc(b * SG_HID * 2));
  SCHK(s, s->cpart.alloc(((b + 63) / 64) * SG_HID * 4));
  SCHK(s, s->dZf.alloc(b * SG_HID * 4));
  SCHK(s, s->dZ.alloc(b * SG_HID * 2));
  SCHK(s, s->P.alloc((size_t)TN_CHUNKS * SG_HID * SG_HID * 4));
  SCHK(s, s->lpart.alloc(((b * SG_NOUTP + GB - 1) / GB) * 8));
  s->wb = B;
  return VG_OK;
}

unsigned nb(int64_t n) { return (unsigned)((n + GB - 1) / GB); }

// Forward of a batch already in X0 (and targets in Yn): per-layer Z and BF16
// activations.  Then, with grads, the backward into s->grad.  Returns the
// summed squared error of the batch in *sse.
int fwd_bwd(vg_surr *s, int B, bool grads, double *sse) {
  hipStream_t st = s->stream;
  const float *par = s->par.as<float>();
  const __bf16 *pbf = s->parbf.as<__bf16>();
  float *Zl[SG_LAYERS];
  __bf16 *Al[SG_LAYERS];  // Al[l] = input of layer l (Al[0] = X0)
  for (int l = 0; l < SG_LAYERS; l++) {
    Zl[l] = s->Zs.as<float>() + (size_t)l * B * SG_HID;
    Al[l] = l == 0 ? s->X0.as<__bf16>() : s->As.as<__bf16>() + (size_t)l * B * SG_HID;
  }
  for (int l = 0; l < SG_LAYERS; l++) {
    const int kp = sg_kp(l), np = sg_np(l);
    const bool last = l == SG_LAYERS - 1;
    Epi ep;
    ep.bias = par + LAY.b[l];
    ep.nreal = sg_n(l);
    ep.act = last ? 0 : 1;
    ep.Z = Zl[l];
    ep.A = last ? (__bf16 *)nullptr : Al[l + 1];
    hipLaunchKernelGGL((k_gemm<false, true, EPI_BIAS_ACT>), dim3((B + 127) / 128, np / 32, 1),
                       dim3(256), 0, st, Al[l], kp, pbf + LAY.total + LAY.w[l], kp,
                       (float *)nullptr, np, B, np, kp, kp, (int64_t)0, ep);
  }
  const unsigned lb = nb((int64_t)B * SG_NOUTP);
  hipLaunchKernelGGL(k_sg_loss, dim3(lb), dim3(GB), 0, st, Zl[SG_LAYERS - 1], s->Yn.as<float>(), B,
                     s->dZf.as<float>(), s->dZ.as<__bf16>(), s->lpart.as<double>());
  if (grads) {
    float *g = s->grad.as<float>();
    for (int l = SG_LAYERS - 1; l >= 0; l--) {
      const int kp = sg_kp(l), np = sg_np(l);
      // dW = A_l^T dZ_l, split over the batch in nz chunks.
      int nz = std::max(1, std::min(TN_CHUNKS, B / TN_ROWS));
      while (B % nz) nz--;
      const int kc = B / nz;
      hipLaunchKernelGGL(k_gemm_tn, dim3(nz), dim3(256), 0, st, Al[l], kp, s->dZ.as<__bf16>(), np,
                         kc, s->P.as<float>());
      hipLaunchKernelGGL(k_sg_reduce_z, dim3(nb((int64_t)kp * np)), dim3(GB), 0, st,
                         s->P.as<float>(), nz, np, sg_k(l), sg_n(l), (int64_t)kp * np,
                         g + LAY.w[l]);
      const int parts = (B + 63) / 64;
      hipLaunchKernelGGL(k_sg_colsum_part, dim3(parts), dim3(GB), 0, st, s->dZf.as<float>(), B,
                         np, s->cpart.as<float>());
      hipLaunchKernelGGL(k_sg_colsum_fin, dim3(np), dim3(GB), 0, st, s->cpart.as<float>(), parts,
                         np, g + LAY.b[l]);
      if (l > 0) {
        // dA_{l-1} = dZ_l W_l^T, then through SiLU'(Z_{l-1}).
        // (dZ of layer l is read by this GEMM and overwritten by its
        // epilogue: the next layer's dZ goes to the other buffer.)
        Epi ep;
        ep.nreal = sg_n(l - 1);
        ep.Zin = Zl[l - 1];
        ep.dZf = s->dZf.as<float>();
        ep.dZ = s->dZ2.as<__bf16>();
        hipLaunchKernelGGL((k_gemm<false, true, EPI_ACT_BACK>), dim3((B + 127) / 128, kp / 32, 1),
                           dim3(256), 0, st, s->dZ.as<__bf16>(), np, pbf + LAY.w[l], np,
                           (float *)nullptr, kp, B, kp, np, np, (int64_t)0, ep);
        std::swap(s->dZ.p, s->dZ2.p);
      }
    }
  }
  SCHK(s, hipGetLastError());
  if (sse) {
    std::vector<double> lp(lb);
    SCHK(s, hipMemcpyAsync(lp.data(), s->lpart.p, lb * 8, hipMemcpyDeviceToHost, st));
    SCHK(s, hipStreamSynchronize(st));
    double t = 0.0;
    for (double v : lp) t += v;
    *sse = t;
  }
  return VG_OK;
}

// Host arrays -> normalised batch in X0/Yn (n <= wb), through a staging copy
// of the raw samples at the end of the dataset buffers' space.
int load_host_batch(vg_surr *s, const float *x, const float *y, int n) {
  DBuf xs, ys;
  SCHK(s, xs.alloc((size_t)n * SG_NIN * 4));
  SCHK(s, ys.alloc((size_t)n * SG_NOUT * 4));
  SCHK(s, hipMemcpy(xs.p, x, (size_t)n * SG_NIN * 4, hipMemcpyHostToDevice));
  if (y) SCHK(s, hipMemcpy(ys.p, y, (size_t)n * SG_NOUT * 4, hipMemcpyHostToDevice));
  else SCHK(s, hipMemset(ys.p, 0, (size_t)n * SG_NOUT * 4));
  hipLaunchKernelGGL(k_sg_gather, dim3(nb(n)), dim3(GB), 0, s->stream, xs.as<float>(),
                     ys.as<float>(), (int64_t)n, (uint64_t)0, (const int64_t *)nullptr, (int64_t)0,
                     n, SCL(s, mu_x),
                     SCL(s, sd_x), SCL(s, mu_y), SCL(s, sd_y),
                     SCL(s, c_y), s->X0.as<__bf16>(), s->Yn.as<float>());
  SCHK(s, hipStreamSynchronize(s->stream));
  return VG_OK;
}

}  // namespace

extern "C" {

int vg_surr_create(uint64_t seed, vg_surr **out) {
  if (!out) return VG_ERR_ARG;
  std::unique_ptr<vg_surr> s(new vg_surr());
  s->seed = seed ? seed : 1;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return VG_ERR_HIP;
  const int64_t n = LAY.total;
  std::vector<float> p(n, 0.0f);
  std::vector<unsigned char> mk(n, 0);
  // He-style uniform initialisation of the real weights, zero biases.
  uint64_t st = s->seed;
  auto urand = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return ((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
  };
  for (int l = 0; l < SG_LAYERS; l++) {
    const double a = std::sqrt(6.0 / sg_k(l));
    for (int k = 0; k < sg_kp(l); k++)
      for (int j = 0; j < sg_np(l); j++) {
        const int64_t i = LAY.w[l] + (int64_t)k * sg_np(l) + j;
        if (k < sg_k(l) && j < sg_n(l)) {
          p[i] = (float)(a * urand() * (l == SG_LAYERS - 1 ? 0.1 : 1.0));
          mk[i] = 1;
        }
      }
    for (int j = 0; j < sg_n(l); j++) mk[LAY.b[l] + j] = 1;
  }
  vg_surr *S = s.get();
  for (DBuf *b : {&S->par, &S->grad, &S->m1, &S->m2})
    if (b->alloc(n * 4) != hipSuccess || hipMemsetAsync(b->p, 0, n * 4, S->stream) != hipSuccess)
      return VG_ERR_HIP;
  if (hipStreamSynchronize(S->stream) != hipSuccess) return VG_ERR_HIP;
  if (S->parbf.alloc(n * 4) != hipSuccess || S->mask.alloc(n) != hipSuccess) return VG_ERR_HIP;
  if (hipMemcpy(S->par.p, p.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  if (hipMemcpy(S->mask.p, mk.data(), n, hipMemcpyHostToDevice) != hipSuccess) return VG_ERR_HIP;
  hipLaunchKernelGGL(k_sg_to_bf16, dim3(nb(n)), dim3(GB), 0, S->stream, S->par.as<float>(), n,
                     S->parbf.as<__bf16>());
  if (S->scl.alloc(sizeof(vg_surr_scaling)) != hipSuccess) return VG_ERR_HIP;
  if (S->ts.alloc(sizeof(TrainState)) != hipSuccess ||
      hipMemsetAsync(S->ts.p, 0, sizeof(TrainState), S->stream) != hipSuccess)
    return VG_ERR_HIP;
  if (hipStreamSynchronize(S->stream) != hipSuccess) return VG_ERR_HIP;
  *out = s.release();
  return VG_OK;
}

void vg_surr_destroy(vg_surr *s) { delete s; }

const char *vg_surr_last_error(const vg_surr *s) { return s ? s->err.c_str() : "null"; }

int vg_surr_reserve(vg_surr *s, int64_t capacity) {
  if (!s || capacity < 1) return VG_ERR_ARG;
  if (capacity <= s->cap) return VG_OK;
  DBuf nx, ny;
  SCHK(s, nx.alloc((size_t)capacity * SG_NIN * 4));
  SCHK(s, ny.alloc((size_t)capacity * SG_NOUT * 4));
  if (s->count > 0) {
    SCHK(s, hipMemcpy(nx.p, s->X.p, (size_t)s->count * SG_NIN * 4, hipMemcpyDeviceToDevice));
    SCHK(s, hipMemcpy(ny.p, s->Y.p, (size_t)s->count * SG_NOUT * 4, hipMemcpyDeviceToDevice));
  }
  std::swap(s->X.p, nx.p);
  std::swap(s->Y.p, ny.p);
  s->cap = capacity;
  return VG_OK;
}

int vg_surr_add_samples(vg_surr *s, const float *x, const float *y, int64_t n) {
  if (!s || !x || !y || n < 0) return VG_ERR_ARG;
  if (s->count + n > s->cap) {
    const int rc = vg_surr_reserve(s, std::max<int64_t>(s->count + n, 2 * s->cap));
    if (rc) return rc;
  }
  SCHK(s, hipMemcpy(s->X.as<float>() + s->count * SG_NIN, x, (size_t)n * SG_NIN * 4,
                    hipMemcpyHostToDevice));
  SCHK(s, hipMemcpy(s->Y.as<float>() + s->count * SG_NOUT, y, (size_t)n * SG_NOUT * 4,
                    hipMemcpyHostToDevice));
  s->count += n;
  return VG_OK;
}

int vg_surr_num_samples(vg_surr *s, int64_t *n) {
  if (!s || !n) return VG_ERR_ARG;
  *n = s->count;
  return VG_OK;
}

int vg_surr_get_samples(vg_surr *s, int64_t first, int64_t n, float *x, float *y) {
  if (!s || first < 0 || n < 0 || first + n > s->count) return VG_ERR_ARG;
  if (x) SCHK(s, hipMemcpy(x, s->X.as<float>() + first * SG_NIN, (size_t)n * SG_NIN * 4, hipMemcpyDeviceToHost));
  if (y) SCHK(s, hipMemcpy(y, s->Y.as<float>() + first * SG_NOUT, (size_t)n * SG_NOUT * 4, hipMemcpyDeviceToHost));
  return VG_OK;
}

// Mean, spread and envelope of the inputs, then the outputs' transform:
// c = a tenth of the mean |y|, mean and spread of asinh(y / c), and the
// envelope of the normalised transformed outputs; over the dataset.
// Envelopes are [min, max] widened by margin * (max - min).
int vg_surr_fit_scaling(vg_surr *s, double margin) {
  if (!s || s->count < 2) return VG_ERR_ARG;
  const int G = 1024;
  DBuf part;
  vg_surr_scaling sc{};
  // Pass 0: inputs; pass 1: raw outputs (for c); pass 2: transformed outputs.
  for (int pass = 0; pass < 3; pass++) {
    const int F = pass ? SG_NOUT : SG_NIN;
    SCHK(s, part.alloc((size_t)G * F * 5 * 8));
    hipLaunchKernelGGL(k_sg_stats, dim3(G), dim3(64), 0, s->stream,
                       pass ? s->Y.as<float>() : s->X.as<float>(), F, s->count,
                       pass == 2 ? (const float *)SCL(s, c_y) : (const float *)nullptr,
                       part.as<double>());
    std::vector<double> h((size_t)G * F * 5);
    SCHK(s, hipMemcpyAsync(h.data(), part.p, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
    SCHK(s, hipStreamSynchronize(s->stream));
    for (int f = 0; f < F; f++) {
      double sum = 0, sq = 0, mn = 1e300, mxv = -1e300, ab = 0;
      for (int g = 0; g < G; g++) {
        const double *o = &h[((size_t)g * F + f) * 5];
        sum += o[0];
        sq += o[1];
        mn = std::min(mn, o[2]);
        mxv = std::max(mxv, o[3]);
        ab += o[4];
      }
      const double mean = sum / s->count;
      const double var = std::max(0.0, sq / s->count - mean * mean);
      const double sd = std::max(std::sqrt(var), 1e-6 + 1e-3 * std::fabs(mean));
      const double w = mxv - mn;
      if (pass == 0) {
        sc.mu_x[f] = (float)mean;
        sc.sd_x[f] = (float)sd;
        sc.lo[f] = (float)(mn - margin * w);
        sc.hi[f] = (float)(mxv + margin * w);
      } else if (pass == 1) {
        sc.c_y[f] = (float)std::max(0.1 * ab / s->count, 1e-30);
      } else {
        sc.mu_y[f] = (float)mean;
        sc.sd_y[f] = (float)sd;
        sc.t_lo[f] = (float)((mn - margin * w - mean) / sd);
        sc.t_hi[f] = (float)((mxv + margin * w - mean) / sd);
      }
    }
    SCHK(s, hipMemcpy(s->scl.p, &sc, sizeof sc, hipMemcpyHostToDevice));
  }
  s->scaled = true;
  return VG_OK;
}

int vg_surr_set_upstroke_fallback(vg_surr *s, double v_min, double h_min) {
  if (!s) return VG_ERR_ARG;
  if (std::isnan(v_min) || std::isnan(h_min)) return s->fail(VG_ERR_ARG, "thresholds must not be NaN");
  s->up_v = (float)v_min;
  s->up_h = (float)h_min;
  return VG_OK;
}

int vg_surr_set_band_fallback(vg_surr *s, double v_lo, double v_hi) {
  if (!s) return VG_ERR_ARG;
  if (std::isnan(v_lo) || std::isnan(v_hi)) return s->fail(VG_ERR_ARG, "band edges must not be NaN");
  s->band_lo = (float)v_lo;
  s->band_hi = (float)v_hi;
  return VG_OK;
}

int vg_surr_get_scaling(vg_surr *s, vg_surr_scaling *out) {
  if (!s || !out || !s->scaled) return VG_ERR_ARG;
  SCHK(s, hipMemcpy(out, s->scl.p, sizeof *out, hipMemcpyDeviceToHost));
  return VG_OK;
}

int vg_surr_set_scaling(vg_surr *s, const vg_surr_scaling *in) {
  if (!s || !in) return VG_ERR_ARG;
  for (int f = 0; f < SG_NOUT; f++)
    if (!(in->c_y[f] > 0.0f) || !(in->sd_y[f] > 0.0f))
      return s->fail(VG_ERR_ARG, "c_y and sd_y must be > 0");
  for (int f = 0; f < SG_NIN; f++)
    if (!(in->sd_x[f] > 0.0f)) return s->fail(VG_ERR_ARG, "sd_x must be > 0");
  SCHK(s, hipMemcpy(s->scl.p, in, sizeof *in, hipMemcpyHostToDevice));
  s->scaled = true;
  return VG_OK;
}

// Parameters in the unpadded layout: per layer W[k][n] row-major, then b[n].
int vg_surr_num_params(int64_t *n) {
  if (!n) return VG_ERR_ARG;
  *n = n_params_unpadded();
  return VG_OK;
}

int vg_surr_get_params(vg_surr *s, float *out) {
  if (!s || !out) return VG_ERR_ARG;
  std::vector<float> p(LAY.total);
  SCHK(s, hipMemcpy(p.data(), s->par.p, p.size() * 4, hipMemcpyDeviceToHost));
  int64_t o = 0;
  for (int l = 0; l < SG_LAYERS; l++) {
    for (int k = 0; k < sg_k(l); k++)
      for (int j = 0; j < sg_n(l); j++) out[o++] = p[LAY.w[l] + (int64_t)k * sg_np(l) + j];
    for (int j = 0; j < sg_n(l); j++) out[o++] = p[LAY.b[l] + j];
  }
  return VG_OK;
}

int vg_surr_set_params(vg_surr *s, const float *in) {
  if (!s || !in) return VG_ERR_ARG;
  std::vector<float> p(LAY.total, 0.0f);
  int64_t o = 0;
  for (int l = 0; l < SG_LAYERS; l++) {
    for (int k = 0; k < sg_k(l); k++)
      for (int j = 0; j < sg_n(l); j++) p[LAY.w[l] + (int64_t)k * sg_np(l) + j] = in[o++];
    for (int j = 0; j < sg_n(l); j++) p[LAY.b[l] + j] = in[o++];
  }
  SCHK(s, hipMemcpy(s->par.p, p.data(), p.size() * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sg_to_bf16, dim3(nb(LAY.total)), dim3(GB), 0, s->stream, s->par.as<float>(),
                     LAY.total, s->parbf.as<__bf16>());
  SCHK(s, hipStreamSynchronize(s->stream));
  s->adam_t = 0;
  SCHK(s, hipMemset(s->ts.p, 0, sizeof(TrainState)));
  SCHK(s, hipMemsetAsync(s->m1.p, 0, LAY.total * 4, s->stream));
  SCHK(s, hipMemsetAsync(s->m2.p, 0, LAY.total * 4, s->stream));
  SCHK(s, hipStreamSynchronize(s->stream));
  return VG_OK;
}

// Mean loss and its gradient (unpadded layout) over n host samples (n a
// multiple of 128), with the network's BF16 arithmetic.
int vg_surr_loss_grad(vg_surr *s, const float *x, const float *y, int n, double *loss, float *grad) {
  if (!s || !x || !y || n < 128 || n % 128 || !s->scaled) return VG_ERR_ARG;
  int rc = ensure_work(s, n);
  if (rc) return rc;
  if ((rc = load_host_batch(s, x, y, n))) return rc;
  double sse = 0.0;
  if ((rc = fwd_bwd(s, n, grad != nullptr, &sse))) return rc;
  if (loss) *loss = sse / ((double)n * SG_NOUT);
  if (grad) {
    std::vector<float> g(LAY.total);
    SCHK(s, hipMemcpy(g.data(), s->grad.p, g.size() * 4, hipMemcpyDeviceToHost));
    int64_t o = 0;
    for (int l = 0; l < SG_LAYERS; l++) {
      for (int k = 0; k < sg_k(l); k++)
        for (int j = 0; j < sg_n(l); j++) grad[o++] = g[LAY.w[l] + (int64_t)k * sg_np(l) + j];
      for (int j = 0; j < sg_n(l); j++) grad[o++] = g[LAY.b[l] + j];
    }
  }
  return VG_OK;
}

// Raw inputs -> raw increments with the training forward (n multiple of 128).
int vg_surr_predict(vg_surr *s, const float *x, int n, float *y) {
  if (!s || !x || !y || n < 128 || n % 128 || !s->scaled) return VG_ERR_ARG;
  int rc = ensure_work(s, n);
  if (rc) return rc;
  if ((rc = load_host_batch(s, x, nullptr, n))) return rc;
  if ((rc = fwd_bwd(s, n, false, nullptr))) return rc;
  // The stream is non-blocking: the copy below does not wait for it.
  SCHK(s, hipStreamSynchronize(s->stream));
  std::vector<float> z((size_t)n * SG_HID);
  SCHK(s, hipMemcpy(z.data(), s->Zs.as<float>() + (size_t)(SG_LAYERS - 1) * n * SG_HID,
                    (size_t)n * SG_NOUTP * 4, hipMemcpyDeviceToHost));
  vg_surr_scaling sc;
  SCHK(s, hipMemcpy(&sc, s->scl.p, sizeof sc, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++)
    for (int f = 0; f < SG_NOUT; f++)
      y[(size_t)i * SG_NOUT + f] =
          sc.c_y[f] * std::sinh(z[(size_t)i * SG_NOUTP + f] * sc.sd_y[f] + sc.mu_y[f]);
  return VG_OK;
}

// Adam over `iters` batches of `batch` samples drawn from the dataset;
// *loss = mean normalised squared error over the last min(iters, 50) batches.
int vg_surr_train(vg_surr *s, int iters, int batch, double lr, double *loss) {
  if (!s || iters < 1 || batch < 128 || batch % 128 || !s->scaled || s->count < 1) return VG_ERR_ARG;
  int rc = ensure_work(s, batch);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f;
  // The step count and learning rate go to device memory; every iteration
  // then launches the same kernels with the same arguments.
  const TrainState ts0{s->adam_t, (float)lr};
  SCHK(s, hipMemcpyAsync(s->ts.p, &ts0, sizeof ts0, hipMemcpyHostToDevice, s->stream));
  auto iteration = [&](double *sse) -> int {
    hipLaunchKernelGGL(k_sg_gather, dim3(nb(batch)), dim3(GB), 0, s->stream, s->X.as<float>(),
                       s->Y.as<float>(), s->count, s->seed, (const int64_t *)s->ts.p,
                       (int64_t)-1, batch, SCL(s, mu_x), SCL(s, sd_x), SCL(s, mu_y),
                       SCL(s, sd_y), SCL(s, c_y), s->X0.as<__bf16>(), s->Yn.as<float>());
    if (const int r = fwd_bwd(s, batch, true, sse)) return r;
    hipLaunchKernelGGL(k_sg_adam, dim3(nb(LAY.total)), dim3(GB), 0, s->stream, s->par.as<float>(),
                       s->grad.as<float>(), s->m1.as<float>(), s->m2.as<float>(),
                       s->mask.as<unsigned char>(), LAY.total, s->ts.as<TrainState>(), b1, b2,
                       1e-8f, s->parbf.as<__bf16>());
    hipLaunchKernelGGL(k_sg_tick, dim3(1), dim3(64), 0, s->stream, s->ts.as<TrainState>());
    return VG_OK;
  };
  // All but the last 50 iterations replay one captured iteration (the host
  // launches one graph instead of ~35 kernels); the last 50 run eagerly and
  // report their loss.  (fwd_bwd swaps the two dZ buffers an even number of
  // times, so a replay finds them where the capture did.)
  const int tail = std::min(iters, 50);
  const int replays = iters - tail;
  if (replays > 0) {
    if (!s->graph || s->graph_batch != batch || s->graph_x != s->X.p ||
        s->graph_count != s->count || s->graph_gen != s->work_gen) {
      if (s->graph) SCHK(s, hipGraphExecDestroy(s->graph));
      s->graph = nullptr;
      SCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      const int r = iteration(nullptr);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s->stream, &g);
      if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      SCHK(s, e);
      const hipError_t ei = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      SCHK(s, ei);
      s->graph_batch = batch;
      s->graph_x = s->X.p;
      s->graph_count = s->count;
      s->graph_gen = s->work_gen;
    }
    for (int it = 0; it < replays; it++) SCHK(s, hipGraphLaunch(s->graph, s->stream));
  }
  double acc = 0.0;
  for (int it = 0; it < tail; it++) {
    double sse = 0.0;
    if ((rc = iteration(&sse))) return rc;
    acc += sse / ((double)batch * SG_NOUT);
  }
  s->adam_t += iters;
  SCHK(s, hipStreamSynchronize(s->stream));
  if (loss) *loss = acc / tail;
  return VG_OK;
}

}  // extern "C"
