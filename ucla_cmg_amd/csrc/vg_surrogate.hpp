// vg_surrogate.hpp -- learned ten Tusscher ionic step (device side shared by
// the training library and the time step).
//
// The network maps one node's ionic-step inputs to the increments of its 19
// states over the step:
//   inputs  (25): V and the 18 other states at the start of the ionic update,
//                 log(dt), and the node's five conductance scalings
//                 (gNa, gKr, gKs, gCaL, gto), normalised by the training data's
//                 per-feature mean and spread, and entered as two BF16 halves
//                 (x = hi + lo, hi = bf16(x), lo = bf16(x - hi): ~16 bits of
//                 each input instead of 8, so the network sees the plateau's
//                 small changes of V), layer 0 reading 50 values;
//   hidden      : 4 layers of 128 with SiLU;
//   outputs (19): the state increments y through the transform
//                 t = (asinh(y / c) - mu) / sd per output, c a small scale
//                 (a tenth of the mean |y|): the loss then weighs relative
//                 error, so the small increments of the plateau and of rest
//                 are fitted as well as the upstroke's large ones.
// Every layer is a BF16 GEMM on MFMA (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation; widths are padded to the MFMA tile (layer 0's 50 input values
// -> 64: his 0-24, los 25-49, zeros; 19 -> 32 outputs; the padding weights
// kept at zero).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/voomgpu.h"

namespace vg {

constexpr int SG_NIN = 25, SG_NOUT = 19, SG_HID = 128, SG_LAYERS = 5;
constexpr int SG_NIN2 = 2 * SG_NIN;  // layer 0's input values: his, then los
constexpr int SG_KIN = 64, SG_NOUTP = 32;
// Padded (K, N) of each layer.
__host__ __device__ constexpr int sg_kp(int l) { return l == 0 ? SG_KIN : SG_HID; }
__host__ __device__ constexpr int sg_np(int l) { return l == SG_LAYERS - 1 ? SG_NOUTP : SG_HID; }
// Unpadded (K, N).
__host__ __device__ constexpr int sg_k(int l) { return l == 0 ? SG_NIN2 : SG_HID; }
__host__ __device__ constexpr int sg_n(int l) { return l == SG_LAYERS - 1 ? SG_NOUT : SG_HID; }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Operand and result layout of v_mfma_f32_32x32x16_bf16 for lane l of a wave:
// A (32 x 16) holds row l%32, k = 8*(l/32) + q (q < 8); B (16 x 32) holds
// column l%32, the same k; the 16 results are column l%32, rows
// 8*(j/4) + 4*(l/32) + j%4.
__device__ inline int mfma_row(int lane, int j) { return 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3); }

// SiLU z * sigmoid(z) with the hardware reciprocal (training and the step
// evaluate the same expression).
__device__ inline float sg_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float sg_silu(float z) { return z * sg_sigmoid(z); }

// The weights as the step's kernel keeps them in LDS: per layer W^T[np][kp]
// (row = output column) with rows padded by 8 (16-byte reads of 32 rows then
// fall on distinct bank groups), layers back to back.
__host__ __device__ constexpr int sg_wt_ld(int l) { return sg_kp(l) + 8; }
__host__ __device__ constexpr int sg_wt_off(int l) {
  return l == 0 ? 0 : sg_wt_off(l - 1) + sg_np(l - 1) * sg_wt_ld(l - 1);
}
constexpr int SG_WT_TOTAL = sg_wt_off(SG_LAYERS);
constexpr int SG_ACT_LD = SG_HID + 8;  // activation rows in LDS

// Everything the step's surrogate kernel needs, all device pointers.
struct SurrDevice {
  const __bf16 *wt;            // [SG_WT_TOTAL] the LDS image above
  const float *b[SG_LAYERS];   // [np], zero padded
  const float *mu_x, *sd_x;    // [SG_NIN]
  const float *lo, *hi;        // [SG_NIN] envelope of the inputs
  const float *mu_y, *sd_y;    // [SG_NOUT] of the transformed outputs
  const float *c_y;            // [SG_NOUT] scale of the asinh transform
  const float *t_lo, *t_hi;    // [SG_NOUT] envelope of the network's outputs
  // Upstroke fallback: rows with V > up_v and h > up_h (depolarised, the
  // sodium channel's inactivation gate still open) go to the plain model;
  // +inf turns it off.
  float up_v, up_h;
  // Voltage-band fallback: rows with band_lo < V < band_hi go to the plain
  // model (an empty band, lo >= hi, turns it off).
  float band_lo, band_hi;
};
// Index of the sodium inactivation gate h among the ten Tusscher states.
constexpr int SG_STATE_H = 8;

// Normalised input (x - mu) / sd, formed as (x - mu) * (1 / sd) with the
// reciprocal correctly rounded: training and the step evaluate the same
// expression, so their BF16 inputs agree bit for bit.
__device__ inline float sg_norm_in(float x, float mu, float inv_sd) { return (x - mu) * inv_sd; }

// A normalised input as layer 0 takes it: hi = bf16(x) and lo = bf16(x - hi)
// (x - hi is exact in FP32).  Training and the step share this.
__device__ inline void sg_split(float x, float &hi, float &lo) {
  hi = (float)(__bf16)x;
  lo = (float)(__bf16)(x - hi);
}
// Layer 0 input value f (0 <= f < SG_KIN) from the normalised inputs xn.
template <class XN>
__device__ inline float sg_in_value(const XN &xn, int f) {
  float hi, lo;
  if (f < SG_NIN) {
    sg_split(xn[f], hi, lo);
    return hi;
  }
  if (f < SG_NIN2) {
    sg_split(xn[f - SG_NIN], hi, lo);
    return lo;
  }
  return 0.0f;
}

// sinh for the output transform: Taylor terms below |u| = 1/16 (truncation
// below 1e-9 relative), else (e^|u| - e^-|u|) / 2 from the hardware
// exponential and reciprocal, the sign restored; ~1e-6 relative against
// sinh, in ~12 instructions where the library's sinhf takes ~70.
__device__ inline float sg_sinh(float u) {
  const float a = fabsf(u), a2 = a * a;
  const float e = __expf(a);
  const float big = 0.5f * (e - __builtin_amdgcn_rcpf(e));
  const float small = a * (1.0f + a2 * (1.0f / 6.0f + a2 * (1.0f / 120.0f)));
  return copysignf(a < 0.0625f ? small : big, u);
}

// Increment -> network target and back.
__device__ inline float sg_to_target(float y, float mu, float sd, float c) {
  return (asinhf(y / c) - mu) / sd;
}
__device__ inline float sg_from_target(float t, float mu, float sd, float c) {
  return c * sg_sinh(t * sd + mu);
}

// Hooks between the surrogate object and the time step (same library).
bool surr_dataset(vg_surr *s, float **x, float **y, int64_t *count, int64_t *cap);
void surr_set_count(vg_surr *s, int64_t count);
bool surr_device(vg_surr *s, SurrDevice *d);  // builds the LDS image

}  // namespace vg
