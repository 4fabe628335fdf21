// Learned boundary extension: the MLP's shape, its parameter layout and the
// exact inference arithmetic shared by the sift kernel (in-kernel inference
// at every sift) and the batch inference kernel.
//
// No counterpart exists in the reference (its GPU path has natural ends only
// and its CPU path mirrors, emd_ref.py:92-95).  The model reads the 16
// extrema nearest to one end of both envelopes and predicts the next 4
// extrema of each envelope beyond that end, which replace the single mirror
// knot of emd_ref.py:92-95.
//
// Inference arithmetic (bit-reproducible on the CPU): BF16 parameters
// widened to FP32; each neuron is an FP32 fma chain over its inputs in index
// order starting from its bias; GELU in tanh form with an exponential built
// from a Cody-Waite reduction and a degree-7 Taylor polynomial (only IEEE
// operations: fma, mul, add, div, rint, ldexp).
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace eemd {

constexpr int MLP_IN = 64;    // 2 envelopes x (16 gaps + 16 values)
constexpr int MLP_HID = 128;  // three hidden layers
constexpr int MLP_OUT = 16;   // 2 envelopes x (4 gaps + 4 values)
constexpr int MLP_CTX = 16;   // extrema per envelope read by the model
constexpr int MLP_PRED = 4;   // virtual extrema per envelope and end

// Parameter layout (row-major W[in][out], then bias[out], per layer)
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_IN * MLP_HID;
constexpr int MLP_W2 = MLP_B1 + MLP_HID;
constexpr int MLP_B2 = MLP_W2 + MLP_HID * MLP_HID;
constexpr int MLP_W3 = MLP_B2 + MLP_HID;
constexpr int MLP_B3 = MLP_W3 + MLP_HID * MLP_HID;
constexpr int MLP_W4 = MLP_B3 + MLP_HID;
constexpr int MLP_B4 = MLP_W4 + MLP_HID * MLP_OUT;
constexpr int MLP_NPARAM = MLP_B4 + MLP_OUT;  // 43408
// LDS image of the model used by the sift kernel: per layer one row per
// neuron, W^T[out][in + MLP_LDS_PAD] (the pad skews consecutive rows across
// LDS banks; rows stay 16-byte aligned), then the biases b1 b2 b3 b4.
constexpr int MLP_LDS_PAD = 8;
constexpr int MLP_LDS_W1 = 0;
constexpr int MLP_LDS_W2 = MLP_LDS_W1 + MLP_HID * (MLP_IN + MLP_LDS_PAD);
constexpr int MLP_LDS_W3 = MLP_LDS_W2 + MLP_HID * (MLP_HID + MLP_LDS_PAD);
constexpr int MLP_LDS_W4 = MLP_LDS_W3 + MLP_HID * (MLP_HID + MLP_LDS_PAD);
constexpr int MLP_LDS_B1 = MLP_LDS_W4 + MLP_OUT * (MLP_HID + MLP_LDS_PAD);
constexpr int MLP_LDS_B2 = MLP_LDS_B1 + MLP_HID;
constexpr int MLP_LDS_B3 = MLP_LDS_B2 + MLP_HID;
constexpr int MLP_LDS_B4 = MLP_LDS_B3 + MLP_HID;
constexpr int MLP_LDS_N = MLP_LDS_B4 + MLP_OUT;
constexpr size_t MLP_LDS_BYTES = (size_t)MLP_LDS_N * 2;
static_assert(MLP_LDS_BYTES % 16 == 0 && ((MLP_IN + MLP_LDS_PAD) * 2) % 16 == 0 &&
                  ((MLP_HID + MLP_LDS_PAD) * 2) % 16 == 0,
              "rows are read in 16-byte pieces");

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// Round to nearest even (finite inputs).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// exp(x) for x in [-87, 88]: x = k ln2 + r with ln2 split in two (Cody and
// Waite), |r| <= ln2/2, exp(r) by its Taylor series to degree 7 (relative
// error < 3e-7), scaled by 2^k exactly.
__device__ __forceinline__ float mlp_exp(float x) {
  x = fminf(fmaxf(x, -87.0f), 88.0f);
  const float k = rintf(x * 1.44269504088896341f);
  float r = fmaf(-k, 0.693145751953125f, x);
  r = fmaf(-k, 1.42860682030941723e-6f, r);
  float p = 1.0f / 5040.0f;
  p = fmaf(p, r, 1.0f / 720.0f);
  p = fmaf(p, r, 1.0f / 120.0f);
  p = fmaf(p, r, 1.0f / 24.0f);
  p = fmaf(p, r, 1.0f / 6.0f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))),
// tanh(u) = 1 - 2 / (1 + exp(2u)) with u clamped to [-15, 15].
__device__ __forceinline__ float mlp_gelu(float x) {
  const float x3 = (x * x) * x;
  const float inner = fmaf(0.044715f, x3, x);
  float u = 0.7978845608028654f * inner;
  u = fminf(fmaxf(u, -15.0f), 15.0f);
  const float e = mlp_exp(u + u);
  const float t = 1.0f - 2.0f / (1.0f + e);
  return (0.5f * x) * (1.0f + t);
}

// One neuron: bias + sum_i w[i*ld] * in[i] as two fma chains in index
// order, lo over i < n/2 starting from the bias and hi over i >= n/2
// starting from 0, then lo + hi.
__device__ __forceinline__ float mlp_neuron(const uint16_t* w, int ld, float bias, const float* in,
                                            int n) {
  float lo = bias, hi = 0.0f;
  for (int i = 0; i < n / 2; ++i) lo = fmaf(bf16_to_f32(w[(int64_t)i * ld]), in[i], lo);
  for (int i = n / 2; i < n; ++i) hi = fmaf(bf16_to_f32(w[(int64_t)i * ld]), in[i], hi);
  return lo + hi;
}

// One half chain of a neuron from its LDS row: acc + sum over the n inputs
// row[0..n) * in[0..n) (n a multiple of 8), fma in index order, 16-byte
// loads of 8 weights and 8 inputs at a time.
__device__ __forceinline__ float mlp_half_row(const uint16_t* row, float acc, const float* in,
                                              int n) {
#pragma unroll 4
  for (int i = 0; i < n; i += 8) {
    const uint4 wq = *reinterpret_cast<const uint4*>(row + i);
    const float4 xa = *reinterpret_cast<const float4*>(in + i);
    const float4 xb = *reinterpret_cast<const float4*>(in + i + 4);
    acc = fmaf(__builtin_bit_cast(float, wq.x << 16), xa.x, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.x & 0xFFFF0000u), xa.y, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.y << 16), xa.z, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.y & 0xFFFF0000u), xa.w, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.z << 16), xb.x, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.z & 0xFFFF0000u), xb.y, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.w << 16), xb.z, acc);
    acc = fmaf(__builtin_bit_cast(float, wq.w & 0xFFFF0000u), xb.w, acc);
  }
  return acc;
}

// mlp_half_row for a compile-time n: all of the row's weights are loaded
// before the fma chain, so the loads are in flight together (the row may be
// in global memory); the arithmetic is mlp_half_row's.
template <int N>
__device__ __forceinline__ float mlp_half_row_n(const uint16_t* row, float acc, const float* in) {
  static_assert(N % 8 == 0, "rows are read in 16-byte pieces");
  uint4 wq[N / 8];
#pragma unroll
  for (int q = 0; q < N / 8; ++q) wq[q] = *reinterpret_cast<const uint4*>(row + 8 * q);
#pragma unroll
  for (int q = 0; q < N / 8; ++q) {
    const float4 xa = *reinterpret_cast<const float4*>(in + 8 * q);
    const float4 xb = *reinterpret_cast<const float4*>(in + 8 * q + 4);
    acc = fmaf(__builtin_bit_cast(float, wq[q].x << 16), xa.x, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].x & 0xFFFF0000u), xa.y, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].y << 16), xa.z, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].y & 0xFFFF0000u), xa.w, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].z << 16), xb.x, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].z & 0xFFFF0000u), xb.y, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].w << 16), xb.z, acc);
    acc = fmaf(__builtin_bit_cast(float, wq[q].w & 0xFFFF0000u), xb.w, acc);
  }
  return acc;
}

}  // namespace eemd
