// Device helpers shared by the ConvNetAutoregressive2D kernels (fused
// forward and fast sampler): MFMA precision traits, activations, register
// D->B fragment chaining, tap tables.
#pragma once
#include "fk_common.h"

namespace fk {

template <int PREC>
struct Prec;

// 16-bit precisions (bf16, fp16) share the layout: 8 elements per lane per
// fragment, 32 input-channel slots per k-step.
template <>
struct Prec<FK_PREC_BF16> {
  using elem = __bf16;
  using pair = bf16x2;
  using frag = bf16x8;
  using quad = bf16x4;
  static constexpr bool H16 = true;
  static constexpr int ESZ = 2;
  static constexpr int KW = 32;
  __device__ static __forceinline__ f32x4 mfma(frag a, frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Prec<FK_PREC_FP16> {
  using elem = _Float16;
  using pair = f16x2;
  using frag = f16x8;
  using quad = f16x4;
  static constexpr bool H16 = true;
  static constexpr int ESZ = 2;
  static constexpr int KW = 32;
  __device__ static __forceinline__ f32x4 mfma(frag a, frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Prec<FK_PREC_FP32> {
  using frag = float;
  static constexpr bool H16 = false;
  static constexpr int ESZ = 4;
  static constexpr int KW = 4;
  __device__ static __forceinline__ f32x4 mfma(frag a, frag b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// 4 fp32 values -> one quad of 16-bit elements and back
template <int PREC>
__device__ __forceinline__ typename Prec<PREC>::quad to_quad(f32x4 v) {
  using E = typename Prec<PREC>::elem;
  typename Prec<PREC>::quad h;
  h[0] = (E)v.x;
  h[1] = (E)v.y;
  h[2] = (E)v.z;
  h[3] = (E)v.w;
  return h;
}
template <int PREC>
__device__ __forceinline__ f32x4 from_quad(typename Prec<PREC>::quad h) {
  f32x4 v;
  v.x = (float)h[0];
  v.y = (float)h[1];
  v.z = (float)h[2];
  v.w = (float)h[3];
  return v;
}

// ReLU on the bit pattern: a float with the sign bit set reads as a negative
// int32, so max(bits, 0) sends every negative value (and -0) to +0 and keeps
// the rest -- one integer max, where fmaxf on an MFMA result also costs a
// NaN-quieting max of the input with itself.
template <int ACT>
__device__ __forceinline__ float act1(float x) {
  if constexpr (ACT == FK_ACT_RELU) return __int_as_float(max(__float_as_int(x), 0));
  else return x > 0.0f ? x : expm1f(x);
}
template <int ACT>
__device__ __forceinline__ f32x4 act4(f32x4 v) {
  f32x4 r;
  r.x = act1<ACT>(v.x);
  r.y = act1<ACT>(v.y);
  r.z = act1<ACT>(v.z);
  r.w = act1<ACT>(v.w);
  return r;
}

// Two fp32 values -> one dword of two 16-bit elements (round to nearest even).
template <int PREC>
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, typename Prec<PREC>::pair));
}
// ReLU of two packed 16-bit floats: the sign is the top bit of bf16 and fp16
// alike, so a packed signed 16-bit max with 0 is the activation.
__device__ __forceinline__ uint32_t relu2(uint32_t w) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, w), (s16x2)0));
}
// act(a), act(b) rounded to 16 bits.  For ReLU the rounding comes first:
// round-to-nearest is monotone and maps 0 to 0, so relu(round(x)) ==
// round(relu(x)) for every x (both give +0 for x <= 0), at half the
// instructions.
template <int PREC, int ACT>
__device__ __forceinline__ uint32_t act_pack2(float a, float b) {
  if constexpr (ACT == FK_ACT_RELU) return relu2(pack2<PREC>(a, b));
  else return pack2<PREC>(act1<ACT>(a), act1<ACT>(b));
}
// act(v) as one quad of 16-bit elements
template <int PREC, int ACT>
__device__ __forceinline__ typename Prec<PREC>::quad act_quad(f32x4 v) {
  const u32x2 w = {act_pack2<PREC, ACT>(v.x, v.y), act_pack2<PREC, ACT>(v.z, v.w)};
  return __builtin_bit_cast(typename Prec<PREC>::quad, w);
}
// act of two D fragments as one 16-bit B fragment (lo -> elements 0..3, hi
// -> 4..7: the order of d_to_b and of a stored pair of quads)
template <int PREC, int ACT>
__device__ __forceinline__ typename Prec<PREC>::frag act_frag(f32x4 lo, f32x4 hi) {
  const u32x4 w = {act_pack2<PREC, ACT>(lo.x, lo.y), act_pack2<PREC, ACT>(lo.z, lo.w),
                   act_pack2<PREC, ACT>(hi.x, hi.y), act_pack2<PREC, ACT>(hi.z, hi.w)};
  return __builtin_bit_cast(typename Prec<PREC>::frag, w);
}

// D fragments (MT tiles of 4 channels per lane) -> B fragments of the next
// 1x1 contraction, in registers.
template <int PREC, int MT, int KS>
__device__ __forceinline__ void d_to_b(const f32x4 (&d)[MT], typename Prec<PREC>::frag (&b)[KS]) {
  if constexpr (Prec<PREC>::H16) {
    using E = typename Prec<PREC>::elem;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      typename Prec<PREC>::frag f;
      f[0] = (E)d[2 * ks].x;
      f[1] = (E)d[2 * ks].y;
      f[2] = (E)d[2 * ks].z;
      f[3] = (E)d[2 * ks].w;
      f[4] = (E)d[2 * ks + 1].x;
      f[5] = (E)d[2 * ks + 1].y;
      f[6] = (E)d[2 * ks + 1].z;
      f[7] = (E)d[2 * ks + 1].w;
      b[ks] = f;
    }
  } else {
#pragma unroll
    for (int t = 0; t < KS; ++t) b[t] = d[t >> 2][t & 3];
  }
}

// Tap tables: KIND 0 = vertical kxk (rows -2..0, cols -1..1), 1 = horizontal
// 1xk (cols -2..0), 2 = h3conv kxk on the concatenation (rows -2..0, cols
// -2..0).  Tap index = di*kw + dj in the Keras kernel.
template <int KIND>
struct Taps {
  static constexpr int N = (KIND == 1) ? 3 : 9;
  __device__ static __forceinline__ int dr(int t) { return KIND == 1 ? 0 : (t / 3) - 2; }
  __device__ static __forceinline__ int dc(int t) {
    return KIND == 0 ? (t % 3) - 1 : (KIND == 1 ? t - 2 : (t % 3) - 2);
  }
  // bit of the (dr, dc) validity mask
  __device__ static __forceinline__ int bit(int t) { return (dr(t) + 2) * 4 + (dc(t) + 2); }
};

__device__ __forceinline__ f32x4 load_bias4(const float* b, int ch, int limit) {
  f32x4 v;
  v.x = (ch + 0 < limit) ? b[ch + 0] : 0.0f;
  v.y = (ch + 1 < limit) ? b[ch + 1] : 0.0f;
  v.z = (ch + 2 < limit) ? b[ch + 2] : 0.0f;
  v.w = (ch + 3 < limit) ? b[ch + 3] : 0.0f;
  return v;
}

enum BlockKind { BK_FIRST = 0, BK_A = 1, BK_B = 2, BK_FINAL = 3 };

__device__ __forceinline__ int block_kind(int blk, int nb) {
  return (blk == 0) ? BK_FIRST : (blk == nb - 1 ? BK_FINAL : ((blk & 1) ? BK_A : BK_B));
}

}  // namespace fk
