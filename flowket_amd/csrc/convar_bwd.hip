// Per-sample log-psi gradients O_k = d log psi(s) / d theta of
// ConvNetAutoregressive2D (the quantity of Machine.predictions_jacobian,
// machines/abstract_machine.py:24-28, re-paired into complex O_k as in
// optimizers/complex_values_optimizer.py:8-9 and computed per example like
// utils/jacobian.py:114-123).
//
// One workgroup per sample, FP32 throughout.  A forward pass stores every
// post-activation tensor of the sample to a per-sample workspace; the backward
// walks the causal blocks in reverse twice -- once with the cotangent of
// Re log psi and once with that of Im log psi -- and writes d/d theta of both
// into O_re and O_im (batch, P) in the canonical flat parameter order.
// Derivatives of the activations are taken from the stored outputs
// (relu' = [y > 0], elu' = y > 0 ? 1 : y + 1).
#include "fk_common.h"
#include "convar_layout.h"

namespace fk {

template <int ACT>
__device__ __forceinline__ float bact(float x) {
  if constexpr (ACT == FK_ACT_RELU) return fmaxf(x, 0.0f);
  else return x > 0.0f ? x : expm1f(x);
}
template <int ACT>
__device__ __forceinline__ float bdact(float y) {  // derivative from the output
  if constexpr (ACT == FK_ACT_RELU) return y > 0.0f ? 1.0f : 0.0f;
  else return y > 0.0f ? 1.0f : y + 1.0f;
}

struct BwdArgs {
  const float* params;  // flat parameters (effective kernels), P floats
  int H, W, HW, C, nb;
  int64_t batch;
  const int8_t* spins;  // (batch, HW)
  float* work;          // per-workgroup activation store
  int64_t work_floats;  // floats per workgroup
  float* o_re;          // (batch, P)
  float* o_im;
  float* logpsi;        // (batch, 2) or null
  int64_t P;
  int64_t ld;           // row stride of o_re / o_im (>= P)
  const float* cot;     // GRAD_VJP: (batch, 2) cotangents of (Re, Im) log psi
  float* partial;       // GRAD_VJP: (gridDim.x, P) per-workgroup partial gradients
  const float* tangent; // JVP: direction, P floats
  float* jvp;           // JVP: (batch, 2) output
  const float* kt;      // the conv kernels transposed (transpose_kernels_kernel), P floats
  const float* tkt;     // JVP: the tangent's conv kernels transposed
  int stage;            // LDS row stride (floats) of the two staging tiles, 0: no staging
};

// Offsets of layer parameters in the flat vector.
struct ParamOffsets {
  int64_t kv, bv, kh, bh, kx, bx, ky, by, kc, bc;
};
__device__ __forceinline__ ParamOffsets param_offsets(int blk, int C) {
  // block 0 has cin = 1; blocks >= 1 have cin = C
  const int64_t blk0 = 9LL * C + C + 3LL * C + C + (int64_t)C * (C / 2) * 2 + C + 9LL * C * C + C;
  const int64_t blkn = 9LL * C * C + C + 3LL * C * C + C + (int64_t)C * (C / 2) * 2 + C + 9LL * C * C + C;
  const int64_t base = blk == 0 ? 0 : blk0 + (int64_t)(blk - 1) * blkn;
  const int cin = blk == 0 ? 1 : C;
  ParamOffsets o;
  o.kv = base;
  o.bv = o.kv + 9LL * cin * C;
  o.kh = o.bv + C;
  o.bh = o.kh + 3LL * cin * C;
  o.kx = o.bh + C;
  o.bx = o.kx + (int64_t)C * (C / 2);
  o.ky = o.bx + C / 2;
  o.by = o.ky + (int64_t)C * (C / 2);
  o.kc = o.by + C / 2;
  o.bc = o.kc + 9LL * C * C;
  return o;
}
__device__ __forceinline__ int64_t head_offset(int nb, int C) {
  const ParamOffsets o = param_offsets(nb - 1, C);
  return o.bc + C;
}

// Tap offsets: vertical kxk rows -2..0 cols -1..1; horizontal 1xk cols -2..0;
// h3 kxk rows -2..0 cols -2..0.
__device__ __forceinline__ void tap_rc(int kind, int t, int& dr, int& dc) {
  if (kind == 0) {
    dr = t / 3 - 2;
    dc = t % 3 - 1;
  } else if (kind == 1) {
    dr = 0;
    dc = t - 2;
  } else {
    dr = t / 3 - 2;
    dc = t % 3 - 2;
  }
}

// out[p][co] = b[co] + sum_{tap,ci} K[tap][ci][co] * in[p + off][ci]
// (cin channels; in/out are [HW][cs] with channel strides cs_in/cs_out).
__device__ void conv_fwd(float* out, int cs_out, const float* in, int cs_in, int cin, int cout, const float* K,
                         const float* b, int kind, int ntap, int H, int W, bool accumulate = false) {
  const int HW = H * W;
  for (int idx = threadIdx.x; idx < HW * cout; idx += blockDim.x) {
    const int p = idx / cout, co = idx - (idx / cout) * cout;
    const int i = p / W, j = p - (p / W) * W;
    float acc = b ? b[co] : 0.0f;
    for (int t = 0; t < ntap; ++t) {
      int dr, dc;
      tap_rc(kind, t, dr, dc);
      const int r = i + dr, c = j + dc;
      if (r < 0 || r >= H || c < 0 || c >= W) continue;
      const float* src = in + (r * W + c) * cs_in;
      const float* k = K + (int64_t)t * cin * cout + co;
      for (int ci = 0; ci < cin; ++ci) acc += k[(int64_t)ci * cout] * src[ci];
    }
    if (accumulate) out[p * cs_out + co] += acc;
    else out[p * cs_out + co] = acc;
  }
}

// din[q][ci] (+)= sum_{tap,co} K[tap][ci][co] * dout[q - off][co]
__device__ void conv_bwd_data(float* din, int cs_in, bool accumulate, const float* dout, int cs_out, int cin,
                              int cout, const float* K, int kind, int ntap, int H, int W) {
  const int HW = H * W;
  for (int idx = threadIdx.x; idx < HW * cin; idx += blockDim.x) {
    const int q = idx / cin, ci = idx - (idx / cin) * cin;
    const int i = q / W, j = q - (q / W) * W;
    float acc = 0.0f;
    for (int t = 0; t < ntap; ++t) {
      int dr, dc;
      tap_rc(kind, t, dr, dc);
      const int r = i - dr, c = j - dc;
      if (r < 0 || r >= H || c < 0 || c >= W) continue;
      const float* g = dout + (r * W + c) * cs_out;
      const float* k = K + ((int64_t)t * cin + ci) * cout;
      for (int co = 0; co < cout; ++co) acc += k[co] * g[co];
    }
    if (accumulate) din[q * cs_in + ci] += acc;
    else din[q * cs_in + ci] = acc;
  }
}

// dK[tap][ci][co] = sum_p in[p + off][ci] * dout[p][co]; db[co] = sum_p dout[p][co]
__device__ void conv_bwd_weight(float* dK, float* db, const float* in, int cs_in, const float* dout, int cs_out,
                                int cin, int cout, int kind, int ntap, int H, int W, bool accumulate = false) {
  const int HW = H * W;
  const int nk = ntap * cin * cout;
  for (int idx = threadIdx.x; idx < nk + cout; idx += blockDim.x) {
    if (idx >= nk) {
      const int co = idx - nk;
      float acc = 0.0f;
      for (int p = 0; p < HW; ++p) acc += dout[p * cs_out + co];
      if (accumulate) db[co] += acc;
      else db[co] = acc;
      continue;
    }
    const int t = idx / (cin * cout);
    const int rem = idx - t * cin * cout;
    const int ci = rem / cout, co = rem - (rem / cout) * cout;
    int dr, dc;
    tap_rc(kind, t, dr, dc);
    float acc = 0.0f;
    for (int i = 0; i < H; ++i) {
      const int r = i + dr;
      if (r < 0 || r >= H) continue;
      for (int j = 0; j < W; ++j) {
        const int c = j + dc;
        if (c < 0 || c >= W) continue;
        acc += in[(r * W + c) * cs_in + ci] * dout[(i * W + j) * cs_out + co];
      }
    }
    if (accumulate) dK[idx] += acc;
    else dK[idx] = acc;
  }
}

// Per-sample workspace layout (floats): spins HW, then per block
// VA (act(v')), XA (act(hx)), CAT (x2|y2), VO (vertical output), HO
// (horizontal output), each HW*C.
// (the spins region is padded to 4 floats so that every tensor is 16-byte
// aligned)
__device__ __forceinline__ float* ws_tensor(float* ws, int HW, int C, int blk, int which) {
  return ws + ((HW + 3) & ~3) + ((int64_t)blk * 5 + which) * HW * C;
}
enum { T_VA = 0, T_XA = 1, T_CAT = 2, T_VO = 3, T_HO = 4 };

template <int ACT>
__global__ __launch_bounds__(256) void convar_grad_kernel(BwdArgs a) {
  extern __shared__ float sm[];
  const int H = a.H, W = a.W, HW = a.HW, C = a.C, nb = a.nb;
  const int F = nb - 1;
  const int64_t b = blockIdx.x;
  if (b >= a.batch) return;
  float* ws = a.work + (int64_t)blockIdx.x * a.work_floats;
  // LDS: gradient buffers
  float* dV = sm;                // HW*C  gradient w.r.t. the vertical input of the current block
  float* dH = dV + HW * C;       // HW*C
  float* dVr = dH + HW * C;      // HW*C  pending residual gradient (vertical)
  float* dHr = dVr + HW * C;     // HW*C
  float* t0 = dHr + HW * C;      // HW*C  scratch
  float* t1 = t0 + HW * C;       // HW*C  scratch
  float* dz = t1 + HW * C;       // HW*4
  const float* prm = a.params;

  // ------------------------------------------------------------ forward
  float* sp = ws;
  for (int q = threadIdx.x; q < HW; q += blockDim.x) sp[q] = (float)a.spins[b * HW + q];
  __syncthreads();
  for (int blk = 0; blk < nb; ++blk) {
    const ParamOffsets o = param_offsets(blk, C);
    const bool first = blk == 0, fin = blk == F;
    const bool kb = !first && !fin && ((blk & 1) == 0);  // second block of a residual pair
    const int cin = first ? 1 : C;
    const float* Vin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_VO);
    const float* Hin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_HO);
    const int cs = first ? 1 : C;
    float* VA = ws_tensor(ws, HW, C, blk, T_VA);
    float* XA = ws_tensor(ws, HW, C, blk, T_XA);
    float* CAT = ws_tensor(ws, HW, C, blk, T_CAT);
    float* VO = ws_tensor(ws, HW, C, blk, T_VO);
    float* HO = ws_tensor(ws, HW, C, blk, T_HO);
    // v' -> t0, hx -> t1
    conv_fwd(t0, C, Vin, cs, cin, C, prm + o.kv, prm + o.bv, 0, 9, H, W);
    conv_fwd(t1, C, Hin, cs, cin, C, prm + o.kh, prm + o.bh, 1, 3, H, W);
    __syncthreads();
    for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
      const float v = t0[idx];
      VA[idx] = bact<ACT>(v);
      XA[idx] = bact<ACT>(t1[idx]);
      if (!fin) {
        float vo;
        if (kb) vo = bact<ACT>(ws_tensor(ws, HW, C, blk - 2, T_VO)[idx] + v);
        else vo = bact<ACT>(v);
        VO[idx] = vo;
      }
    }
    __syncthreads();
    // x2 = act(1x1(shift?(xa))), y2 = act(1x1(downshift(va)))
    for (int idx = threadIdx.x; idx < HW * (C / 2); idx += blockDim.x) {
      const int p = idx / (C / 2), co = idx - (idx / (C / 2)) * (C / 2);
      const int i = p / W, j = p - (p / W) * W;
      float x = prm[o.bx + co];
      if (!fin || j > 0) {
        const float* src = XA + (fin ? p - 1 : p) * C;
        for (int ci = 0; ci < C; ++ci) x += prm[o.kx + (int64_t)ci * (C / 2) + co] * src[ci];
      }
      float y = prm[o.by + co];
      if (i > 0) {
        const float* src = VA + (p - W) * C;
        for (int ci = 0; ci < C; ++ci) y += prm[o.ky + (int64_t)ci * (C / 2) + co] * src[ci];
      }
      CAT[p * C + co] = bact<ACT>(x);
      CAT[p * C + C / 2 + co] = bact<ACT>(y);
    }
    __syncthreads();
    conv_fwd(t0, C, CAT, C, C, C, prm + o.kc, prm + o.bc, 2, 9, H, W);
    __syncthreads();
    for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
      const float h = t0[idx];
      float ho;
      if (kb) ho = bact<ACT>(ws_tensor(ws, HW, C, blk - 2, T_HO)[idx] + h);
      else ho = bact<ACT>(h);
      HO[idx] = ho;
    }
    __syncthreads();
  }

  // head: z = Khead act(h_F) + bhead; cotangents of Re / Im log psi
  const int64_t ho = head_offset(nb, C);
  const float* HF = ws_tensor(ws, HW, C, F, T_HO);
  __shared__ float s_lp[2];
  if (threadIdx.x == 0) s_lp[0] = s_lp[1] = 0.0f;
  __syncthreads();
  for (int part = 0; part < 2; ++part) {
    // per-pixel cotangent of the 4 logits
    for (int p = threadIdx.x; p < HW; p += blockDim.x) {
      float z[4];
      for (int k = 0; k < 4; ++k) {
        float acc = prm[ho + 4 * C + k];
        for (int ci = 0; ci < C; ++ci) acc += prm[ho + ci * 4 + k] * HF[p * C + ci];
        z[k] = acc;
      }
      const float mx = fmaxf(z[0], z[1]);
      const float e0 = expf(2.0f * (z[0] - mx)), e1 = expf(2.0f * (z[1] - mx));
      const float p0 = e0 / (e0 + e1), p1 = e1 / (e0 + e1);
      const bool up = sp[p] > 0.0f;
      if (part == 0) {
        dz[p * 4 + 0] = (up ? 1.0f : 0.0f) - p0;
        dz[p * 4 + 1] = (up ? 0.0f : 1.0f) - p1;
        dz[p * 4 + 2] = 0.0f;
        dz[p * 4 + 3] = 0.0f;
        const float lse = mx + 0.5f * logf(e0 + e1);
        atomicAdd(&s_lp[0], (up ? z[0] : z[1]) - lse);
        atomicAdd(&s_lp[1], up ? z[2] : z[3]);
      } else {
        dz[p * 4 + 0] = 0.0f;
        dz[p * 4 + 1] = 0.0f;
        dz[p * 4 + 2] = up ? 1.0f : 0.0f;
        dz[p * 4 + 3] = up ? 0.0f : 1.0f;
      }
    }
    __syncthreads();
    float* O = (part == 0 ? a.o_re : a.o_im) + b * a.ld;
    // head weight / bias gradients and d act(h_F)
    for (int idx = threadIdx.x; idx < 4 * C + 4; idx += blockDim.x) {
      float acc = 0.0f;
      if (idx < 4 * C) {
        const int ci = idx / 4, k = idx - (idx / 4) * 4;
        for (int p = 0; p < HW; ++p) acc += HF[p * C + ci] * dz[p * 4 + k];
      } else {
        for (int p = 0; p < HW; ++p) acc += dz[p * 4 + idx - 4 * C];
      }
      O[ho + idx] = acc;
    }
    for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
      const int p = idx / C, ci = idx - (idx / C) * C;
      float g = 0.0f;
      for (int k = 0; k < 4; ++k) g += prm[ho + ci * 4 + k] * dz[p * 4 + k];
      dH[idx] = g;   // gradient w.r.t. HO of the final block
      dV[idx] = 0.0f;  // the final block's vertical output is discarded
      dVr[idx] = 0.0f;
      dHr[idx] = 0.0f;
    }
    __syncthreads();

    // ---------------------------------------------------------- backward
    for (int blk = F; blk >= 0; --blk) {
      const ParamOffsets o = param_offsets(blk, C);
      const bool first = blk == 0, fin = blk == F;
      const bool kb = !first && !fin && ((blk & 1) == 0);
      const bool ka = !first && !fin && ((blk & 1) == 1);
      const int cin = first ? 1 : C;
      const float* Vin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_VO);
      const float* Hin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_HO);
      const int cs = first ? 1 : C;
      const float* VA = ws_tensor(ws, HW, C, blk, T_VA);
      const float* XA = ws_tensor(ws, HW, C, blk, T_XA);
      const float* CAT = ws_tensor(ws, HW, C, blk, T_CAT);
      const float* VO = ws_tensor(ws, HW, C, blk, T_VO);
      const float* HO = ws_tensor(ws, HW, C, blk, T_HO);
      // dh' -> t1 ; dv' -> t0 (from the vertical output; 0 for the final block)
      for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
        const float gh = dH[idx] * bdact<ACT>(HO[idx]);
        t1[idx] = gh;
        float gv = 0.0f;
        if (!fin) gv = dV[idx] * bdact<ACT>(VO[idx]);
        t0[idx] = gv;
        if (kb) {  // the residual inputs receive the pre-activation gradient
          dVr[idx] = gv;
          dHr[idx] = gh;
        }
      }
      __syncthreads();
      // h3conv: weights, bias, and d cat -> dV (reused as scratch)
      conv_bwd_weight(O + o.kc, O + o.bc, CAT, C, t1, C, C, C, 2, 9, H, W);
      conv_bwd_data(dV, C, false, t1, C, C, C, prm + o.kc, 2, 9, H, W);
      __syncthreads();
      // d x2p (in dV[:, :C/2]) and d y2p (dV[:, C/2:]) through act
      for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) dV[idx] *= bdact<ACT>(CAT[idx]);
      __syncthreads();
      // x1x1 / y1x1 weight and bias gradients
      for (int idx = threadIdx.x; idx < 2 * (C * (C / 2) + C / 2); idx += blockDim.x) {
        const int which = idx / (C * (C / 2) + C / 2);  // 0: x, 1: y
        const int r = idx - which * (C * (C / 2) + C / 2);
        float acc = 0.0f;
        if (r < C * (C / 2)) {
          const int ci = r / (C / 2), co = r - (r / (C / 2)) * (C / 2);
          for (int p = 0; p < HW; ++p) {
            const int i = p / W, j = p - (p / W) * W;
            float in;
            if (which == 0) in = (fin ? (j > 0 ? XA[(p - 1) * C + ci] : 0.0f) : XA[p * C + ci]);
            else in = (i > 0 ? VA[(p - W) * C + ci] : 0.0f);
            acc += in * dV[p * C + which * (C / 2) + co];
          }
          O[(which == 0 ? o.kx : o.ky) + r] = acc;
        } else {
          const int co = r - C * (C / 2);
          for (int p = 0; p < HW; ++p) acc += dV[p * C + which * (C / 2) + co];
          O[(which == 0 ? o.bx : o.by) + co] = acc;
        }
      }
      // d xa (with the RightShift adjoint) -> dH, d va (DownShift adjoint)
      // added into dv' (t0)
      for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
        const int p = idx / C, ci = idx - (idx / C) * C;
        const int i = p / W, j = p - (p / W) * W;
        float gx = 0.0f;
        const int px = fin ? p + 1 : p;
        if (!fin || j + 1 < W)
          for (int co = 0; co < C / 2; ++co) gx += prm[o.kx + (int64_t)ci * (C / 2) + co] * dV[px * C + co];
        float gy = 0.0f;
        if (i + 1 < H)
          for (int co = 0; co < C / 2; ++co)
            gy += prm[o.ky + (int64_t)ci * (C / 2) + co] * dV[(p + W) * C + C / 2 + co];
        dH[idx] = gx * bdact<ACT>(XA[idx]);   // d hx
        t0[idx] += gy * bdact<ACT>(VA[idx]);  // d v'
      }
      __syncthreads();
      // horizontal 1xk and vertical kxk: weights, biases
      conv_bwd_weight(O + o.kh, O + o.bh, Hin, cs, dH, C, cin, C, 1, 3, H, W);
      conv_bwd_weight(O + o.kv, O + o.bv, Vin, cs, t0, C, cin, C, 0, 9, H, W);
      if (first) break;
      // gradients w.r.t. the block inputs (= the previous block's outputs)
      conv_bwd_data(t1, C, false, dH, C, C, C, prm + o.kh, 1, 3, H, W);
      __syncthreads();
      conv_bwd_data(dV, C, false, t0, C, C, C, prm + o.kv, 0, 9, H, W);
      __syncthreads();
      for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
        float gh = t1[idx], gv = dV[idx];
        if (ka) {  // first block of a residual pair: add the residual path
          gh += dHr[idx];
          gv += dVr[idx];
        }
        dH[idx] = gh;
        dV[idx] = gv;
      }
      __syncthreads();
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && a.logpsi) {
    a.logpsi[2 * b] = s_lp[0];
    a.logpsi[2 * b + 1] = s_lp[1];
  }
}

// ===========================================================================
// MFMA form (channels a multiple of 32): the same forward / backward
// schedule, with every C-channel convolution, its data adjoint and its
// per-sample weight gradient an implicit GEMM on v_mfma_f32_16x16x4f32
// (FP32 in, FP32 accumulate).  Operands stream from the per-sample workspace
// (L1/L2 resident); tap validity comes from per-lattice source tables in LDS.
// MFMA 16x16x4 lane roles: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane j + 16*(i/4), register i%4.

// source-table rows: vertical taps 0-8, horizontal 9-11, h3 12-20, x1x1 21,
// x1x1 behind the RightShift (final block) 22, y1x1 behind the DownShift 23
constexpr int GT_V = 0, GT_H = 9, GT_C = 12, GT_X = 21, GT_XF = 22, GT_Y = 23, GT_N = 24;

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Each wave walks 2x2 blocks of 16x16 D tiles; the K loop is (outer, inner)
// so that tap and channel-quad indices need no division.
template <class FA, class FB, class FS>
__device__ __forceinline__ void wave_gemm(int mtiles, int ntiles, int kouter, int kinner, FA fa, FB fb, FS fs) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int mb = (mtiles + 1) >> 1, nbk = (ntiles + 1) >> 1;
  for (int blk = wave; blk < mb * nbk; blk += nw) {
    const int mt0 = 2 * (blk / nbk), nt0 = 2 * (blk % nbk);
    const bool m1 = mt0 + 1 < mtiles, n1 = nt0 + 1 < ntiles;
    f32x4 c00 = {0.f, 0.f, 0.f, 0.f}, c01 = c00, c10 = c00, c11 = c00;
    for (int t = 0; t < kouter; ++t) {
#pragma unroll 4
      for (int k = 0; k < kinner; ++k) {
        const float a0 = fa(mt0, t, k), a1 = m1 ? fa(mt0 + 1, t, k) : 0.0f;
        const float b0 = fb(nt0, t, k), b1 = n1 ? fb(nt0 + 1, t, k) : 0.0f;
        c00 = mfma4(a0, b0, c00);
        c01 = mfma4(a0, b1, c01);
        c10 = mfma4(a1, b0, c10);
        c11 = mfma4(a1, b1, c11);
      }
    }
    fs(mt0, nt0, c00);
    if (n1) fs(mt0, nt0 + 1, c01);
    if (m1) fs(mt0 + 1, nt0, c10);
    if (m1 && n1) fs(mt0 + 1, nt0 + 1, c11);
  }
}

// The same tiling over the contraction in quads: quad j of lane group kk
// holds the contraction indices 16 j + 4 kk + r, r = 0..3, which feed four
// consecutive MFMAs.  Operands contiguous along the contraction then come
// in one 16-byte load per lane per four MFMAs (instead of four 4-byte
// loads touching a cache line each).
template <class FA, class FB, class FS>
__device__ __forceinline__ void wave_gemm4(int mtiles, int ntiles, int kouter, int kquads, FA fa, FB fb, FS fs) {
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int mb = (mtiles + 1) >> 1, nbk = (ntiles + 1) >> 1;
  for (int blk = wave; blk < mb * nbk; blk += nw) {
    const int mt0 = 2 * (blk / nbk), nt0 = 2 * (blk % nbk);
    const bool m1 = mt0 + 1 < mtiles, n1 = nt0 + 1 < ntiles;
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    f32x4 c00 = z, c01 = z, c10 = z, c11 = z;
    for (int t = 0; t < kouter; ++t) {
#pragma unroll 2
      for (int j = 0; j < kquads; ++j) {
        const f32x4 a0 = fa(mt0, t, j), a1 = m1 ? fa(mt0 + 1, t, j) : z;
        const f32x4 b0 = fb(nt0, t, j), b1 = n1 ? fb(nt0 + 1, t, j) : z;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          c00 = mfma4(a0[r], b0[r], c00);
          c01 = mfma4(a0[r], b1[r], c01);
          c10 = mfma4(a1[r], b0[r], c10);
          c11 = mfma4(a1[r], b1[r], c11);
        }
      }
    }
    fs(mt0, nt0, c00);
    if (n1) fs(mt0, nt0 + 1, c01);
    if (m1) fs(mt0 + 1, nt0, c10);
    if (m1 && n1) fs(mt0 + 1, nt0 + 1, c11);
  }
}

// Conv kernels with the contraction index innermost: KT[tap][co][ci] =
// K[tap][ci][co] for every kernel with C input channels, at the
// kernel's own offset in the flat vector (other entries untouched).  One
// grid row per (block, kernel), threads over the kernel's entries.
__global__ void transpose_kernels_kernel(const float* __restrict__ prm, float* __restrict__ kt, int C, int nb) {
  const int row = blockIdx.y, blk = row / 5, which = row % 5;
  // block 0's vertical and horizontal kernels read the 1-channel input
  // (scalar path, untransposed)
  if (blk >= nb || (blk == 0 && which < 2)) return;
  const ParamOffsets o = param_offsets(blk, C);
  int64_t off;
  int ntap, cout;
  switch (which) {
    case 0: off = o.kv; ntap = 9; cout = C; break;
    case 1: off = o.kh; ntap = 3; cout = C; break;
    case 2: off = o.kx; ntap = 1; cout = C / 2; break;
    case 3: off = o.ky; ntap = 1; cout = C / 2; break;
    default: off = o.kc; ntap = 9; cout = C; break;
  }
  const int cin = C;
  const int64_t n = (int64_t)ntap * cin * cout;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(e / ((int64_t)cin * cout));
    const int rem = (int)(e - (int64_t)t * cin * cout);
    const int ci = rem / cout, co = rem - (rem / cout) * cout;
    kt[off + ((int64_t)t * cout + co) * cin + ci] = prm[off + e];
  }
}

// Elementwise passes over HW x C tensors (C a multiple of 4) in 16-byte
// vectors: element i of a tensor as a quad, the activation and its
// derivative per component.
__device__ __forceinline__ f32x4 L4(const float* p, int i) { return *reinterpret_cast<const f32x4*>(p + 4 * i); }
__device__ __forceinline__ f32x4& S4(float* p, int i) { return *reinterpret_cast<f32x4*>(p + 4 * i); }
template <int ACT>
__device__ __forceinline__ f32x4 bact4(f32x4 x) {
  return f32x4{bact<ACT>(x[0]), bact<ACT>(x[1]), bact<ACT>(x[2]), bact<ACT>(x[3])};
}
template <int ACT>
__device__ __forceinline__ f32x4 bdact4(f32x4 y) {
  return f32x4{bdact<ACT>(y[0]), bdact<ACT>(y[1]), bdact<ACT>(y[2]), bdact<ACT>(y[3])};
}

// Copies n channels (from channel off, row stride cs) of an HW-pixel tensor
// into an LDS tile of row stride SC (C + 4 floats: consecutive pixels start
// 36 banks apart, so 16 lanes reading 16 bytes of 16 pixels do not
// conflict).  The 3x3 convolutions, their data adjoints and weight gradients
// then read their tap-shifted operands from LDS instead of L2.
__device__ __forceinline__ void stage_in(float* dst, int SC, const float* src, int cs, int off, int n, int HW) {
  const int q4 = n / 4;
  for (int idx = threadIdx.x; idx < HW * q4; idx += blockDim.x) {
    const int p = idx / q4, c = idx - p * q4;
    *reinterpret_cast<f32x4*>(dst + p * SC + 4 * c) =
        *reinterpret_cast<const f32x4*>(src + (int64_t)p * cs + off + 4 * c);
  }
}

// out[p][co_off + co] = bias[co] + sum_{tap, ci} K[tap][ci][co] in[src(tap, p)][ci],
// the kernel given transposed (KT[tap][co][ci], transpose_kernels_kernel)
// The same contraction with the store left to post(p, co_off + co, r), r =
// bias + the sums of channels co..co+3 of pixel p: the elementwise work that
// follows a convolution (activation, residual) runs in its epilogue.
template <class Post>
__device__ __forceinline__ void mconv_fwd_post(int co_off, const float* in, int cs_in, int cin, int cout,
                                               const float* KT, const float* bias, const int16_t* tab, int ntap,
                                               int HW, Post post) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, kk = lane >> 4;
  const f32x4 z = {0.f, 0.f, 0.f, 0.f};
  wave_gemm4(
      cout / 16, (HW + 15) / 16, ntap, cin / 16,
      [&](int mt, int t, int j) {
        return *reinterpret_cast<const f32x4*>(KT + ((int64_t)t * cout + 16 * mt + l16) * cin + 16 * j + 4 * kk);
      },
      [&](int nt, int t, int j) {
        const int p = 16 * nt + l16;
        const int src = p < HW ? tab[t * HW + p] : -1;
        return src >= 0 ? *reinterpret_cast<const f32x4*>(in + (int64_t)src * cs_in + 16 * j + 4 * kk) : z;
      },
      [&](int mt, int nt, f32x4 c) {
        const int p = 16 * nt + l16;
        if (p >= HW) return;
        const int co = 16 * mt + 4 * kk;
        f32x4 r = c;
        if (bias) r += *reinterpret_cast<const f32x4*>(bias + co);
        post(p, co_off + co, r);
      });
}

__device__ void mconv_fwd(float* out, int cs_out, int co_off, const float* in, int cs_in, int cin, int cout,
                          const float* KT, const float* bias, const int16_t* tab, int ntap, int HW,
                          bool accumulate = false) {
  mconv_fwd_post(co_off, in, cs_in, cin, cout, KT, bias, tab, ntap, HW, [&](int p, int co, f32x4 r) {
    f32x4* dst = reinterpret_cast<f32x4*>(out + (int64_t)p * cs_out + co);
    if (accumulate) r += *dst;
    *dst = r;
  });
}

// din[q][ci] = sum_{tap, co} K[tap][ci][co] dout[inv(tap, q)][co_off + co]
__device__ void mconv_bwd_data(float* din, int cs_in, const float* dout, int cs_out, int co_off, int cin, int cout,
                               const float* K, const int16_t* inv, int ntap, int HW) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, kk = lane >> 4;
  const f32x4 z = {0.f, 0.f, 0.f, 0.f};
  wave_gemm4(
      cin / 16, (HW + 15) / 16, ntap, cout / 16,
      [&](int mt, int t, int j) {
        return *reinterpret_cast<const f32x4*>(K + (int64_t)(t * cin + 16 * mt + l16) * cout + 16 * j + 4 * kk);
      },
      [&](int nt, int t, int j) {
        const int q = 16 * nt + l16;
        const int src = q < HW ? inv[t * HW + q] : -1;
        return src >= 0 ? *reinterpret_cast<const f32x4*>(dout + (int64_t)src * cs_out + co_off + 16 * j + 4 * kk)
                        : z;
      },
      [&](int mt, int nt, f32x4 c) {
        const int q = 16 * nt + l16;
        if (q >= HW) return;
        *reinterpret_cast<f32x4*>(din + (int64_t)q * cs_in + 16 * mt + 4 * kk) = c;
      });
}

// dK[tap][ci][co] = sum_p in[src(tap, p)][ci] dout[p][co_off + co]  (D[co][ci]),
// every tap's 2x2 tile blocks shared out over the waves; db[co] = sum_p dout.
__device__ void mconv_bwd_weight(float* dK, float* db, const float* in, int cs_in, const float* dout, int cs_out,
                                 int co_off, int cin, int cout, const int16_t* tab, int ntap, int HW,
                                 bool accumulate = false) {
  const int lane = threadIdx.x & 63, l16 = lane & 15, kk = lane >> 4;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int mtiles = cout / 16, ntiles = cin / 16;
  const int mb = (mtiles + 1) >> 1, nbk = (ntiles + 1) >> 1;
  const int ks = (HW + 3) / 4;
  for (int blk = wave; blk < ntap * mb * nbk; blk += nw) {
    const int t = blk / (mb * nbk), r = blk - t * (mb * nbk);
    const int mt0 = 2 * (r / nbk), nt0 = 2 * (r % nbk);
    const bool m1 = mt0 + 1 < mtiles, n1 = nt0 + 1 < ntiles;
    const int16_t* tt = tab + t * HW;
    f32x4 c00 = {0.f, 0.f, 0.f, 0.f}, c01 = c00, c10 = c00, c11 = c00;
#pragma unroll 4
    for (int k = 0; k < ks; ++k) {
      const int p = 4 * k + kk;
      const bool pin = p < HW;
      const float* dp = dout + (int64_t)p * cs_out + co_off + l16;
      const float a0 = pin ? dp[16 * mt0] : 0.0f;
      const float a1 = (pin && m1) ? dp[16 * mt0 + 16] : 0.0f;
      const int src = pin ? tt[p] : -1;
      const float* ip = in + (int64_t)src * cs_in + l16;
      const float b0 = src >= 0 ? ip[16 * nt0] : 0.0f;
      const float b1 = (src >= 0 && n1) ? ip[16 * nt0 + 16] : 0.0f;
      c00 = mfma4(a0, b0, c00);
      c01 = mfma4(a0, b1, c01);
      c10 = mfma4(a1, b0, c10);
      c11 = mfma4(a1, b1, c11);
    }
    auto st = [&](int mt, int nt, f32x4 c) {
      const int ci = 16 * nt + l16, co = 16 * mt + 4 * kk;
      f32x4* dst = reinterpret_cast<f32x4*>(dK + ((int64_t)t * cin + ci) * cout + co);
      *dst = accumulate ? *dst + c : c;
    };
    st(mt0, nt0, c00);
    if (n1) st(mt0, nt0 + 1, c01);
    if (m1) st(mt0 + 1, nt0, c10);
    if (m1 && n1) st(mt0 + 1, nt0 + 1, c11);
  }
  // bias: db[co] = sum_p dout[p][co] on the last wave (it has the fewest
  // tile blocks when they do not divide evenly): lane l sums the pixels
  // p = l / cout (mod 64 / cout) of channel l % cout, and the partial sums
  // are added across the lanes in a fixed order (deterministic, no barrier)
  if (wave == nw - 1 && cout <= 64) {
    const int nparts = 64 / cout;
    const int co = lane % cout, part = lane / cout;
    float acc = 0.0f;
    if (part < nparts)
      for (int p = part; p < HW; p += nparts) acc += dout[(int64_t)p * cs_out + co_off + co];
    for (int k = 1; k < nparts; ++k) {
      const float o = __shfl(acc, lane + k * cout, 64);
      if (part == 0) acc += o;
    }
    if (part == 0) {
      if (accumulate) db[co] += acc;
      else db[co] = acc;
    }
  } else if (cout > 64) {
    for (int co = threadIdx.x; co < cout; co += blockDim.x) {
      float acc = 0.0f;
      for (int p = 0; p < HW; ++p) acc += dout[(int64_t)p * cs_out + co_off + co];
      if (accumulate) db[co] += acc;
      else db[co] = acc;
    }
  }
}

// Per-sample workspace of the MFMA kernel: the scalar kernel's layout plus six
// gradient tensors (dV, dH, dVr, dHr, t0, t1) of HW*C floats.
__device__ __forceinline__ float* ws_grad(float* ws, int HW, int C, int nb, int which) {
  return ws + ((HW + 3) & ~3) + ((int64_t)nb * 5 + which) * HW * C;
}

// Source tables (and their inverses) of every tap / shift of the lattice.
__device__ void build_tap_tables(int16_t* ftab, int16_t* itab, int H, int W) {
  const int HW = H * W;
  for (int idx = threadIdx.x; idx < GT_N * HW; idx += blockDim.x) itab[idx] = -1;
  for (int idx = threadIdx.x; idx < GT_N * HW; idx += blockDim.x) {
    const int row = idx / HW, p = idx - row * HW;
    const int i = p / W, j = p - (p / W) * W;
    int dr = 0, dc = 0;
    if (row < GT_H) tap_rc(0, row - GT_V, dr, dc);
    else if (row < GT_C) tap_rc(1, row - GT_H, dr, dc);
    else if (row < GT_X) tap_rc(2, row - GT_C, dr, dc);
    else if (row == GT_XF) dc = -1;
    else if (row == GT_Y) dr = -1;
    const int r = i + dr, c = j + dc;
    const bool ok = r >= 0 && r < H && c >= 0 && c < W;
    ftab[idx] = ok ? (int16_t)(r * W + c) : (int16_t)-1;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < GT_N * HW; idx += blockDim.x) {
    const int row = idx / HW, p = idx - row * HW;
    const int src = ftab[idx];
    if (src >= 0) itab[row * HW + src] = (int16_t)p;
  }
  __syncthreads();
}

// Forward of one sample (MFMA form), every block's tensors stored in ws.
// With a tangent direction (tan != null) the forward-mode derivative of every
// tensor is carried along in wt (same layout as ws): for y = K*x + b,
// y' = K*x' + Kt*x + bt, and act(y)' = act'(y) y'.  t0/t1 and u0/u1 are
// scratch tensors of the primal and the tangent.
// STG: the kxk and 1xk convolutions read their inputs from the LDS staging
// tiles S0, S1 (row stride SC); otherwise from L2.
template <int ACT, bool STG, bool TAN>
__device__ __forceinline__ void mfma_forward(const float* prm, const float* tan, const float* kt, const float* tkt,
                                             float* ws, float* wt, const int16_t* ftab, int H, int W, int C, int nb,
                                             float* t0, float* t1, float* u0, float* u1, float* S0, float* S1,
                                             int SC) {
  const int HW = H * W, F = nb - 1;
  const float* sp = ws;
  for (int blk = 0; blk < nb; ++blk) {
    const ParamOffsets o = param_offsets(blk, C);
    const bool first = blk == 0, fin = blk == F;
    const bool kb = !first && !fin && ((blk & 1) == 0);
    const float* Vin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_VO);
    const float* Hin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_HO);
    float* VA = ws_tensor(ws, HW, C, blk, T_VA);
    float* XA = ws_tensor(ws, HW, C, blk, T_XA);
    float* CAT = ws_tensor(ws, HW, C, blk, T_CAT);
    float* VO = ws_tensor(ws, HW, C, blk, T_VO);
    float* HO = ws_tensor(ws, HW, C, blk, T_HO);
    // C-channel tensors in 16-byte units: element (p, co..co+3) is quad (p C + co) / 4
    auto q4 = [&](int p, int co) { return (p * C + co) >> 2; };
    const float* VOr = kb ? ws_tensor(ws, HW, C, blk - 2, T_VO) : nullptr;  // the residuals
    const float* HOr = kb ? ws_tensor(ws, HW, C, blk - 2, T_HO) : nullptr;
    if (!TAN && !first) {
      // primal only: the activations and residuals run in the convolutions'
      // epilogues (the same operations on the same values as the passes of
      // the tangent form below)
      const float* Vb = STG ? S0 : Vin;
      const float* Hb = STG ? S1 : Hin;
      const int vs = STG ? SC : C;
      if constexpr (STG) {
        stage_in(S0, SC, Vin, C, 0, C, HW);
        stage_in(S1, SC, Hin, C, 0, C, HW);
        __syncthreads();
      }
      mconv_fwd_post(0, Vb, vs, C, C, kt + o.kv, prm + o.bv, ftab + GT_V * HW, 9, HW, [&](int p, int co, f32x4 v) {
        const int i = q4(p, co);
        const f32x4 va = bact4<ACT>(v);
        S4(VA, i) = va;
        if (!fin) S4(VO, i) = kb ? bact4<ACT>(L4(VOr, i) + v) : va;
      });
      mconv_fwd_post(0, Hb, vs, C, C, kt + o.kh, prm + o.bh, ftab + GT_H * HW, 3, HW,
                     [&](int p, int co, f32x4 h) { S4(XA, q4(p, co)) = bact4<ACT>(h); });
    } else if (first) {
      conv_fwd(t0, C, Vin, 1, 1, C, prm + o.kv, prm + o.bv, 0, 9, H, W);
      conv_fwd(t1, C, Hin, 1, 1, C, prm + o.kh, prm + o.bh, 1, 3, H, W);
      if (tan) {  // the spins carry no tangent
        conv_fwd(u0, C, Vin, 1, 1, C, tan + o.kv, tan + o.bv, 0, 9, H, W);
        conv_fwd(u1, C, Hin, 1, 1, C, tan + o.kh, tan + o.bh, 1, 3, H, W);
      }
    } else {
      const float* Vb = STG ? S0 : Vin;
      const float* Hb = STG ? S1 : Hin;
      const int vs = STG ? SC : C;
      if constexpr (STG) {
        stage_in(S0, SC, Vin, C, 0, C, HW);
        stage_in(S1, SC, Hin, C, 0, C, HW);
        __syncthreads();
      }
      mconv_fwd(t0, C, 0, Vb, vs, C, C, kt + o.kv, prm + o.bv, ftab + GT_V * HW, 9, HW);
      mconv_fwd(t1, C, 0, Hb, vs, C, C, kt + o.kh, prm + o.bh, ftab + GT_H * HW, 3, HW);
      if (tan) {
        const float* Vd = ws_tensor(wt, HW, C, blk - 1, T_VO);
        const float* Hd = ws_tensor(wt, HW, C, blk - 1, T_HO);
        mconv_fwd(u0, C, 0, Vd, C, C, C, kt + o.kv, tan + o.bv, ftab + GT_V * HW, 9, HW);
        mconv_fwd(u1, C, 0, Hd, C, C, C, kt + o.kh, tan + o.bh, ftab + GT_H * HW, 3, HW);
        __syncthreads();
        mconv_fwd(u0, C, 0, Vb, vs, C, C, tkt + o.kv, nullptr, ftab + GT_V * HW, 9, HW, true);
        mconv_fwd(u1, C, 0, Hb, vs, C, C, tkt + o.kh, nullptr, ftab + GT_H * HW, 3, HW, true);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ((TAN || first) ? HW * C / 4 : 0); i += blockDim.x) {
      const f32x4 v = L4(t0, i);
      const f32x4 va = bact4<ACT>(v), xa = bact4<ACT>(L4(t1, i));
      S4(VA, i) = va;
      S4(XA, i) = xa;
      f32x4 vo = {0.f, 0.f, 0.f, 0.f};
      if (!fin) {
        vo = kb ? bact4<ACT>(L4(VOr, i) + v) : va;
        S4(VO, i) = vo;
      }
      if (tan) {
        const f32x4 dv = L4(u0, i);
        S4(ws_tensor(wt, HW, C, blk, T_VA), i) = bdact4<ACT>(va) * dv;
        S4(ws_tensor(wt, HW, C, blk, T_XA), i) = bdact4<ACT>(xa) * L4(u1, i);
        if (!fin) {
          const f32x4 din = kb ? L4(ws_tensor(wt, HW, C, blk - 2, T_VO), i) + dv : dv;
          S4(ws_tensor(wt, HW, C, blk, T_VO), i) = bdact4<ACT>(vo) * din;
        }
      }
    }
    if (TAN || first) __syncthreads();
    if constexpr (!TAN) {
      auto cat_post = [&](int p, int co, f32x4 c) { S4(CAT, q4(p, co)) = bact4<ACT>(c); };
      mconv_fwd_post(0, XA, C, C, C / 2, kt + o.kx, prm + o.bx, ftab + (fin ? GT_XF : GT_X) * HW, 1, HW, cat_post);
      mconv_fwd_post(C / 2, VA, C, C, C / 2, kt + o.ky, prm + o.by, ftab + GT_Y * HW, 1, HW, cat_post);
    } else {
      mconv_fwd(CAT, C, 0, XA, C, C, C / 2, kt + o.kx, prm + o.bx, ftab + (fin ? GT_XF : GT_X) * HW, 1, HW);
      mconv_fwd(CAT, C, C / 2, VA, C, C, C / 2, kt + o.ky, prm + o.by, ftab + GT_Y * HW, 1, HW);
    }
    if (TAN) {
      float* CATd = ws_tensor(wt, HW, C, blk, T_CAT);
      const float* XAd = ws_tensor(wt, HW, C, blk, T_XA);
      const float* VAd = ws_tensor(wt, HW, C, blk, T_VA);
      mconv_fwd(CATd, C, 0, XAd, C, C, C / 2, kt + o.kx, tan + o.bx, ftab + (fin ? GT_XF : GT_X) * HW, 1, HW);
      mconv_fwd(CATd, C, C / 2, VAd, C, C, C / 2, kt + o.ky, tan + o.by, ftab + GT_Y * HW, 1, HW);
      __syncthreads();
      mconv_fwd(CATd, C, 0, XA, C, C, C / 2, tkt + o.kx, nullptr, ftab + (fin ? GT_XF : GT_X) * HW, 1, HW, true);
      mconv_fwd(CATd, C, C / 2, VA, C, C, C / 2, tkt + o.ky, nullptr, ftab + GT_Y * HW, 1, HW, true);
    }
    __syncthreads();
    if constexpr (TAN) {
      for (int i = threadIdx.x; i < HW * C / 4; i += blockDim.x) {
        const f32x4 c = bact4<ACT>(L4(CAT, i));
        S4(CAT, i) = c;
        S4(ws_tensor(wt, HW, C, blk, T_CAT), i) *= bdact4<ACT>(c);
      }
      __syncthreads();
    }
    const float* Cb = STG ? S0 : CAT;
    const int cs = STG ? SC : C;
    if constexpr (STG) {
      stage_in(S0, SC, CAT, C, 0, C, HW);
      __syncthreads();
    }
    if constexpr (!TAN) {
      mconv_fwd_post(0, Cb, cs, C, C, kt + o.kc, prm + o.bc, ftab + GT_C * HW, 9, HW, [&](int p, int co, f32x4 h) {
        const int i = q4(p, co);
        S4(HO, i) = kb ? bact4<ACT>(L4(HOr, i) + h) : bact4<ACT>(h);
      });
      __syncthreads();
      continue;
    }
    mconv_fwd(t0, C, 0, Cb, cs, C, C, kt + o.kc, prm + o.bc, ftab + GT_C * HW, 9, HW);
    {
      mconv_fwd(u0, C, 0, ws_tensor(wt, HW, C, blk, T_CAT), C, C, C, kt + o.kc, tan + o.bc, ftab + GT_C * HW, 9,
                HW);
      __syncthreads();
      mconv_fwd(u0, C, 0, Cb, cs, C, C, tkt + o.kc, nullptr, ftab + GT_C * HW, 9, HW, true);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HW * C / 4; i += blockDim.x) {
      const f32x4 h = L4(t0, i);
      const f32x4 ho = kb ? bact4<ACT>(L4(HOr, i) + h) : bact4<ACT>(h);
      S4(HO, i) = ho;
      {
        const f32x4 din = kb ? L4(ws_tensor(wt, HW, C, blk - 2, T_HO), i) + L4(u0, i) : L4(u0, i);
        S4(ws_tensor(wt, HW, C, blk, T_HO), i) = bdact4<ACT>(ho) * din;
      }
    }
    __syncthreads();
  }
}

// Head logits z[k] (k = 0, 1: Re for spin +1 / -1; 2, 3: Im) of pixel p.
__device__ __forceinline__ void head_logits(const float* prm, int64_t ho, const float* HF, int C, int p, float z[4]) {
  for (int k = 0; k < 4; ++k) {
    float acc = prm[ho + 4 * C + k];
    for (int ci = 0; ci < C; ++ci) acc += prm[ho + ci * 4 + k] * HF[p * C + ci];
    z[k] = acc;
  }
}

// Backward of one sample from the head cotangent dz (HW x 4): d/d theta into
// O (overwritten, or added to when accumulate).
template <int ACT, bool ACC, bool STG>
__device__ __forceinline__ void mfma_backward(const float* prm, float* ws, const int16_t* ftab, const int16_t* itab,
                                              const float* dz, int H, int W, int C, int nb, float* O, float* dV,
                                              float* dH, float* dVr, float* dHr, float* t0, float* t1, float* S0,
                                              float* S1, int SC) {
  constexpr bool accumulate = ACC;
  const int HW = H * W, F = nb - 1;
  const float* sp = ws;
  const int64_t ho = head_offset(nb, C);
  const float* HF = ws_tensor(ws, HW, C, F, T_HO);
  for (int idx = threadIdx.x; idx < 4 * C + 4; idx += blockDim.x) {
    float acc = 0.0f;
    if (idx < 4 * C) {
      const int ci = idx / 4, k = idx - (idx / 4) * 4;
      for (int p = 0; p < HW; ++p) acc += HF[p * C + ci] * dz[p * 4 + k];
    } else {
      for (int p = 0; p < HW; ++p) acc += dz[p * 4 + idx - 4 * C];
    }
    if (accumulate) O[ho + idx] += acc;
    else O[ho + idx] = acc;
  }
  for (int idx = threadIdx.x; idx < HW * C; idx += blockDim.x) {
    const int p = idx / C, ci = idx - (idx / C) * C;
    float g = 0.0f;
    for (int k = 0; k < 4; ++k) g += prm[ho + ci * 4 + k] * dz[p * 4 + k];
    dH[idx] = g;
    dV[idx] = 0.0f;
    dVr[idx] = 0.0f;
    dHr[idx] = 0.0f;
  }
  __syncthreads();
  for (int blk = F; blk >= 0; --blk) {
    const ParamOffsets o = param_offsets(blk, C);
    const bool first = blk == 0, fin = blk == F;
    const bool kb = !first && !fin && ((blk & 1) == 0);
    const bool ka = !first && !fin && ((blk & 1) == 1);
    const float* Vin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_VO);
    const float* Hin = first ? sp : ws_tensor(ws, HW, C, blk - 1, T_HO);
    const float* VA = ws_tensor(ws, HW, C, blk, T_VA);
    const float* XA = ws_tensor(ws, HW, C, blk, T_XA);
    const float* CAT = ws_tensor(ws, HW, C, blk, T_CAT);
    const float* VO = ws_tensor(ws, HW, C, blk, T_VO);
    const float* HO = ws_tensor(ws, HW, C, blk, T_HO);
    for (int i = threadIdx.x; i < HW * C / 4; i += blockDim.x) {
      const f32x4 gh = L4(dH, i) * bdact4<ACT>(L4(HO, i));
      S4(t1, i) = gh;
      const f32x4 gv = fin ? f32x4{0.f, 0.f, 0.f, 0.f} : L4(dV, i) * bdact4<ACT>(L4(VO, i));
      S4(t0, i) = gv;
      if (kb) {
        S4(dVr, i) = gv;
        S4(dHr, i) = gh;
      }
    }
    __syncthreads();
    {
      const float* ib = STG ? S1 : CAT;
      const float* db = STG ? S0 : t1;
      const int is = STG ? SC : C, ds = STG ? SC : C;
      if constexpr (STG) {
        stage_in(S1, SC, CAT, C, 0, C, HW);
        stage_in(S0, SC, t1, C, 0, C, HW);
        __syncthreads();
      }
      mconv_bwd_weight(O + o.kc, O + o.bc, ib, is, db, ds, 0, C, C, ftab + GT_C * HW, 9, HW, accumulate);
      mconv_bwd_data(dV, C, db, ds, 0, C, C, prm + o.kc, itab + GT_C * HW, 9, HW);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HW * C / 4; i += blockDim.x) S4(dV, i) *= bdact4<ACT>(L4(CAT, i));
    __syncthreads();
    const int xr = fin ? GT_XF : GT_X;
    mconv_bwd_weight(O + o.kx, O + o.bx, XA, C, dV, C, 0, C, C / 2, ftab + xr * HW, 1, HW, accumulate);
    mconv_bwd_weight(O + o.ky, O + o.by, VA, C, dV, C, C / 2, C, C / 2, ftab + GT_Y * HW, 1, HW, accumulate);
    mconv_bwd_data(dH, C, dV, C, 0, C, C / 2, prm + o.kx, itab + xr * HW, 1, HW);
    mconv_bwd_data(t1, C, dV, C, C / 2, C, C / 2, prm + o.ky, itab + GT_Y * HW, 1, HW);
    __syncthreads();
    for (int i = threadIdx.x; i < HW * C / 4; i += blockDim.x) {
      S4(dH, i) *= bdact4<ACT>(L4(XA, i));
      S4(t0, i) += L4(t1, i) * bdact4<ACT>(L4(VA, i));
    }
    __syncthreads();
    if (first) {
      conv_bwd_weight(O + o.kh, O + o.bh, Hin, 1, dH, C, 1, C, 1, 3, H, W, accumulate);
      conv_bwd_weight(O + o.kv, O + o.bv, Vin, 1, t0, C, 1, C, 0, 9, H, W, accumulate);
      break;
    }
    {
      const float* ib = STG ? S1 : Vin;
      const float* db = STG ? S0 : t0;
      const int is = STG ? SC : C, ds = STG ? SC : C;
      if constexpr (STG) {
        stage_in(S1, SC, Vin, C, 0, C, HW);
        stage_in(S0, SC, t0, C, 0, C, HW);
        __syncthreads();
      }
      mconv_bwd_weight(O + o.kh, O + o.bh, Hin, C, dH, C, 0, C, C, ftab + GT_H * HW, 3, HW, accumulate);
      mconv_bwd_weight(O + o.kv, O + o.bv, ib, is, db, ds, 0, C, C, ftab + GT_V * HW, 9, HW, accumulate);
      mconv_bwd_data(t1, C, dH, C, 0, C, C, prm + o.kh, itab + GT_H * HW, 3, HW);
      mconv_bwd_data(dV, C, db, ds, 0, C, C, prm + o.kv, itab + GT_V * HW, 9, HW);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HW * C / 4; i += blockDim.x) {
      f32x4 gh = L4(t1, i), gv = L4(dV, i);
      if (ka) {
        gh += L4(dHr, i);
        gv += L4(dVr, i);
      }
      S4(dH, i) = gh;
      S4(dV, i) = gv;
    }
    __syncthreads();
  }
  __syncthreads();
}

enum { GRAD_O = 0, GRAD_VJP = 1 };

// MODE GRAD_O: one workgroup per sample, O_re / O_im rows from two backward
// passes (cotangents of Re and Im log psi).  MODE GRAD_VJP: a persistent grid,
// each workgroup walks samples b = blockIdx.x, blockIdx.x + gridDim.x, ...
// and adds cot_re[b] d Re log psi_b + cot_im[b] d Im log psi_b (one backward
// pass with the combined head cotangent) into its own partial gradient row.
// STG: staging tiles in LDS (a.stage is their row stride).
// WPC: workgroups per CU the register budget is set for.
template <int ACT, int MODE, bool STG, int WPC = 4>
__global__ __launch_bounds__(256, WPC) void convar_grad_mfma_kernel(BwdArgs a) {
  extern __shared__ int16_t tabs[];  // GT_N forward rows, then GT_N inverse rows, of HW entries
  __shared__ float s_lp[2];
  const int H = a.H, W = a.W, HW = a.HW, C = a.C, nb = a.nb;
  const int F = nb - 1;
  if (MODE == GRAD_O && (int64_t)blockIdx.x >= a.batch) return;
  float* ws = a.work + (int64_t)blockIdx.x * a.work_floats;
  int16_t* ftab = tabs;
  int16_t* itab = tabs + GT_N * HW;
  float* dz = reinterpret_cast<float*>(tabs + 2 * GT_N * HW + (2 * GT_N * HW & 1));
  // staging tiles after dz (4 HW floats; 16-byte aligned: 2 GT_N HW int16
  // is a multiple of 16 bytes)
  float* S0 = dz + 4 * HW;
  float* S1 = S0 + HW * a.stage;
  float* dV = ws_grad(ws, HW, C, nb, 0);
  float* dH = ws_grad(ws, HW, C, nb, 1);
  float* dVr = ws_grad(ws, HW, C, nb, 2);
  float* dHr = ws_grad(ws, HW, C, nb, 3);
  float* t0 = ws_grad(ws, HW, C, nb, 4);
  float* t1 = ws_grad(ws, HW, C, nb, 5);
  const float* prm = a.params;
  build_tap_tables(ftab, itab, H, W);
  float* part = nullptr;
  if (MODE == GRAD_VJP) {
    part = a.partial + (int64_t)blockIdx.x * a.P;
    for (int64_t q = threadIdx.x; q < a.P; q += blockDim.x) part[q] = 0.0f;
    __syncthreads();
  }
  const int64_t step = MODE == GRAD_VJP ? (int64_t)gridDim.x : a.batch;
  for (int64_t b = blockIdx.x; b < a.batch; b += step) {
    float* sp = ws;
    for (int q = threadIdx.x; q < HW; q += blockDim.x) sp[q] = (float)a.spins[b * HW + q];
    if (threadIdx.x == 0) s_lp[0] = s_lp[1] = 0.0f;
    __syncthreads();
    mfma_forward<ACT, STG, false>(prm, nullptr, a.kt, nullptr, ws, nullptr, ftab, H, W, C, nb, t0, t1, nullptr,
                                  nullptr,
                           S0, S1, a.stage);
    const int64_t ho = head_offset(nb, C);
    const float* HF = ws_tensor(ws, HW, C, F, T_HO);
    const float cr = MODE == GRAD_VJP ? a.cot[2 * b] : 1.0f;
    const float ci = MODE == GRAD_VJP ? a.cot[2 * b + 1] : 1.0f;
    for (int part_i = 0; part_i < (MODE == GRAD_VJP ? 1 : 2); ++part_i) {
      for (int p = threadIdx.x; p < HW; p += blockDim.x) {
        float z[4];
        head_logits(prm, ho, HF, C, p, z);
        const float mx = fmaxf(z[0], z[1]);
        const float e0 = expf(2.0f * (z[0] - mx)), e1 = expf(2.0f * (z[1] - mx));
        const float p0 = e0 / (e0 + e1), p1 = e1 / (e0 + e1);
        const bool up = sp[p] > 0.0f;
        const float wr = (MODE == GRAD_VJP || part_i == 0) ? cr : 0.0f;
        const float wi = (MODE == GRAD_VJP || part_i == 1) ? ci : 0.0f;
        dz[p * 4 + 0] = wr * ((up ? 1.0f : 0.0f) - p0);
        dz[p * 4 + 1] = wr * ((up ? 0.0f : 1.0f) - p1);
        dz[p * 4 + 2] = wi * (up ? 1.0f : 0.0f);
        dz[p * 4 + 3] = wi * (up ? 0.0f : 1.0f);
        if (part_i == 0) {
          const float lse = mx + 0.5f * logf(e0 + e1);
          atomicAdd(&s_lp[0], (up ? z[0] : z[1]) - lse);
          atomicAdd(&s_lp[1], up ? z[2] : z[3]);
        }
      }
      __syncthreads();
      float* O = MODE == GRAD_VJP ? part : (part_i == 0 ? a.o_re : a.o_im) + b * a.ld;
      mfma_backward<ACT, MODE == GRAD_VJP, STG>(prm, ws, ftab, itab, dz, H, W, C, nb, O, dV, dH, dVr, dHr, t0, t1,
                                                S0, S1, a.stage);
    }
    if (threadIdx.x == 0 && a.logpsi) {
      a.logpsi[2 * b] = s_lp[0];
      a.logpsi[2 * b + 1] = s_lp[1];
    }
    __syncthreads();
  }
}

// Forward-mode derivative (JVP) of log psi along a parameter direction:
// out[b] = (d Re log psi_b, d Im log psi_b) . tan, a persistent grid of
// workgroups over the samples, primal and tangent tensors in the workspace.
template <int ACT, bool STG>
__global__ __launch_bounds__(256, 4) void convar_jvp_mfma_kernel(BwdArgs a) {
  extern __shared__ int16_t tabs[];
  __shared__ float s_acc[4];
  const int H = a.H, W = a.W, HW = a.HW, C = a.C, nb = a.nb;
  const int F = nb - 1;
  float* ws = a.work + (int64_t)blockIdx.x * a.work_floats;
  float* wt = ws + a.work_floats / 2;  // tangent region, same layout
  int16_t* ftab = tabs;
  int16_t* itab = tabs + GT_N * HW;
  float* S0 = reinterpret_cast<float*>(tabs + 2 * GT_N * HW) + 4 * HW;
  float* S1 = S0 + HW * a.stage;
  float* t0 = ws_grad(ws, HW, C, nb, 0);
  float* t1 = ws_grad(ws, HW, C, nb, 1);
  float* u0 = ws_grad(ws, HW, C, nb, 2);
  float* u1 = ws_grad(ws, HW, C, nb, 3);
  const float* prm = a.params;
  const float* tan = a.tangent;
  build_tap_tables(ftab, itab, H, W);
  for (int64_t b = blockIdx.x; b < a.batch; b += gridDim.x) {
    float* sp = ws;
    for (int q = threadIdx.x; q < HW; q += blockDim.x) sp[q] = (float)a.spins[b * HW + q];
    if (threadIdx.x < 4) s_acc[threadIdx.x] = 0.0f;
    __syncthreads();
    mfma_forward<ACT, STG, true>(prm, tan, a.kt, a.tkt, ws, wt, ftab, H, W, C, nb, t0, t1, u0, u1, S0, S1, a.stage);
    const int64_t ho = head_offset(nb, C);
    const float* HF = ws_tensor(ws, HW, C, F, T_HO);
    const float* HFd = ws_tensor(wt, HW, C, F, T_HO);
    for (int p = threadIdx.x; p < HW; p += blockDim.x) {
      float z[4], zd[4];
      head_logits(prm, ho, HF, C, p, z);
      for (int k = 0; k < 4; ++k) {
        float acc = tan[ho + 4 * C + k];
        for (int ci = 0; ci < C; ++ci)
          acc += prm[ho + ci * 4 + k] * HFd[p * C + ci] + tan[ho + ci * 4 + k] * HF[p * C + ci];
        zd[k] = acc;
      }
      const float mx = fmaxf(z[0], z[1]);
      const float e0 = expf(2.0f * (z[0] - mx)), e1 = expf(2.0f * (z[1] - mx));
      const float p0 = e0 / (e0 + e1), p1 = e1 / (e0 + e1);
      const bool up = sp[p] > 0.0f;
      // d/dz of z_s - 0.5 log(e^{2 z0} + e^{2 z1}) is delta_s - p
      atomicAdd(&s_acc[0], (up ? zd[0] : zd[1]) - (p0 * zd[0] + p1 * zd[1]));
      atomicAdd(&s_acc[1], up ? zd[2] : zd[3]);
      const float lse = mx + 0.5f * logf(e0 + e1);
      atomicAdd(&s_acc[2], (up ? z[0] : z[1]) - lse);
      atomicAdd(&s_acc[3], up ? z[2] : z[3]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      a.jvp[2 * b] = s_acc[0];
      a.jvp[2 * b + 1] = s_acc[1];
      if (a.logpsi) {
        a.logpsi[2 * b] = s_acc[2];
        a.logpsi[2 * b + 1] = s_acc[3];
      }
    }
    __syncthreads();
  }
}

// Sum of the per-workgroup partial gradients in a fixed order (FP64), scaled.
__global__ void reduce_partials_kernel(const float* __restrict__ part, int64_t nparts, int64_t P, double scale,
                                       float* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P) return;
  double acc = 0.0;
  for (int64_t g = 0; g < nparts; ++g) acc += (double)part[g * P + q];
  out[q] = (float)(scale * acc);
}

int64_t grad_work_floats(int HW, int C, int nb) {
  return (int64_t)((HW + 3) & ~3) + (((int64_t)nb * 5 + 6) * HW * C + 3) / 4 * 4;
}

int64_t flat_param_count(int C, int nb) {
  const int64_t c = C;
  return (9 * c + c + 3 * c + c + c * (c / 2) * 2 + c + 9 * c * c + c) +
         (int64_t)(nb - 1) * (9 * c * c + c + 3 * c * c + c + c * (c / 2) * 2 + c + 9 * c * c + c) + 4 * c + 4;
}

// Workgroups of the persistent VJP / JVP grids: as many 4-wave workgroups
// per CU as the kernel's registers allow (VJP: vjp_per_cu; JVP 4), at most
// one per sample.
int64_t implicit_grid(int64_t batch, int per_cu) {
  const int64_t g = (int64_t)per_cu * device_cu_count();
  return batch < g ? (batch > 0 ? batch : 1) : g;
}
constexpr int JVP_PER_CU = 4;

size_t grad_lds_bytes(int HW) {
  return (size_t)2 * GT_N * HW * sizeof(int16_t) + 4 + (size_t)4 * HW * sizeof(float);
}

// Staging tiles (two of HW x (C + 4) floats) when they keep a workgroup
// within a quarter of the CU's LDS (four workgroups per CU): the stride, or 0.
constexpr size_t GRAD_LDS_PER_WG = 40 * 1024;
int stage_stride(int HW, int C) {
  const int SC = C + 4;
  const size_t b = grad_lds_bytes(HW) + (size_t)2 * HW * SC * sizeof(float);
  return (C % 4 == 0 && b <= GRAD_LDS_PER_WG) ? SC : 0;
}
// Workgroups per CU of the batched VJP: four with staging tiles (128 VGPRs,
// 24.1 ms per 2048 samples at 10x10/C32 against 27.9 at three), three
// without (the unstaged 20x20/C64 implicit products of 16384 samples took
// 430 us per sample at three, 472 at four).
int vjp_per_cu(int HW, int C) { return stage_stride(HW, C) ? 4 : 3; }
size_t grad_lds_total(int HW, int stage) {
  return grad_lds_bytes(HW) + (stage ? (size_t)2 * HW * stage * sizeof(float) : 0);
}

template <class K>
int launch_with_lds(K kern, int64_t grid, size_t lds, hipStream_t s, const BwdArgs& a) {
  FK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), lds, s, a);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

// Floats reserved at the start of a workspace for transposed kernels (P
// rounded up to 64 floats, so what follows stays 256-byte aligned).
int64_t kt_floats(int C, int nb) { return (flat_param_count(C, nb) + 63) / 64 * 64; }

template <int ACT, int MODE>
int launch_grad(int64_t grid, size_t lds, hipStream_t s, const BwdArgs& a) {
  if (a.stage) return launch_with_lds(convar_grad_mfma_kernel<ACT, MODE, true>, grid, lds, s, a);
  if (MODE == GRAD_VJP) return launch_with_lds(convar_grad_mfma_kernel<ACT, MODE, false, 3>, grid, lds, s, a);
  return launch_with_lds(convar_grad_mfma_kernel<ACT, MODE, false>, grid, lds, s, a);
}

int launch_transpose(const float* prm, float* kt, int C, int nb, hipStream_t s) {
  const int64_t n = 9LL * C * C;
  const unsigned gx = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(transpose_kernels_kernel, dim3(gx, (unsigned)(nb * 5)), dim3(256), 0, s, prm, kt, C, nb);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // namespace fk

using namespace fk;

extern "C" {

int64_t fk_convar_grad_workspace_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int64_t batch) {
  if (H < 1 || W < 1 || depth < 1 || channels < 2 || (channels % 2) || batch < 0) return -1;
  const int nb = num_causal_blocks(depth);
  return (kt_floats(channels, nb) + batch * grad_work_floats(H * W, channels, nb)) * (int64_t)sizeof(float);
}

int fk_convar_grad(const float* params, int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t activation,
                   const int8_t* spins, int64_t batch, void* workspace, float* o_re, float* o_im, int64_t row_stride,
                   float* logpsi, void* stream) {
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1 && channels >= 2 && channels % 2 == 0, "fk_convar_grad: bad shape");
  FK_REQUIRE(activation == FK_ACT_RELU || activation == FK_ACT_ELU, "fk_convar_grad: bad activation");
  if (batch <= 0) return FK_OK;
  FK_REQUIRE(params && spins && workspace && o_re && o_im, "fk_convar_grad: null pointer");
  FK_REQUIRE(batch <= 0x7fffffff, "fk_convar_grad: batch too large");
  BwdArgs a{};
  a.params = params;
  a.H = H;
  a.W = W;
  a.HW = H * W;
  a.C = channels;
  a.nb = num_causal_blocks(depth);
  a.batch = batch;
  a.spins = spins;
  a.kt = reinterpret_cast<float*>(workspace);
  a.work = reinterpret_cast<float*>(workspace) + kt_floats(channels, a.nb);
  a.work_floats = grad_work_floats(a.HW, channels, a.nb);
  a.o_re = o_re;
  a.o_im = o_im;
  a.logpsi = logpsi;
  a.P = flat_param_count(channels, a.nb);
  FK_REQUIRE(row_stride == 0 || row_stride >= a.P, "fk_convar_grad: row stride shorter than the parameters");
  a.ld = row_stride > 0 ? row_stride : a.P;
  hipStream_t s = (hipStream_t)stream;
  if (channels % 32 == 0) {
    // MFMA kernel: LDS holds the tap source tables and the head cotangents
    a.stage = stage_stride(a.HW, channels);
    const size_t lds = grad_lds_total(a.HW, a.stage);
    FK_REQUIRE(lds <= 160 * 1024 && a.HW <= 32767, "fk_convar_grad: lattice too large");
    const int rc = launch_transpose(params, const_cast<float*>(a.kt), channels, a.nb, s);
    if (rc != FK_OK) return rc;
    return activation == FK_ACT_ELU
               ? launch_grad<FK_ACT_ELU, GRAD_O>(batch, lds, s, a)
               : launch_grad<FK_ACT_RELU, GRAD_O>(batch, lds, s, a);
  }
  const size_t lds = ((size_t)6 * a.HW * channels + 4 * a.HW) * sizeof(float);
  FK_REQUIRE(lds <= 160 * 1024, "fk_convar_grad: lattice too large for the LDS gradient buffers");
  if (activation == FK_ACT_ELU) {
    FK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(convar_grad_kernel<FK_ACT_ELU>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(convar_grad_kernel<FK_ACT_ELU>, dim3((unsigned)batch), dim3(256), lds, s, a);
  } else {
    FK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(convar_grad_kernel<FK_ACT_RELU>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(convar_grad_kernel<FK_ACT_RELU>, dim3((unsigned)batch), dim3(256), lds, s, a);
  }
  FK_CHECK_LAUNCH();
  return FK_OK;
}

/* ---- implicit products: batched VJP and JVP (no per-sample O) ---------- */

int64_t fk_convar_vjp_workspace_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int64_t batch) {
  if (H < 1 || W < 1 || depth < 1 || channels < 32 || (channels % 32) || batch < 0) return -1;
  const int nb = num_causal_blocks(depth);
  const int64_t g = implicit_grid(batch, vjp_per_cu(H * W, channels));
  return (kt_floats(channels, nb) + g * (grad_work_floats(H * W, channels, nb) + flat_param_count(channels, nb))) *
         (int64_t)sizeof(float);
}

int fk_convar_vjp(const float* params, int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t activation,
                  const int8_t* spins, int64_t batch, const float* cot, double scale, void* workspace, float* grad,
                  float* logpsi, void* stream) {
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1 && channels >= 32 && channels % 32 == 0,
             "fk_convar_vjp: the batched VJP needs a channel count that is a multiple of 32");
  FK_REQUIRE(activation == FK_ACT_RELU || activation == FK_ACT_ELU, "fk_convar_vjp: bad activation");
  FK_REQUIRE(batch >= 1 && batch <= 0x7fffffff, "fk_convar_vjp: bad batch");
  FK_REQUIRE(params && spins && cot && workspace && grad, "fk_convar_vjp: null pointer");
  BwdArgs a{};
  a.params = params;
  a.H = H;
  a.W = W;
  a.HW = H * W;
  a.C = channels;
  a.nb = num_causal_blocks(depth);
  a.batch = batch;
  a.spins = spins;
  a.kt = reinterpret_cast<float*>(workspace);
  a.work = reinterpret_cast<float*>(workspace) + kt_floats(channels, a.nb);
  a.work_floats = grad_work_floats(a.HW, channels, a.nb);
  a.logpsi = logpsi;
  a.P = flat_param_count(channels, a.nb);
  a.cot = cot;
  const int64_t g = implicit_grid(batch, vjp_per_cu(H * W, channels));
  a.partial = a.work + g * a.work_floats;
  a.stage = stage_stride(a.HW, channels);
  const size_t lds = grad_lds_total(a.HW, a.stage);
  FK_REQUIRE(lds <= 160 * 1024 && a.HW <= 32767, "fk_convar_vjp: lattice too large");
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_transpose(params, const_cast<float*>(a.kt), channels, a.nb, s);
  if (rc != FK_OK) return rc;
  rc = activation == FK_ACT_ELU ? launch_grad<FK_ACT_ELU, GRAD_VJP>(g, lds, s, a)
                                : launch_grad<FK_ACT_RELU, GRAD_VJP>(g, lds, s, a);
  if (rc != FK_OK) return rc;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(div_up(a.P, 256)), dim3(256), 0, s, a.partial, g, a.P, scale,
                     grad);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_convar_jvp_workspace_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int64_t batch) {
  if (H < 1 || W < 1 || depth < 1 || channels < 32 || (channels % 32) || batch < 0) return -1;
  const int nb = num_causal_blocks(depth);
  return (2 * kt_floats(channels, nb) + implicit_grid(batch, JVP_PER_CU) * 2 * grad_work_floats(H * W, channels, nb)) *
         (int64_t)sizeof(float);
}

int fk_convar_jvp(const float* params, const float* tangent, int32_t H, int32_t W, int32_t depth, int32_t channels,
                  int32_t activation, const int8_t* spins, int64_t batch, void* workspace, float* jvp, float* logpsi,
                  void* stream) {
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1 && channels >= 32 && channels % 32 == 0,
             "fk_convar_jvp: the JVP needs a channel count that is a multiple of 32");
  FK_REQUIRE(activation == FK_ACT_RELU || activation == FK_ACT_ELU, "fk_convar_jvp: bad activation");
  FK_REQUIRE(batch >= 1 && batch <= 0x7fffffff, "fk_convar_jvp: bad batch");
  FK_REQUIRE(params && tangent && spins && workspace && jvp, "fk_convar_jvp: null pointer");
  BwdArgs a{};
  a.params = params;
  a.H = H;
  a.W = W;
  a.HW = H * W;
  a.C = channels;
  a.nb = num_causal_blocks(depth);
  a.batch = batch;
  a.spins = spins;
  a.kt = reinterpret_cast<float*>(workspace);
  a.tkt = a.kt + kt_floats(channels, a.nb);
  a.work = reinterpret_cast<float*>(workspace) + 2 * kt_floats(channels, a.nb);
  a.work_floats = 2 * grad_work_floats(a.HW, channels, a.nb);
  a.logpsi = logpsi;
  a.P = flat_param_count(channels, a.nb);
  a.tangent = tangent;
  a.jvp = jvp;
  a.stage = stage_stride(a.HW, channels);
  const size_t lds = grad_lds_total(a.HW, a.stage);
  FK_REQUIRE(lds <= 160 * 1024 && a.HW <= 32767, "fk_convar_jvp: lattice too large");
  const int64_t g = implicit_grid(batch, JVP_PER_CU);
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_transpose(params, const_cast<float*>(a.kt), channels, a.nb, s);
  if (rc == FK_OK) rc = launch_transpose(tangent, const_cast<float*>(a.tkt), channels, a.nb, s);
  if (rc != FK_OK) return rc;
  if (a.stage)
    return activation == FK_ACT_ELU ? launch_with_lds(convar_jvp_mfma_kernel<FK_ACT_ELU, true>, g, lds, s, a)
                                    : launch_with_lds(convar_jvp_mfma_kernel<FK_ACT_RELU, true>, g, lds, s, a);
  return activation == FK_ACT_ELU ? launch_with_lds(convar_jvp_mfma_kernel<FK_ACT_ELU, false>, g, lds, s, a)
                                  : launch_with_lds(convar_jvp_mfma_kernel<FK_ACT_RELU, false>, g, lds, s, a);
}

}  // extern "C"
