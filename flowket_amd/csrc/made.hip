// Learned autoregressive proposal for the Markov chains: a masked MLP (MADE)
// x (N spins) -> h1 (Hd) -> h2 (Hd) -> N logits, ReLU, log q(x) =
// sum_i log sigmoid(+-logit_i).  Autoregressive masks from unit degrees:
// input i has degree i+1, hidden unit h has degree 1 + h mod (N-1); a
// hidden unit sees inputs of degree <= its own, a second-layer unit sees
// first-layer units of degree <= its own, and logit i sees hidden units of
// degree <= i (so logit 0 is its bias alone).
//
// Weights are stored in BF16 (forward, sampling), FP32 master copies are
// trained by maximum likelihood on the chains' samples (Adam); the biases are
// read straight from the FP32 master vector.  Log q and sampling run the
// degree-ordered per-lane evaluation (made_lane_kernel); the fit's backward
// uses a masked tile forward, 8 configurations per 256-thread workgroup with
// activations in LDS and coalesced weight reads (second layer stored
// transposed).
#include "fk_common.h"

namespace fk {

constexpr int MADE_S = 8;      // configurations per workgroup
constexpr int MADE_T = 256;    // threads per workgroup

struct MadeW {
  const __bf16 *w1;   // (Hd, N)   [h][i]
  const float* b1;    // (Hd)
  const __bf16* w2t;  // (Hd, Hd)  [k][t] = W2[t][k]
  const float* b2;    // (Hd)
  const __bf16* w3;   // (N, Hd)   [o][h]
  const float* b3;    // (N)
  int N, Hd;
};

__device__ __forceinline__ int deg_hidden(int h, int N) { return 1 + h % (N - 1); }

// forward of the MADE for the MADE_S configurations in s_x; leaves h1, h2 (post
// ReLU) and the logits in LDS.  Masks are applied through the degrees.
__device__ void made_forward_tile(const MadeW& w, const float* s_x, float* s_h1, float* s_h2, float* s_l) {
  const int N = w.N, Hd = w.Hd;
  for (int t = threadIdx.x; t < Hd; t += blockDim.x) {
    const int mh = deg_hidden(t, N);
    float acc[MADE_S];
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) acc[s] = w.b1[t];
    for (int i = 0; i < N && i + 1 <= mh; ++i) {
      const float wv = (float)w.w1[(int64_t)t * N + i];
#pragma unroll
      for (int s = 0; s < MADE_S; ++s) acc[s] += wv * s_x[s * N + i];
    }
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) s_h1[s * Hd + t] = fmaxf(acc[s], 0.0f);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < Hd; t += blockDim.x) {
    const int mt = deg_hidden(t, N);
    float acc[MADE_S];
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) acc[s] = w.b2[t];
    for (int k = 0; k < Hd; ++k) {
      if (deg_hidden(k, N) > mt) continue;
      const float wv = (float)w.w2t[(int64_t)k * Hd + t];
#pragma unroll
      for (int s = 0; s < MADE_S; ++s) acc[s] += wv * s_h1[s * Hd + k];
    }
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) s_h2[s * Hd + t] = fmaxf(acc[s], 0.0f);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < N * MADE_S; idx += blockDim.x) {
    const int o = idx % N, s = idx / N;
    float acc = w.b3[o];
    for (int h = 0; h < Hd; ++h)
      if (deg_hidden(h, N) <= o) acc += (float)w.w3[(int64_t)o * Hd + h] * s_h2[s * Hd + h];
    s_l[s * N + o] = acc;
  }
  __syncthreads();
}

__device__ __forceinline__ float log_sigmoid(float z) { return z >= 0.0f ? -log1pf(expf(-z)) : z - log1pf(expf(z)); }

// Per-sample backward of L = -log q(x): stores the layer inputs and the
// gradients of the pre-activations (dl = sigmoid(l) - [x == +1],
// dh2 = (W3^T dl) [h2 > 0], dh1 = (W2^T dh2) [h1 > 0]) for the weight
// reduction.
__global__ __launch_bounds__(MADE_T) void made_backward_kernel(MadeW w, const int8_t* __restrict__ x, int64_t B,
                                                               float* __restrict__ g_h1, float* __restrict__ g_h2,
                                                               float* __restrict__ g_dl, float* __restrict__ g_dh2,
                                                               float* __restrict__ g_dh1) {
  extern __shared__ float sm[];
  const int N = w.N, Hd = w.Hd;
  float* s_x = sm;
  float* s_h1 = s_x + MADE_S * N;
  float* s_h2 = s_h1 + MADE_S * Hd;
  float* s_l = s_h2 + MADE_S * Hd;
  float* s_d2 = s_l + MADE_S * N;
  const int64_t b0 = (int64_t)blockIdx.x * MADE_S;
  for (int idx = threadIdx.x; idx < MADE_S * N; idx += blockDim.x) {
    const int64_t b = b0 + idx / N;
    s_x[idx] = b < B ? (float)x[b * N + idx % N] : 0.0f;
  }
  __syncthreads();
  made_forward_tile(w, s_x, s_h1, s_h2, s_l);
  // dl = sigmoid(l) - t ; rows of absent configurations are zero
  for (int idx = threadIdx.x; idx < MADE_S * N; idx += blockDim.x) {
    const int s = idx / N;
    const float t = s_x[idx] > 0.0f ? 1.0f : 0.0f;
    const float d = (b0 + s < B) ? (1.0f / (1.0f + expf(-s_l[idx])) - t) : 0.0f;
    s_l[idx] = d;
  }
  __syncthreads();
  // dh2 = (W3^T dl) * [h2 > 0]
  for (int t = threadIdx.x; t < Hd; t += blockDim.x) {
    const int mt = deg_hidden(t, N);
    float acc[MADE_S];
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) acc[s] = 0.0f;
    for (int o = mt; o < N; ++o) {
      const float wv = (float)w.w3[(int64_t)o * Hd + t];
#pragma unroll
      for (int s = 0; s < MADE_S; ++s) acc[s] += wv * s_l[s * N + o];
    }
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) s_d2[s * Hd + t] = s_h2[s * Hd + t] > 0.0f ? acc[s] : 0.0f;
  }
  __syncthreads();
  // dh1 = (W2^T dh2) * [h1 > 0]
  for (int k = threadIdx.x; k < Hd; k += blockDim.x) {
    const int mk = deg_hidden(k, N);
    float acc[MADE_S];
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) acc[s] = 0.0f;
    for (int t = 0; t < Hd; ++t) {
      if (deg_hidden(t, N) < mk) continue;
      const float wv = (float)w.w2t[(int64_t)k * Hd + t];
#pragma unroll
      for (int s = 0; s < MADE_S; ++s) acc[s] += wv * s_d2[s * Hd + t];
    }
#pragma unroll
    for (int s = 0; s < MADE_S; ++s) {
      const int64_t b = b0 + s;
      if (b < B) g_dh1[b * Hd + k] = s_h1[s * Hd + k] > 0.0f ? acc[s] : 0.0f;
    }
  }
  for (int idx = threadIdx.x; idx < MADE_S * Hd; idx += blockDim.x) {
    const int64_t b = b0 + idx / Hd;
    if (b < B) {
      g_h1[b * Hd + idx % Hd] = s_h1[idx];
      g_h2[b * Hd + idx % Hd] = s_h2[idx];
      g_dh2[b * Hd + idx % Hd] = s_d2[idx];
    }
  }
  for (int idx = threadIdx.x; idx < MADE_S * N; idx += blockDim.x) {
    const int64_t b = b0 + idx / N;
    if (b < B) g_dl[b * N + idx % N] = s_l[idx];
  }
}

// Weight gradients as tiled products over the batch: part[r][c] =
// sum_{b in this split} A[b][r] * Bm[b][c] (Bm = +-1 spins when xs is given,
// ones when both are null: the bias sums).  A 64x64 output tile per
// workgroup, a 4x4 block per thread, the batch in chunks of 64 staged
// through LDS; blockIdx.z is the batch split, whose partial sums are added
// in split order by the Adam step (deterministic).
constexpr int WG_TILE = 64, WG_KC = 64;
__global__ __launch_bounds__(256) void made_wgrad_kernel(const float* __restrict__ A, int lda,
                                                         const float* __restrict__ Bm, const int8_t* __restrict__ xs,
                                                         int ldb, int rows, int cols, int64_t B, int64_t per_split,
                                                         float* __restrict__ out, int64_t split_stride) {
  __shared__ __align__(16) float sa[WG_KC][WG_TILE];
  __shared__ __align__(16) float sb[WG_KC][WG_TILE];
  const int r0 = blockIdx.y * WG_TILE, c0 = blockIdx.x * WG_TILE;
  const int64_t k0 = (int64_t)blockIdx.z * per_split;
  const int64_t k1 = (k0 + per_split < B) ? k0 + per_split : B;
  const int tid = threadIdx.x, tr = tid / 16, tc = tid % 16;
  float acc[4][4] = {};
  for (int64_t kb = k0; kb < k1; kb += WG_KC) {
    // all of a chunk's loads are issued before its first LDS store
    constexpr int PER = WG_KC * WG_TILE / 256;
    const int j = tid % WG_TILE;
    float va[PER], vb[PER];
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int kk = it * (256 / WG_TILE) + tid / WG_TILE;
      const int64_t b = kb + kk;
      const bool kin = b < k1;
      va[it] = (kin && r0 + j < rows) ? A[b * lda + r0 + j] : 0.0f;
      float v = 0.0f;
      if (kin && c0 + j < cols) v = xs ? (float)xs[b * ldb + c0 + j] : (Bm ? Bm[b * ldb + c0 + j] : 1.0f);
      vb[it] = v;
    }
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int kk = it * (256 / WG_TILE) + tid / WG_TILE;
      sa[kk][j] = va[it];
      sb[kk][j] = vb[it];
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < WG_KC; ++kk) {
      const f32x4 av = *reinterpret_cast<const f32x4*>(&sa[kk][4 * tr]);
      const f32x4 bv = *reinterpret_cast<const f32x4*>(&sb[kk][4 * tc]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
    }
    __syncthreads();
  }
  float* o = out + blockIdx.z * split_stride;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + 4 * tr + i, c = c0 + 4 * tc + j;
      if (r < rows && c < cols) o[(int64_t)r * cols + c] = acc[i][j];
    }
}

// Masked gradients (mean over the batch, from the split partial sums) and
// one Adam step on the FP32 master weights; the BF16 copies are refreshed.
// One thread per parameter of the packed master vector
// [w1 (Hd*N) | b1 | w2t (Hd*Hd) | b2 | w3 (N*Hd) | b3].
__global__ void made_adam_kernel(float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
                                 __bf16* __restrict__ w1, __bf16* __restrict__ w2t, __bf16* __restrict__ w3,
                                 const float* __restrict__ part, int splits, int64_t B, int N, int Hd, float lr,
                                 float beta1, float beta2, float eps, float bc1, float bc2) {
  const int64_t n1 = (int64_t)Hd * N, n2 = (int64_t)Hd * Hd, n3 = (int64_t)N * Hd;
  const int64_t total = n1 + Hd + n2 + Hd + n3 + N;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  bool masked = false;
  int64_t r = k;
  if (r < n1) {  // w1[h][i]
    const int h = (int)(r / N), i = (int)(r % N);
    masked = !(i + 1 <= deg_hidden(h, N));
  } else if ((r -= n1) < Hd) {  // b1
  } else if ((r -= Hd) < n2) {  // w2t[kk][t] = W2[t][kk]
    const int kk = (int)(r / Hd), t = (int)(r % Hd);
    masked = deg_hidden(kk, N) > deg_hidden(t, N);
  } else if ((r -= n2) < Hd) {  // b2
  } else if ((r -= Hd) < n3) {  // w3[o][h]
    const int o = (int)(r / Hd), h = (int)(r % Hd);
    masked = deg_hidden(h, N) > o;
  }
  double g = 0.0;
  for (int sidx = 0; sidx < splits; ++sidx) g += part[(int64_t)sidx * total + k];
  const float gf = masked ? 0.0f : (float)(g / (double)B);
  const float mm = beta1 * m1[k] + (1.0f - beta1) * gf;
  const float vv = beta2 * m2[k] + (1.0f - beta2) * gf * gf;
  m1[k] = mm;
  m2[k] = vv;
  float p = master[k];
  if (!masked) p -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
  else p = 0.0f;
  master[k] = p;
  // refresh the BF16 copy
  r = k;
  if (r < n1) w1[r] = (__bf16)p;
  else if ((r -= n1) < Hd) {
  } else if ((r -= Hd) < n2) w2t[r] = (__bf16)p;
  else if ((r -= n2) < Hd) {
  } else if ((r -= Hd) < n3) w3[r] = (__bf16)p;
}

size_t made_lds(int N, int Hd) { return (size_t)MADE_S * (2 * N + 3 * Hd) * sizeof(float); }

// ---------------------------------------------------------------------------
// Log q and sampling on MFMA.  Hidden units are renumbered in degree order, so
// the units a site finalises are one contiguous block and everything they
// read is a contiguous prefix.  A unit of degree d depends only on inputs
// x_0..x_{d-1}: at site i the units of degree i are computed once (first
// layer from x_<i, second layer from first-layer units of degree <= i) and
// logit i reads second-layer units of degree <= i -- each a 16-row block of
// v_mfma_f32_16x16x32_bf16 with the 16 samples of a wave as its columns.
// fk_made_pack writes the degree-sorted, masked BF16 matrices once per
// refit; hidden activations are held in LDS as BF16 (part of the proposal's
// definition: log q and sampling share this code path, so the density the
// acceptance uses is the one the sampler drew from).
//
// Packed layout (16-byte aligned sections): A1 (Hd+16, K1P) bf16 first-layer
// rows; A2 (Hd+16, HP) bf16 second layer [t][k]; A3 (N+16, HP) bf16 logits;
// b1 (Hd+16), b2 (Hd+16), b3 (N+16) fp32 -- K1P = N, HP = Hd rounded up to 32.
struct MadePacked {
  int N, Hd, K1P, HP;
  int64_t o_a1, o_a2, o_a3, o_b1, o_b2, o_b3, bytes;
  __host__ __device__ MadePacked(int n, int hd) : N(n), Hd(hd) {
    K1P = (n + 31) / 32 * 32;
    HP = (hd + 31) / 32 * 32;
    auto al = [](int64_t v) { return (v + 15) / 16 * 16; };
    o_a1 = 0;
    o_a2 = al(o_a1 + (int64_t)(hd + 16) * K1P * 2);
    o_a3 = al(o_a2 + (int64_t)(hd + 16) * HP * 2);
    o_b1 = al(o_a3 + (int64_t)(n + 16) * HP * 2);
    o_b2 = al(o_b1 + (int64_t)(hd + 16) * 4);
    o_b3 = al(o_b2 + (int64_t)(hd + 16) * 4);
    bytes = al(o_b3 + (int64_t)(n + 16) * 4);
  }
};

// first sorted position of the units of degree d (1-based degrees)
__host__ __device__ inline int made_deg_start(int d, int N, int Hd) {
  const int step = N - 1;
  int s = 0;
  for (int e = 1; e < d; ++e) s += (e - 1 < Hd) ? (Hd - (e - 1) + step - 1) / step : 0;
  return s;
}
// unit index at sorted position r
__device__ inline int made_sorted_unit(int r, int N, int Hd) {
  const int step = N - 1;
  int s = 0;
  for (int d = 1; d < N; ++d) {
    const int c = (d - 1 < Hd) ? (Hd - (d - 1) + step - 1) / step : 0;
    if (r < s + c) return (d - 1) + (r - s) * step;
    s += c;
  }
  return -1;
}

__global__ void made_pack_kernel(MadeW w, uint8_t* __restrict__ out) {
  const MadePacked L(w.N, w.Hd);
  const int N = w.N, Hd = w.Hd;
  __bf16* a1 = reinterpret_cast<__bf16*>(out + L.o_a1);
  __bf16* a2 = reinterpret_cast<__bf16*>(out + L.o_a2);
  __bf16* a3 = reinterpret_cast<__bf16*>(out + L.o_a3);
  float* b1 = reinterpret_cast<float*>(out + L.o_b1);
  float* b2 = reinterpret_cast<float*>(out + L.o_b2);
  float* b3 = reinterpret_cast<float*>(out + L.o_b3);
  const int64_t n1 = (int64_t)(Hd + 16) * L.K1P, n2 = (int64_t)(Hd + 16) * L.HP, n3 = (int64_t)(N + 16) * L.HP;
  const int64_t total = n1 + n2 + n3 + 2 * (Hd + 16) + (N + 16);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = idx;
    if (r < n1) {
      const int row = (int)(r / L.K1P), j = (int)(r % L.K1P);
      float v = 0.0f;
      if (row < Hd && j < N) {
        const int h = made_sorted_unit(row, N, Hd);
        if (j + 1 <= deg_hidden(h, N)) v = (float)w.w1[(int64_t)h * N + j];
      }
      a1[r] = (__bf16)v;
    } else if ((r -= n1) < n2) {
      const int row = (int)(r / L.HP), c = (int)(r % L.HP);
      float v = 0.0f;
      if (row < Hd && c < Hd) {
        const int t = made_sorted_unit(row, N, Hd), k = made_sorted_unit(c, N, Hd);
        if (deg_hidden(k, N) <= deg_hidden(t, N)) v = (float)w.w2t[(int64_t)k * Hd + t];
      }
      a2[r] = (__bf16)v;
    } else if ((r -= n2) < n3) {
      const int o = (int)(r / L.HP), c = (int)(r % L.HP);
      float v = 0.0f;
      if (o < N && c < Hd) {
        const int t = made_sorted_unit(c, N, Hd);
        if (deg_hidden(t, N) <= o) v = (float)w.w3[(int64_t)o * Hd + t];
      }
      a3[r] = (__bf16)v;
    } else if ((r -= n3) < Hd + 16) {
      b1[r] = r < Hd ? w.b1[made_sorted_unit((int)r, N, Hd)] : 0.0f;
    } else if ((r -= Hd + 16) < Hd + 16) {
      b2[r] = r < Hd ? w.b2[made_sorted_unit((int)r, N, Hd)] : 0.0f;
    } else {
      r -= Hd + 16;
      b3[r] = r < N ? w.b3[r] : 0.0f;
    }
  }
}

constexpr int MADE_MS = 16;  // samples per wave (the MFMA N dimension)

// One wave per 16 samples.  LDS: X (16, K1P), H1 (16, HP), H2 (16, HP) bf16.
__global__ __launch_bounds__(64) void made_mfma_kernel(const uint8_t* __restrict__ packed, int N, int Hd,
                                                        const int8_t* __restrict__ x_in,
                                                        const double* __restrict__ u, int64_t B,
                                                        int8_t* __restrict__ x_out, float* __restrict__ logq) {
  const MadePacked L(N, Hd);
  extern __shared__ __align__(16) uint8_t made_sm[];
  __bf16* X = reinterpret_cast<__bf16*>(made_sm);
  __bf16* H1 = X + MADE_MS * L.K1P;
  __bf16* H2 = H1 + MADE_MS * L.HP;
  const __bf16* A1 = reinterpret_cast<const __bf16*>(packed + L.o_a1);
  const __bf16* A2 = reinterpret_cast<const __bf16*>(packed + L.o_a2);
  const __bf16* A3 = reinterpret_cast<const __bf16*>(packed + L.o_a3);
  const float* b1 = reinterpret_cast<const float*>(packed + L.o_b1);
  const float* b2 = reinterpret_cast<const float*>(packed + L.o_b2);
  const float* b3 = reinterpret_cast<const float*>(packed + L.o_b3);
  const int lane = threadIdx.x, n = lane & 15, g = lane >> 4;
  const int64_t b = (int64_t)blockIdx.x * MADE_MS + n;
  const bool live = b < B;
  for (int idx = lane; idx < MADE_MS * L.K1P; idx += 64) {
    const int sn = idx / L.K1P, j = idx - sn * L.K1P;
    const int64_t bb = (int64_t)blockIdx.x * MADE_MS + sn;
    float v = 0.0f;
    if (x_in && j < N && bb < B) v = (float)x_in[bb * N + j];
    X[idx] = (__bf16)v;
  }
  for (int idx = lane; idx < 2 * MADE_MS * L.HP; idx += 64) H1[idx] = (__bf16)0.0f;
  __syncthreads();

  // sorted units of degree < d start at pos; degree-i block is [s0, s1)
  int s0 = 0;
  float lq = 0.0f;
  for (int i = 0; i < N; ++i) {
    const int s1 = (i >= 1) ? s0 + ((i - 1 < Hd) ? (Hd - (i - 1) + N - 2) / (N - 1) : 0) : 0;
    if (i >= 1 && s1 > s0) {
      // first layer: rows [s0, s1), inputs j < i
      for (int rb = s0; rb < s1; rb += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks * 32 < i; ++ks) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(A1 + (int64_t)(rb + n) * L.K1P + 32 * ks + 8 * g);
          const bf16x8 bx = *reinterpret_cast<const bf16x8*>(X + n * L.K1P + 32 * ks + 8 * g);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bx, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rb + 4 * g + r;
          if (row < s1) H1[n * L.HP + row] = (__bf16)fmaxf(acc[r] + b1[row], 0.0f);
        }
      }
      __syncthreads();
      // second layer: rows [s0, s1), first-layer units [0, s1)
      for (int rb = s0; rb < s1; rb += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks * 32 < s1; ++ks) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(A2 + (int64_t)(rb + n) * L.HP + 32 * ks + 8 * g);
          const bf16x8 bh = *reinterpret_cast<const bf16x8*>(H1 + n * L.HP + 32 * ks + 8 * g);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rb + 4 * g + r;
          if (row < s1) H2[n * L.HP + row] = (__bf16)fmaxf(acc[r] + b2[row], 0.0f);
        }
      }
      __syncthreads();
    }
    // logit i from second-layer units [0, s1) (row 0 of the block starting at i)
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks * 32 < s1; ++ks) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(A3 + (int64_t)(i + n) * L.HP + 32 * ks + 8 * g);
      const bf16x8 bh = *reinterpret_cast<const bf16x8*>(H2 + n * L.HP + 32 * ks + 8 * g);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bh, acc, 0, 0, 0);
    }
    if (g == 0) {
      const float z = acc[0] + b3[i];
      float xi;
      if (x_in) {
        xi = live ? (float)x_in[b * N + i] : 1.0f;
      } else {
        const double uu = live ? u[b * N + i] : 0.5;
        const float p = 1.0f / (1.0f + expf(-z));
        xi = ((double)p > uu) ? 1.0f : -1.0f;
        X[n * L.K1P + i] = (__bf16)xi;
        if (live) x_out[b * N + i] = (int8_t)xi;
      }
      lq += log_sigmoid(xi > 0.0f ? z : -z);
    }
    __syncthreads();
    s0 = s1;
  }
  if (g == 0 && live) logq[b] = lq;
}

size_t made_mfma_lds(int N, int Hd) {
  const MadePacked L(N, Hd);
  return (size_t)MADE_MS * (L.K1P + 2 * L.HP) * sizeof(__bf16);
}

}  // namespace fk

using namespace fk;

extern "C" {

int64_t fk_made_packed_bytes(int32_t n_sites, int32_t hidden) {
  if (n_sites < 2 || hidden < 1) return -1;
  return MadePacked(n_sites, hidden).bytes;
}

int fk_made_pack(const void* w1, const float* b1, const void* w2t, const float* b2, const void* w3, const float* b3,
                 int32_t n_sites, int32_t hidden, void* packed, void* stream) {
  FK_REQUIRE(n_sites >= 2 && hidden >= 1, "fk_made_pack: bad sizes");
  FK_REQUIRE(w1 && b1 && w2t && b2 && w3 && b3 && packed, "fk_made_pack: null pointer");
  MadeW w{(const __bf16*)w1, b1, (const __bf16*)w2t, b2, (const __bf16*)w3, b3, n_sites, hidden};
  hipLaunchKernelGGL(made_pack_kernel, dim3(256), dim3(256), 0, (hipStream_t)stream, w,
                     reinterpret_cast<uint8_t*>(packed));
  FK_CHECK_LAUNCH();
  return FK_OK;
}

static int made_run(const void* packed, int32_t n_sites, int32_t hidden, const int8_t* x_in, const double* u,
                    int64_t batch, int8_t* x_out, float* logq, void* stream) {
  const size_t lds = made_mfma_lds(n_sites, hidden);
  FK_REQUIRE(lds <= 64 * 1024, "MADE: too large for LDS");
  hipLaunchKernelGGL(made_mfma_kernel, dim3(div_up(batch, MADE_MS)), dim3(64), lds, (hipStream_t)stream,
                     reinterpret_cast<const uint8_t*>(packed), n_sites, hidden, x_in, u, batch, x_out, logq);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_made_log_prob(const void* packed, int32_t n_sites, int32_t hidden, const int8_t* x, int64_t batch,
                     float* logq, void* stream) {
  FK_REQUIRE(n_sites >= 2 && hidden >= 1, "fk_made_log_prob: bad sizes");
  if (batch <= 0) return FK_OK;
  FK_REQUIRE(packed && x && logq, "fk_made_log_prob: null pointer");
  return made_run(packed, n_sites, hidden, x, nullptr, batch, nullptr, logq, stream);
}

int fk_made_sample(const void* packed, int32_t n_sites, int32_t hidden, const double* uniforms, int64_t batch,
                   int8_t* x, float* logq, void* stream) {
  FK_REQUIRE(n_sites >= 2 && hidden >= 1, "fk_made_sample: bad sizes");
  if (batch <= 0) return FK_OK;
  FK_REQUIRE(packed && uniforms && x && logq, "fk_made_sample: null pointer");
  return made_run(packed, n_sites, hidden, nullptr, uniforms, batch, x, logq, stream);
}

// batch samples per split of the weight-gradient products
static constexpr int64_t MADE_SPLIT = 512;

int64_t fk_made_fit_workspace_floats(int32_t n_sites, int32_t hidden, int64_t batch) {
  const int64_t Hd = hidden, N = n_sites;
  const int64_t total = Hd * N + Hd + Hd * Hd + Hd + N * Hd + N;
  const int64_t splits = (batch + MADE_SPLIT - 1) / MADE_SPLIT;
  return batch * (3LL * hidden + hidden + n_sites) + splits * total + 2;
}

int fk_made_fit_step(float* master, float* adam_m, float* adam_v, void* w1, float* b1, void* w2t, float* b2, void* w3,
                     float* b3, int32_t n_sites, int32_t hidden, const int8_t* x, int64_t batch, float* workspace,
                     float lr, float beta1, float beta2, float eps, int64_t step, void* stream) {
  FK_REQUIRE(n_sites >= 2 && hidden >= 1 && batch > 0 && step >= 1, "fk_made_fit_step: bad sizes");
  FK_REQUIRE(master && adam_m && adam_v && w1 && b1 && w2t && b2 && w3 && b3 && x && workspace,
             "fk_made_fit_step: null pointer");
  const int64_t Hd = hidden, N = n_sites;
  float* g_h1 = workspace;
  float* g_h2 = g_h1 + batch * Hd;
  float* g_dh2 = g_h2 + batch * Hd;
  float* g_dh1 = g_dh2 + batch * Hd;
  float* g_dl = g_dh1 + batch * Hd;
  MadeW w{(const __bf16*)w1, b1, (const __bf16*)w2t, b2, (const __bf16*)w3, b3, n_sites, hidden};
  const size_t lds = made_lds(n_sites, hidden) + (size_t)MADE_S * hidden * sizeof(float);
  FK_REQUIRE(lds <= 64 * 1024, "fk_made_fit_step: too large for LDS");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(made_backward_kernel, dim3(div_up(batch, MADE_S)), dim3(MADE_T), lds, s, w, x, batch, g_h1,
                     g_h2, g_dl, g_dh2, g_dh1);
  FK_CHECK_LAUNCH();
  const int64_t total = Hd * N + Hd + Hd * Hd + Hd + N * Hd + N;
  const int splits = (int)((batch + MADE_SPLIT - 1) / MADE_SPLIT);
  FK_REQUIRE(splits <= 65535, "fk_made_fit_step: batch too large");
  float* part = g_dl + batch * N;
  // per-split partial sums, laid out like the master vector
  auto wgrad = [&](const float* A, int lda, const float* Bm, const int8_t* xs, int ldb, int rows, int cols,
                   int64_t off) {
    hipLaunchKernelGGL(made_wgrad_kernel, dim3(div_up(cols, WG_TILE), div_up(rows, WG_TILE), splits), dim3(256),
                       0, s, A, lda, Bm, xs, ldb, rows, cols, batch, MADE_SPLIT, part + off, total);
  };
  const int64_t o_b1 = Hd * N, o_w2 = o_b1 + Hd, o_b2 = o_w2 + Hd * Hd, o_w3 = o_b2 + Hd, o_b3 = o_w3 + N * Hd;
  wgrad(g_dh1, (int)Hd, nullptr, x, (int)N, (int)Hd, (int)N, 0);                 // w1[h][i]: dh1^T x
  wgrad(g_dh1, (int)Hd, nullptr, nullptr, 1, (int)Hd, 1, o_b1);                  // b1
  wgrad(g_h1, (int)Hd, g_dh2, nullptr, (int)Hd, (int)Hd, (int)Hd, o_w2);         // w2t[kk][t]: h1^T dh2
  wgrad(g_dh2, (int)Hd, nullptr, nullptr, 1, (int)Hd, 1, o_b2);                  // b2
  wgrad(g_dl, (int)N, g_h2, nullptr, (int)Hd, (int)N, (int)Hd, o_w3);            // w3[o][h]: dl^T h2
  wgrad(g_dl, (int)N, nullptr, nullptr, 1, (int)N, 1, o_b3);                     // b3
  FK_CHECK_LAUNCH();
  const float bc1 = 1.0f - powf(beta1, (float)step), bc2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(made_adam_kernel, dim3(div_up(total, 256)), dim3(256), 0, s, master, adam_m, adam_v,
                     (__bf16*)w1, (__bf16*)w2t, (__bf16*)w3, part, splits, batch, n_sites, hidden, lr, beta1, beta2,
                     eps, bc1, bc2);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
