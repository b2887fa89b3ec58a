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
// read straight from the FP32 master vector.  One workgroup of 256 threads
// handles 8 configurations; activations live in LDS, weights are read
// coalesced (second layer stored transposed).
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

__global__ __launch_bounds__(MADE_T) void made_logq_kernel(MadeW w, const int8_t* __restrict__ x, int64_t B,
                                                           float* __restrict__ logq) {
  extern __shared__ float sm[];
  const int N = w.N, Hd = w.Hd;
  float* s_x = sm;
  float* s_h1 = s_x + MADE_S * N;
  float* s_h2 = s_h1 + MADE_S * Hd;
  float* s_l = s_h2 + MADE_S * Hd;
  const int64_t b0 = (int64_t)blockIdx.x * MADE_S;
  for (int idx = threadIdx.x; idx < MADE_S * N; idx += blockDim.x) {
    const int64_t b = b0 + idx / N;
    s_x[idx] = b < B ? (float)x[b * N + idx % N] : 0.0f;
  }
  __syncthreads();
  made_forward_tile(w, s_x, s_h1, s_h2, s_l);
  if (threadIdx.x < MADE_S) {
    const int s = threadIdx.x;
    if (b0 + s < B) {
      float acc = 0.0f;
      for (int o = 0; o < N; ++o) acc += log_sigmoid(s_x[s * N + o] > 0.0f ? s_l[s * N + o] : -s_l[s * N + o]);
      logq[b0 + s] = acc;
    }
  }
}

// x_i ~ q(x_i | x_<i) in site order: s_i = +1 iff sigmoid(logit_i) > u_i
__global__ __launch_bounds__(MADE_T) void made_sample_kernel(MadeW w, const double* __restrict__ u, int64_t B,
                                                             int8_t* __restrict__ x, float* __restrict__ logq) {
  extern __shared__ float sm[];
  const int N = w.N, Hd = w.Hd;
  float* s_x = sm;
  float* s_h1 = s_x + MADE_S * N;
  float* s_h2 = s_h1 + MADE_S * Hd;
  float* s_l = s_h2 + MADE_S * Hd;
  __shared__ float s_lq[MADE_S];
  const int64_t b0 = (int64_t)blockIdx.x * MADE_S;
  for (int idx = threadIdx.x; idx < MADE_S * N; idx += blockDim.x) s_x[idx] = 0.0f;
  if (threadIdx.x < MADE_S) s_lq[threadIdx.x] = 0.0f;
  __syncthreads();
  for (int i = 0; i < N; ++i) {
    made_forward_tile(w, s_x, s_h1, s_h2, s_l);
    if (threadIdx.x < MADE_S) {
      const int s = threadIdx.x;
      const int64_t b = b0 + s;
      const float z = s_l[s * N + i];
      const double uu = b < B ? u[b * N + i] : 0.5;
      const float p = 1.0f / (1.0f + expf(-z));
      const bool up = (double)p > uu;
      s_x[s * N + i] = up ? 1.0f : -1.0f;
      s_lq[s] += log_sigmoid(up ? z : -z);
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < MADE_S * N; idx += blockDim.x) {
    const int64_t b = b0 + idx / N;
    if (b < B) x[b * N + idx % N] = (int8_t)s_x[idx];
  }
  if (threadIdx.x < MADE_S && b0 + threadIdx.x < B) logq[b0 + threadIdx.x] = s_lq[threadIdx.x];
}

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

// Masked gradients (mean over the batch) and one Adam step on the FP32
// master weights; the BF16 copies are refreshed.  One thread per parameter
// of the packed master vector [w1 (Hd*N) | b1 | w2t (Hd*Hd) | b2 | w3 (N*Hd) | b3].
__global__ void made_adam_kernel(float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
                                 __bf16* __restrict__ w1, __bf16* __restrict__ w2t, __bf16* __restrict__ w3,
                                 const int8_t* __restrict__ x, const float* __restrict__ g_h1,
                                 const float* __restrict__ g_h2, const float* __restrict__ g_dl,
                                 const float* __restrict__ g_dh2, const float* __restrict__ g_dh1, int64_t B, int N,
                                 int Hd, float lr, float beta1, float beta2, float eps, float bc1, float bc2) {
  const int64_t n1 = (int64_t)Hd * N, n2 = (int64_t)Hd * Hd, n3 = (int64_t)N * Hd;
  const int64_t total = n1 + Hd + n2 + Hd + n3 + N;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  double g = 0.0;
  bool masked = false;
  int64_t r = k;
  if (r < n1) {  // w1[h][i]
    const int h = (int)(r / N), i = (int)(r % N);
    masked = !(i + 1 <= deg_hidden(h, N));
    if (!masked)
      for (int64_t b = 0; b < B; ++b) g += (double)g_dh1[b * Hd + h] * (double)x[b * N + i];
  } else if ((r -= n1) < Hd) {  // b1
    for (int64_t b = 0; b < B; ++b) g += g_dh1[b * Hd + r];
  } else if ((r -= Hd) < n2) {  // w2t[kk][t] = W2[t][kk]
    const int kk = (int)(r / Hd), t = (int)(r % Hd);
    masked = deg_hidden(kk, N) > deg_hidden(t, N);
    if (!masked)
      for (int64_t b = 0; b < B; ++b) g += (double)g_dh2[b * Hd + t] * (double)g_h1[b * Hd + kk];
  } else if ((r -= n2) < Hd) {  // b2
    for (int64_t b = 0; b < B; ++b) g += g_dh2[b * Hd + r];
  } else if ((r -= Hd) < n3) {  // w3[o][h]
    const int o = (int)(r / Hd), h = (int)(r % Hd);
    masked = deg_hidden(h, N) > o;
    if (!masked)
      for (int64_t b = 0; b < B; ++b) g += (double)g_dl[b * N + o] * (double)g_h2[b * Hd + h];
  } else {  // b3
    r -= n3;
    for (int64_t b = 0; b < B; ++b) g += g_dl[b * N + r];
  }
  float gf = masked ? 0.0f : (float)(g / (double)B);
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

}  // namespace fk

using namespace fk;

extern "C" {

int fk_made_log_prob(const void* w1, const float* b1, const void* w2t, const float* b2, const void* w3,
                     const float* b3, int32_t n_sites, int32_t hidden, const int8_t* x, int64_t batch, float* logq,
                     void* stream) {
  FK_REQUIRE(n_sites >= 2 && hidden >= 1, "fk_made_log_prob: bad sizes");
  if (batch <= 0) return FK_OK;
  FK_REQUIRE(w1 && b1 && w2t && b2 && w3 && b3 && x && logq, "fk_made_log_prob: null pointer");
  const size_t lds = made_lds(n_sites, hidden);
  FK_REQUIRE(lds <= 64 * 1024, "fk_made_log_prob: too large for LDS");
  MadeW w{(const __bf16*)w1, b1, (const __bf16*)w2t, b2, (const __bf16*)w3, b3, n_sites, hidden};
  hipLaunchKernelGGL(made_logq_kernel, dim3(div_up(batch, MADE_S)), dim3(MADE_T), lds, (hipStream_t)stream, w, x,
                     batch, logq);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_made_sample(const void* w1, const float* b1, const void* w2t, const float* b2, const void* w3, const float* b3,
                   int32_t n_sites, int32_t hidden, const double* uniforms, int64_t batch, int8_t* x, float* logq,
                   void* stream) {
  FK_REQUIRE(n_sites >= 2 && hidden >= 1, "fk_made_sample: bad sizes");
  if (batch <= 0) return FK_OK;
  FK_REQUIRE(w1 && b1 && w2t && b2 && w3 && b3 && uniforms && x && logq, "fk_made_sample: null pointer");
  const size_t lds = made_lds(n_sites, hidden);
  FK_REQUIRE(lds <= 64 * 1024, "fk_made_sample: too large for LDS");
  MadeW w{(const __bf16*)w1, b1, (const __bf16*)w2t, b2, (const __bf16*)w3, b3, n_sites, hidden};
  hipLaunchKernelGGL(made_sample_kernel, dim3(div_up(batch, MADE_S)), dim3(MADE_T), lds, (hipStream_t)stream, w,
                     uniforms, batch, x, logq);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_made_fit_workspace_floats(int32_t n_sites, int32_t hidden, int64_t batch) {
  return batch * (3LL * hidden + hidden + n_sites) + 2;
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
  const float bc1 = 1.0f - powf(beta1, (float)step), bc2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(made_adam_kernel, dim3(div_up(total, 256)), dim3(256), 0, s, master, adam_m, adam_v,
                     (__bf16*)w1, (__bf16*)w2t, (__bf16*)w3, x, g_h1, g_h2, g_dl, g_dh2, g_dh1, batch, n_sites,
                     hidden, lr, beta1, beta2, eps, bc1, bc2);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
