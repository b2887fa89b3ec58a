// Chain rule of weight normalisation (deepar/layers/wrappers.py:123-135):
// a wrapped layer's kernel is k = v / n * s per output channel c, with
// n = sqrt(max(sum_i v_i^2, 1e-12)) over the channel's kh*kw*cin values
// (tf.nn.l2_normalize) and s = exp(g) (exponential_norm) or g.  The gradient
// kernels of this library differentiate with respect to k; these kernels map
// such derivatives onto the trainable (v, g), and parameter directions
// (v', g') onto directions in k.  Small elementwise work: one thread per
// (row, layer, output channel).
#include "fk_common.h"

namespace fk {

// Layer table: 4 int64 per weight-normalised layer (kernel offset in the flat
// parameters, kh*kw*cin, cout, offset of its g in the g vector).
struct WnLayer {
  int64_t koff, nin, cout, goff;
};

__device__ __forceinline__ WnLayer wn_layer(const int64_t* t, int l) {
  return WnLayer{t[4 * l], t[4 * l + 1], t[4 * l + 2], t[4 * l + 3]};
}

// Norm and scale of channel c of layer L.
__device__ __forceinline__ void wn_norm(const float* v, const WnLayer& L, int64_t c, const float* g, int exp_norm,
                                        float& n, float& s, bool& clamped) {
  float ss = 0.0f;
  for (int64_t i = 0; i < L.nin; ++i) {
    const float x = v[L.koff + i * L.cout + c];
    ss += x * x;
  }
  clamped = ss < 1e-12f;
  n = sqrtf(clamped ? 1e-12f : ss);
  s = exp_norm ? expf(g[L.goff + c]) : g[L.goff + c];
}

// In place on rows (nrows, ld): row[:P] holds d/dk on entry; on exit the
// kernel entries of every wrapped layer hold d/dv and row[P + goff + c] holds
// d/dg.  dv_i = s/n (dk_i - v_i (v.dk)/n^2) (s/n dk_i when the norm is
// clamped), dg = s/n (v.dk) for exponential_norm, (v.dk)/n otherwise.
__global__ void wn_grad_rows_kernel(float* __restrict__ rows, int64_t nrows, int64_t ld,
                                    const float* __restrict__ v, const float* __restrict__ g,
                                    const int64_t* __restrict__ layers, int exp_norm, int64_t P) {
  const WnLayer L = wn_layer(layers, blockIdx.y);
  const int64_t per_block = (int64_t)blockDim.x;
  const int64_t idx = (int64_t)blockIdx.x * per_block + threadIdx.x;
  const int64_t r = idx / L.cout, c = idx - (idx / L.cout) * L.cout;
  if (r >= nrows) return;
  float n, s;
  bool clamped;
  wn_norm(v, L, c, g, exp_norm, n, s, clamped);
  float* row = rows + r * ld;
  float dot = 0.0f;
  for (int64_t i = 0; i < L.nin; ++i) dot += v[L.koff + i * L.cout + c] * row[L.koff + i * L.cout + c];
  const float sn = s / n;
  const float proj = clamped ? 0.0f : dot / (n * n);
  for (int64_t i = 0; i < L.nin; ++i) {
    const int64_t q = L.koff + i * L.cout + c;
    row[q] = sn * (row[q] - v[q] * proj);
  }
  row[P + L.goff + c] = exp_norm ? sn * dot : dot / n;
}

// t_eff (P) from a direction t over (v, g): kernel entries
// s/n (t_i - v_i (v.t)/n^2) + dk/dg t_g with dk/dg = k (exponential_norm) or
// v/n; every other entry (biases, the head) is t itself.
__global__ void wn_copy_kernel(const float* __restrict__ t, float* __restrict__ out, int64_t P) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < P) out[q] = t[q];
}

__global__ void wn_tangent_kernel(const float* __restrict__ v, const float* __restrict__ g,
                                  const int64_t* __restrict__ layers, int exp_norm, const float* __restrict__ t,
                                  float* __restrict__ out, int64_t P) {
  const WnLayer L = wn_layer(layers, blockIdx.y);
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.cout) return;
  float n, s;
  bool clamped;
  wn_norm(v, L, c, g, exp_norm, n, s, clamped);
  float dot = 0.0f;
  for (int64_t i = 0; i < L.nin; ++i) dot += v[L.koff + i * L.cout + c] * t[L.koff + i * L.cout + c];
  const float sn = s / n;
  const float proj = clamped ? 0.0f : dot / (n * n);
  const float tg = t[P + L.goff + c];
  for (int64_t i = 0; i < L.nin; ++i) {
    const int64_t q = L.koff + i * L.cout + c;
    const float dkdg = exp_norm ? v[q] * sn : v[q] / n;
    out[q] = sn * (t[q] - v[q] * proj) + dkdg * tg;
  }
}

}  // namespace fk

using namespace fk;

extern "C" {

int fk_wn_grad_rows(float* rows, int64_t nrows, int64_t ld, const float* params, const float* g,
                    const int64_t* layers, int32_t n_layers, int32_t max_cout, int32_t exp_norm, int64_t n_params,
                    void* stream) {
  FK_REQUIRE(nrows >= 0 && n_layers >= 0 && max_cout >= 1 && n_params >= 0, "fk_wn_grad_rows: bad sizes");
  if (nrows == 0 || n_layers == 0) return FK_OK;
  FK_REQUIRE(rows && params && g && layers, "fk_wn_grad_rows: null pointer");
  FK_REQUIRE(ld >= n_params, "fk_wn_grad_rows: row stride shorter than the parameters");
  FK_REQUIRE(n_layers <= 65535, "fk_wn_grad_rows: too many layers");
  const int64_t threads = nrows * (int64_t)max_cout;
  hipLaunchKernelGGL(wn_grad_rows_kernel, dim3((unsigned)div_up(threads, 256), (unsigned)n_layers), dim3(256), 0,
                     (hipStream_t)stream, rows, nrows, ld, params, g, layers, exp_norm, n_params);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_wn_tangent(const float* params, const float* g, const int64_t* layers, int32_t n_layers, int32_t max_cout,
                  int32_t exp_norm, const float* t, float* t_eff, int64_t n_params, void* stream) {
  FK_REQUIRE(n_layers >= 0 && max_cout >= 1 && n_params >= 0, "fk_wn_tangent: bad sizes");
  FK_REQUIRE(params && t && t_eff && (n_layers == 0 || (g && layers)), "fk_wn_tangent: null pointer");
  FK_REQUIRE(n_layers <= 65535, "fk_wn_tangent: too many layers");
  hipStream_t s = (hipStream_t)stream;
  if (n_params > 0) {
    hipLaunchKernelGGL(wn_copy_kernel, dim3(div_up(n_params, 256)), dim3(256), 0, s, t, t_eff, n_params);
    FK_CHECK_LAUNCH();
  }
  if (n_layers == 0) return FK_OK;
  hipLaunchKernelGGL(wn_tangent_kernel, dim3((unsigned)div_up(max_cout, 64), (unsigned)n_layers), dim3(64), 0, s,
                     params, g, layers, exp_norm, t, t_eff, n_params);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
