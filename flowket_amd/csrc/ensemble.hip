// Symmetrised wave functions (machines/ensemble.py:14-69): a lattice
// symmetry applied to a batch of configurations, and the probabilistic
// ensemble of K evaluations of log psi,
//   Re = 0.5 * logsumexp_k(2 Re_k) - 0.5 log K,
//   Im = arg(mean_k exp(i Im_k))                       (ensemble.py:14-20).
#include "fk_common.h"

namespace fk {

// out[b, q] = sign * in[b, src[q]]: one thread per output spin.
__global__ void __launch_bounds__(256) transform_spins_kernel(const int8_t* __restrict__ in, int64_t batch,
                                                              int32_t n_sites, const int32_t* __restrict__ src,
                                                              int32_t sign, int8_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * n_sites) return;
  const int64_t b = i / n_sites;
  const int q = (int)(i - b * n_sites);
  out[i] = (int8_t)(sign * in[b * n_sites + src[q]]);
}

// logpsi holds K blocks of `stride` rows (re, im); rows at or beyond the live
// count (*m_dev when given, else m) are left untouched.
__global__ void __launch_bounds__(256) ensemble_combine_kernel(const float* __restrict__ logpsi, int32_t k_members,
                                                               int64_t stride, int64_t m,
                                                               const int64_t* __restrict__ m_dev,
                                                               float* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t live = m_dev ? min(*m_dev, m) : m;
  if (r >= live) return;
  double mx = -INFINITY;
  for (int k = 0; k < k_members; ++k) mx = fmax(mx, 2.0 * (double)logpsi[2 * (k * stride + r)]);
  double s = 0.0, cs = 0.0, sn = 0.0;
  for (int k = 0; k < k_members; ++k) {
    const float2 z = *reinterpret_cast<const float2*>(logpsi + 2 * (k * stride + r));
    s += exp(2.0 * (double)z.x - mx);
    cs += cos((double)z.y);
    sn += sin((double)z.y);
  }
  const double re = 0.5 * (mx + log(s)) - 0.5 * log((double)k_members);
  out[2 * r] = (float)re;
  out[2 * r + 1] = (float)atan2(sn, cs);
}

}  // namespace fk

using namespace fk;

extern "C" {

int fk_transform_spins(const int8_t* in, int64_t batch, int32_t n_sites, const int32_t* src, int32_t sign,
                       int8_t* out, void* stream) {
  FK_REQUIRE(in && src && out, "fk_transform_spins: null pointer");
  FK_REQUIRE(batch >= 0 && n_sites > 0, "fk_transform_spins: bad sizes");
  FK_REQUIRE(sign == 1 || sign == -1, "fk_transform_spins: sign must be +1 or -1");
  FK_REQUIRE(in != out, "fk_transform_spins: in and out must not alias");
  const int64_t n = batch * n_sites;
  if (n == 0) return FK_OK;
  hipLaunchKernelGGL(transform_spins_kernel, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, in, batch,
                     n_sites, src, sign, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_ensemble_combine(const float* logpsi, int32_t k_members, int64_t stride, int64_t m, const int64_t* m_dev,
                        float* out, void* stream) {
  FK_REQUIRE(logpsi && out, "fk_ensemble_combine: null pointer");
  FK_REQUIRE(k_members >= 1 && m >= 0 && stride >= m, "fk_ensemble_combine: bad sizes");
  if (m == 0) return FK_OK;
  hipLaunchKernelGGL(ensemble_combine_kernel, dim3(div_up(m, 256)), dim3(256), 0, (hipStream_t)stream, logpsi,
                     k_members, stride, m, m_dev, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
