// Exact enumeration path: every one of the 2^N spin configurations, their
// normalised probabilities, exact local energies by index gathers (a
// connected configuration of an enumerated state is itself enumerated, so no
// extra network evaluations), the exact energy, variance and gradient
// coefficients, and categorical sampling from |psi|^2
// (optimization/exact_variational.py:10-161, samplers/exact_sampler.py:7-20,
// exact/utils.py:39-84).
//
// State index k <-> configuration: site i holds +1 when bit i of k is set,
// else -1 (exact/utils.py:49-58, little-endian), so flipping site a is
// k ^ (1 << a) and exchanging antiparallel sites a, b is k ^ (1 << a) ^ (1 << b).
//
// Reductions run over a fixed grid with grid-stride loops and a final
// single-workgroup pass in fixed order, so every result is deterministic.
#include "fk_common.h"

namespace fk {

constexpr int EX_BLOCKS = 256, EX_T = 256;

__device__ __forceinline__ double block_sum(double v, double* sm) {
  const int tid = threadIdx.x;
  sm[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) sm[tid] += sm[tid + s];
    __syncthreads();
  }
  const double r = sm[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_max(double v, double* sm) {
  const int tid = threadIdx.x;
  sm[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) sm[tid] = fmax(sm[tid], sm[tid + s]);
    __syncthreads();
  }
  const double r = sm[0];
  __syncthreads();
  return r;
}

// states[k - first][i] = 2 * bit_i(k) - 1
__global__ void exact_states_kernel(int n_sites, int64_t first, int64_t count, int8_t* __restrict__ states) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * n_sites) return;
  const int64_t k = first + t / n_sites;
  const int i = (int)(t % n_sites);
  states[t] = ((k >> i) & 1) ? (int8_t)1 : (int8_t)-1;
}

// ---- normalisation (exact_variational.py:127-133, exact/utils.py:79-82):
// log |psi|^2 = 2 Re log psi; m = max; log Z = log sum exp(2 Re log psi - m)
// + m; p = exp(2 Re log psi - log Z).
__global__ __launch_bounds__(EX_T) void exact_max_kernel(const float* __restrict__ logpsi, int64_t n,
                                                         double* __restrict__ part) {
  __shared__ double sm[EX_T];
  double m = -INFINITY;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, 2.0 * (double)logpsi[2 * k]);
  m = block_max(m, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(EX_T) void exact_sumexp_kernel(const float* __restrict__ logpsi, int64_t n,
                                                            const double* __restrict__ mpart, int nparts,
                                                            double* __restrict__ part) {
  __shared__ double sm[EX_T];
  double m = -INFINITY;
  for (int i = 0; i < nparts; ++i) m = fmax(m, mpart[i]);
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    s += exp(2.0 * (double)logpsi[2 * k] - m);
  s = block_sum(s, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// scalars: [0] log Z, [1] max of 2 Re log psi
__global__ __launch_bounds__(EX_T) void exact_logz_kernel(const double* __restrict__ mpart,
                                                          const double* __restrict__ spart, int nparts,
                                                          double* __restrict__ scalars) {
  __shared__ double sm[EX_T];
  double m = -INFINITY;
  for (int i = 0; i < nparts; ++i) m = fmax(m, mpart[i]);
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += spart[i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    scalars[0] = log(s) + m;
    scalars[1] = m;
  }
}

__global__ void exact_probs_kernel(const float* __restrict__ logpsi, int64_t n, const double* __restrict__ scalars,
                                   double* __restrict__ probs) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) probs[k] = exp(2.0 * (double)logpsi[2 * k] - scalars[0]);
}

// ---- local energies (exact_variational.py:80-95, the "naive" energies):
// E_loc(k) = sum over used slots c of mel[c][k] * exp(log psi(k') - log psi(k)),
// complex, in FP64 from the FP32 log psi values.
__global__ void exact_local_energy_kernel(const float* __restrict__ logpsi, int64_t n, int kind,
                                          const int32_t* __restrict__ conn_a, const int32_t* __restrict__ conn_b,
                                          const double* __restrict__ mel, const uint8_t* __restrict__ use, int n_conn,
                                          double* __restrict__ eloc) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double r0 = logpsi[2 * k], i0 = logpsi[2 * k + 1];
  double er = 0.0, ei = 0.0;
  for (int c = 0; c < n_conn; ++c) {
    if (!use[(int64_t)c * n + k]) continue;
    int64_t kk = k;
    if (c > 0) {
      const int a = conn_a[c];
      if (kind == FK_CONN_FLIP) {
        kk ^= (int64_t)1 << a;
      } else {
        kk ^= ((int64_t)1 << a) ^ ((int64_t)1 << conn_b[c]);
      }
    }
    const double dr = (double)logpsi[2 * kk] - r0, di = (double)logpsi[2 * kk + 1] - i0;
    const double m = mel[(int64_t)c * n + k] * exp(dr);
    er += m * cos(di);
    ei += m * sin(di);
  }
  eloc[2 * k] = er;
  eloc[2 * k + 1] = ei;
}

// ---- energy, variance and gradient coefficients (exact_variational.py:
// 86-100,135-138): E = sum_k p_k E_loc(k); var = sum_k p_k (Re E_loc(k) -
// Re E)^2; coefficient k = p_k conj(E_loc(k) - E), the exact counterpart of
// the sampled conj(E_loc - E) / B.
__global__ __launch_bounds__(EX_T) void exact_mean_kernel(const double* __restrict__ probs,
                                                          const double* __restrict__ eloc, int64_t n,
                                                          double* __restrict__ part) {
  __shared__ double sm[EX_T];
  double r = 0.0, i = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    r += probs[k] * eloc[2 * k];
    i += probs[k] * eloc[2 * k + 1];
  }
  r = block_sum(r, sm);
  i = block_sum(i, sm);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r;
    part[2 * blockIdx.x + 1] = i;
  }
}

__global__ __launch_bounds__(EX_T) void exact_var_kernel(const double* __restrict__ probs,
                                                         const double* __restrict__ eloc, int64_t n,
                                                         const double* __restrict__ mean, double* __restrict__ coef,
                                                         double* __restrict__ part) {
  __shared__ double sm[EX_T];
  const double er = mean[0], ei = mean[1];
  double v = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const double dr = eloc[2 * k] - er, di = eloc[2 * k + 1] - ei;
    v += probs[k] * dr * dr;
    if (coef) {
      coef[2 * k] = probs[k] * dr;
      coef[2 * k + 1] = -probs[k] * di;
    }
  }
  v = block_sum(v, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// moments: [0] Re E, [1] Im E, [2] variance
__global__ __launch_bounds__(EX_T) void exact_combine_kernel(const double* __restrict__ part, int nparts, int width,
                                                             double* __restrict__ out) {
  __shared__ double sm[EX_T];
  for (int w = 0; w < width; ++w) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i * width + w];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) out[w] = s;
  }
}

// ---- categorical sampling (samplers/exact_sampler.py:14-20, numpy's
// choice: cdf = cumsum(p), cdf /= cdf[-1], index = searchsorted(cdf, u,
// 'right')).  The prefix sum runs in one workgroup: per-thread runs, a
// scan of the run totals, then the runs again.
__global__ __launch_bounds__(1024) void exact_cdf_kernel(const double* __restrict__ probs, int64_t n,
                                                         double* __restrict__ cdf) {
  __shared__ double sm[1024];
  const int tid = threadIdx.x, T = blockDim.x;
  const int64_t per = (n + T - 1) / T;
  const int64_t k0 = tid * per, k1 = (k0 + per < n) ? k0 + per : n;
  double s = 0.0;
  for (int64_t k = k0; k < k1; ++k) s += probs[k];
  sm[tid] = s;
  __syncthreads();
  if (tid == 0) {
    double run = 0.0;
    for (int t = 0; t < T; ++t) {
      const double x = sm[t];
      sm[t] = run;
      run += x;
    }
  }
  __syncthreads();
  double run = sm[tid];
  for (int64_t k = k0; k < k1; ++k) {
    run += probs[k];
    cdf[k] = run;
  }
}

__global__ void exact_draw_kernel(const double* __restrict__ cdf, int64_t n, int n_sites,
                                  const double* __restrict__ u, int64_t batch, int8_t* __restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const double x = u[b] * cdf[n - 1];
  // first index with cdf > x
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (cdf[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  for (int i = 0; i < n_sites; ++i) out[b * n_sites + i] = ((lo >> i) & 1) ? (int8_t)1 : (int8_t)-1;
}

}  // namespace fk

using namespace fk;

extern "C" {

int fk_exact_states(int32_t n_sites, int64_t first, int64_t count, int8_t* states, void* stream) {
  FK_REQUIRE(n_sites >= 1 && n_sites <= 40 && first >= 0 && count >= 0, "fk_exact_states: bad sizes");
  FK_REQUIRE(first + count <= ((int64_t)1 << n_sites), "fk_exact_states: state range beyond 2^n_sites");
  if (count == 0) return FK_OK;
  FK_REQUIRE(states, "fk_exact_states: null pointer");
  const int64_t total = count * n_sites;
  hipLaunchKernelGGL(exact_states_kernel, dim3((unsigned)div_up(total, 256)), dim3(256), 0, (hipStream_t)stream,
                     n_sites, first, count, states);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_exact_workspace_doubles(int64_t n_states) { return 4 * (int64_t)EX_BLOCKS + 8 + (n_states > 0 ? n_states : 0); }

int fk_exact_normalize(const float* logpsi, int64_t n_states, double* probs, double* scalars, double* workspace,
                       void* stream) {
  FK_REQUIRE(n_states >= 1, "fk_exact_normalize: bad sizes");
  FK_REQUIRE(logpsi && probs && scalars && workspace, "fk_exact_normalize: null pointer");
  hipStream_t s = (hipStream_t)stream;
  double* mpart = workspace;
  double* spart = workspace + EX_BLOCKS;
  hipLaunchKernelGGL(exact_max_kernel, dim3(EX_BLOCKS), dim3(EX_T), 0, s, logpsi, n_states, mpart);
  hipLaunchKernelGGL(exact_sumexp_kernel, dim3(EX_BLOCKS), dim3(EX_T), 0, s, logpsi, n_states, mpart, EX_BLOCKS,
                     spart);
  hipLaunchKernelGGL(exact_logz_kernel, dim3(1), dim3(EX_T), 0, s, mpart, spart, EX_BLOCKS, scalars);
  hipLaunchKernelGGL(exact_probs_kernel, dim3((unsigned)div_up(n_states, 256)), dim3(256), 0, s, logpsi, n_states,
                     scalars, probs);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_exact_local_energy(const float* logpsi, int64_t n_states, int32_t n_sites, int32_t kind,
                          const int32_t* conn_a, const int32_t* conn_b, const double* mel, const uint8_t* use,
                          int32_t n_conn, double* eloc, void* stream) {
  FK_REQUIRE(n_sites >= 1 && n_sites <= 40 && n_states == ((int64_t)1 << n_sites) && n_conn >= 1,
             "fk_exact_local_energy: n_states must be 2^n_sites");
  FK_REQUIRE(kind == FK_CONN_SWAP || kind == FK_CONN_SWAP_ZERO_PAD || kind == FK_CONN_FLIP,
             "fk_exact_local_energy: bad kind %d", kind);
  FK_REQUIRE(logpsi && conn_a && conn_b && mel && use && eloc, "fk_exact_local_energy: null pointer");
  hipLaunchKernelGGL(exact_local_energy_kernel, dim3((unsigned)div_up(n_states, 256)), dim3(256), 0,
                     (hipStream_t)stream, logpsi, n_states, kind, conn_a, conn_b, mel, use, n_conn, eloc);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_exact_energy(const double* probs, const double* eloc, int64_t n_states, double* coef, double* moments,
                    double* workspace, void* stream) {
  FK_REQUIRE(n_states >= 1, "fk_exact_energy: bad sizes");
  FK_REQUIRE(probs && eloc && moments && workspace, "fk_exact_energy: null pointer");
  hipStream_t s = (hipStream_t)stream;
  double* mpart = workspace;                  // 2 * EX_BLOCKS
  double* vpart = workspace + 2 * EX_BLOCKS;  // EX_BLOCKS
  hipLaunchKernelGGL(exact_mean_kernel, dim3(EX_BLOCKS), dim3(EX_T), 0, s, probs, eloc, n_states, mpart);
  hipLaunchKernelGGL(exact_combine_kernel, dim3(1), dim3(EX_T), 0, s, mpart, EX_BLOCKS, 2, moments);
  hipLaunchKernelGGL(exact_var_kernel, dim3(EX_BLOCKS), dim3(EX_T), 0, s, probs, eloc, n_states, moments, coef,
                     vpart);
  hipLaunchKernelGGL(exact_combine_kernel, dim3(1), dim3(EX_T), 0, s, vpart, EX_BLOCKS, 1, moments + 2);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_exact_sample(const double* probs, int64_t n_states, int32_t n_sites, const double* uniforms, int64_t batch,
                    int8_t* out, double* workspace, void* stream) {
  FK_REQUIRE(n_sites >= 1 && n_sites <= 40 && n_states == ((int64_t)1 << n_sites) && batch >= 0,
             "fk_exact_sample: n_states must be 2^n_sites");
  if (batch == 0) return FK_OK;
  FK_REQUIRE(probs && uniforms && out && workspace, "fk_exact_sample: null pointer");
  hipStream_t s = (hipStream_t)stream;
  double* cdf = workspace + 4 * EX_BLOCKS + 8;
  hipLaunchKernelGGL(exact_cdf_kernel, dim3(1), dim3(1024), 0, s, probs, n_states, cdf);
  hipLaunchKernelGGL(exact_draw_kernel, dim3((unsigned)div_up(batch, 256)), dim3(256), 0, s, cdf, n_states, n_sites,
                     uniforms, batch, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
