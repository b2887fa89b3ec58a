// Complex RBM wave function and Metropolis-Hastings chains.
//
// RBM (machines/rbm.py:35-48, layers/complex/dense.py:11-50,
// layers/complex/tensorflow_ops.py:69-85):
//   log psi(s) = sum_j a_j s_j + sum_h lncosh(theta_h),  theta_h = b_h + sum_j W_jh s_j,
// complex parameters stored as (real, imag) pairs and used as real - 1j*imag
// (layers/complex/base_layer.py:30-35).  One wavefront per configuration, the
// hidden units spread over the lanes.
//
// MH (samplers/metropolis_hastings.py:15-145): one wavefront per chain keeps
// theta in registers and updates it incrementally for each accepted move
// (a flipped site s_j changes theta by -2 s_j W_j.), so a proposal costs
// O(M) instead of a full O(N M) evaluation.  Acceptance is the reference rule
// exp(min(2 Re(logpsi' - logpsi), 0)) > U.  Random numbers are either
// injected (draw-for-draw parity) or drawn from a counter-based hash keyed by
// (seed, chain, step).
#include "fk_common.h"

namespace fk {

struct RbmParams {
  const float *w_re, *w_im, *b_re, *b_im, *a_re, *a_im;
  int N, M;
};

// Re and Im of lncosh(x + iy) (tensorflow_ops.py:79-85): |x| - log 2 +
// log(e^{z-|x|} + e^{-z-|x|}); the real part is written in the overflow-free
// form |x| - log 2 + 0.5 log1p(e^{-4|x|} + 2 e^{-2|x|} cos 2y).
__device__ __forceinline__ float lncosh_re(float x, float y) {
  const float ax = fabsf(x);
  const float e2 = expf(-2.0f * ax);
  return ax - 0.69314718055994531f + 0.5f * log1pf(e2 * e2 + 2.0f * e2 * cosf(2.0f * y));
}
__device__ __forceinline__ float lncosh_im(float x, float y) {
  // arg(e^{z-|x|} + e^{-z-|x|}) = atan2(Im, Re)
  const float ax = fabsf(x);
  const float e1 = expf(x - ax), e2 = expf(-x - ax);
  float sy, cy;
  sincosf(y, &sy, &cy);
  return atan2f((e1 - e2) * sy, (e1 + e2) * cy);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// log psi of configurations: rows of `spins` or (sample, connection) pairs
// applied with the connection table (kind as fk_find_conn).
__global__ __launch_bounds__(256) void rbm_log_psi_kernel(RbmParams p, const int8_t* __restrict__ spins,
                                                          const int32_t* __restrict__ pairs,
                                                          const int64_t* __restrict__ m_dev, int64_t m_count,
                                                          int conn_kind, const int32_t* __restrict__ conn_a,
                                                          const int32_t* __restrict__ conn_b,
                                                          float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int64_t M = m_dev ? *m_dev : m_count;
  if (M > m_count) M = m_count;
  if (m >= M) return;
  int64_t b = m;
  int c = 0;
  if (pairs) {
    b = pairs[2 * m];
    c = pairs[2 * m + 1];
  }
  const int8_t* row = spins + b * p.N;
  int fa = -1, fb = -1;
  if (c > 0) {
    fa = conn_a[c];
    fb = (conn_kind == FK_CONN_FLIP) ? -1 : conn_b[c];
  }
  auto spin = [&](int j) -> float {
    int8_t v = row[j];
    if (conn_kind == FK_CONN_FLIP) {
      if (j == fa) v = (int8_t)-v;
    } else if (fa >= 0) {
      if (j == fa) v = row[fb];
      else if (j == fb) v = row[fa];
    }
    return (float)v;
  };
  float lr = 0.0f, li = 0.0f;
  for (int h = lane; h < p.M; h += 64) {
    float tr = p.b_re[h], ti = -p.b_im[h];
    for (int j = 0; j < p.N; ++j) {
      const float s = spin(j);
      tr += p.w_re[(int64_t)j * p.M + h] * s;
      ti -= p.w_im[(int64_t)j * p.M + h] * s;
    }
    lr += lncosh_re(tr, ti);
    li += lncosh_im(tr, ti);
  }
  for (int j = lane; j < p.N; j += 64) {
    const float s = spin(j);
    lr += p.a_re[j] * s;
    li -= p.a_im[j] * s;
  }
  lr = wave_sum(lr);
  li = wave_sum(li);
  if (lane == 0) {
    out[2 * m] = lr;
    out[2 * m + 1] = li;
  }
}

// counter-based uniform in [0, 1) (splitmix64 finaliser of the counter)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t chain, uint64_t step, uint64_t k) {
  const uint64_t r = mix64(seed ^ mix64(chain * 0x100000001B3ull ^ mix64(step * 4 + k)));
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

struct MhArgs {
  RbmParams p;
  int H, W;              // lattice (N = H*W)
  int kind;              // 0 local flip, 1 exchange
  int64_t chains;
  int64_t n_burn;        // sweeps before the first record
  int64_t n_record;      // recorded samples per chain
  int64_t unused;        // sweeps between records are unused + 1
  int8_t* state;         // (chains, N) in/out
  int8_t* out;           // (chains, n_record, N)
  int64_t* accepts;      // (chains) accepted proposals during the recorded sweeps
  float* logpsi;         // (chains, 2) final log psi
  // randomness: injected (sweeps, chains) arrays or hashed from seed
  const int32_t* r_site;   // local: flat site; exchange: flat site
  const int32_t* r_move;   // exchange: (dy+1)*3 + (dx+1)
  const double* r_u;       // acceptance uniforms
  uint64_t seed;
  int32_t* nonfinite;      // set to 1 when a proposal's log psi is not finite (or null)
};

template <int HPL>  // hidden units per lane (M <= 64 * HPL)
__global__ __launch_bounds__(64) void mh_rbm_kernel(MhArgs a) {
  const int lane = threadIdx.x;
  const int64_t ch = blockIdx.x;
  if (ch >= a.chains) return;
  const RbmParams& p = a.p;
  const int N = p.N, M = p.M;
  extern __shared__ int8_t s_state[];  // N spins of this chain
  for (int j = lane; j < N; j += 64) s_state[j] = a.state[ch * N + j];
  __syncthreads();
  float tr[HPL], ti[HPL];
#pragma unroll
  for (int k = 0; k < HPL; ++k) {
    const int h = lane + 64 * k;
    float r = 0.0f, i = 0.0f;
    if (h < M) {
      r = p.b_re[h];
      i = -p.b_im[h];
      for (int j = 0; j < N; ++j) {
        const float s = (float)s_state[j];
        r += p.w_re[(int64_t)j * M + h] * s;
        i -= p.w_im[(int64_t)j * M + h] * s;
      }
    }
    tr[k] = r;
    ti[k] = i;
  }
  // Re lncosh(theta) of the current state, kept with theta (recomputed
  // only where theta changes)
  float lc[HPL];
#pragma unroll
  for (int k = 0; k < HPL; ++k) lc[k] = (lane + 64 * k < M) ? lncosh_re(tr[k], ti[k]) : 0.0f;
  const int64_t total = a.n_burn + a.n_record * (a.unused + 1);
  int64_t acc = 0;
  // proposal of a step: sites (j2 = -1 for local moves) and uniform, from
  // the injected arrays or the counter hash; it does not depend on earlier
  // acceptances, so the next step's is drawn and its weight rows and
  // visible biases are requested while the current step computes
  int cj1 = 0, cj2 = -1;
  double cu = 0.0;
  float pw1r[HPL], pw1i[HPL], pw2r[HPL], pw2i[HPL];
  float pa1 = 0.0f, pa2 = 0.0f;
  auto prefetch = [&](int64_t st) {
    int j1, j2 = -1;
    double u;
    if (a.r_u) {
      j1 = a.r_site[st * a.chains + ch];
      u = a.r_u[st * a.chains + ch];
      if (a.kind == 1) {
        const int mv = a.r_move[st * a.chains + ch];
        const int dy = mv / 3 - 1, dx = mv % 3 - 1;
        const int y = j1 / a.W, x = j1 - (j1 / a.W) * a.W;
        j2 = ((y + dy + a.H) % a.H) * a.W + ((x + dx + a.W) % a.W);
      }
    } else {
      j1 = (int)(u01(a.seed, ch, st, 0) * N);
      if (j1 >= N) j1 = N - 1;
      u = u01(a.seed, ch, st, 1);
      if (a.kind == 1) {
        int mv = (int)(u01(a.seed, ch, st, 2) * 9.0);
        if (mv > 8) mv = 8;
        const int dy = mv / 3 - 1, dx = mv % 3 - 1;
        const int y = j1 / a.W, x = j1 - (j1 / a.W) * a.W;
        j2 = ((y + dy + a.H) % a.H) * a.W + ((x + dx + a.W) % a.W);
      }
    }
    cj1 = j1;
    cj2 = j2;
    cu = u;
    const int jb = j2 >= 0 ? j2 : j1;
#pragma unroll
    for (int k = 0; k < HPL; ++k) {
      const int h = lane + 64 * k;
      pw1r[k] = h < M ? p.w_re[(int64_t)j1 * M + h] : 0.0f;
      pw1i[k] = h < M ? p.w_im[(int64_t)j1 * M + h] : 0.0f;
      pw2r[k] = h < M ? p.w_re[(int64_t)jb * M + h] : 0.0f;
      pw2i[k] = h < M ? p.w_im[(int64_t)jb * M + h] : 0.0f;
    }
    pa1 = p.a_re[j1];
    pa2 = p.a_re[jb];
  };
  if (total > 0) prefetch(0);
  for (int64_t step = 0; step < total; ++step) {
    if (step > 0 && (step & 255) == 0) {
      // refresh theta from the state (bounds the drift of the FP32
      // incremental updates)
#pragma unroll
      for (int k = 0; k < HPL; ++k) {
        const int h = lane + 64 * k;
        if (h < M) {
          float r = p.b_re[h], i = -p.b_im[h];
          for (int j = 0; j < N; ++j) {
            const float s = (float)s_state[j];
            r += p.w_re[(int64_t)j * M + h] * s;
            i -= p.w_im[(int64_t)j * M + h] * s;
          }
          tr[k] = r;
          ti[k] = i;
          lc[k] = lncosh_re(r, i);
        }
      }
    }
    // ---- proposal: drawn (and its weight rows requested) one step ahead
    const int j1 = cj1, j2 = cj2;
    const double u = cu;
    float w1r[HPL], w1i[HPL], w2r[HPL], w2i[HPL];
#pragma unroll
    for (int k = 0; k < HPL; ++k) {
      w1r[k] = pw1r[k];
      w1i[k] = pw1i[k];
      w2r[k] = pw2r[k];
      w2i[k] = pw2i[k];
    }
    const float a1 = pa1, a2 = pa2;
    if (step + 1 < total) prefetch(step + 1);
    // sites that change: local -> j1; exchange -> j1 and j2 if they differ
    const float s1 = (float)s_state[j1];
    bool flip1 = true, flip2 = false;
    if (a.kind == 1) {
      const float s2 = (float)s_state[j2];
      flip1 = (j2 != j1) && (s1 != s2);
      flip2 = flip1;
    }
    float dre = 0.0f;
    float nr[HPL], ni[HPL], nl[HPL];
    if (flip1) {
      const float s2 = flip2 ? (float)s_state[j2] : 0.0f;
#pragma unroll
      for (int k = 0; k < HPL; ++k) {
        const int h = lane + 64 * k;
        nr[k] = tr[k];
        ni[k] = ti[k];
        if (h < M) {
          nr[k] -= 2.0f * s1 * w1r[k];
          ni[k] += 2.0f * s1 * w1i[k];
          if (flip2) {
            nr[k] -= 2.0f * s2 * w2r[k];
            ni[k] += 2.0f * s2 * w2i[k];
          }
          nl[k] = lncosh_re(nr[k], ni[k]);
          dre += nl[k] - lc[k];
        }
      }
      dre = wave_sum(dre);
      dre += -2.0f * s1 * a1;
      if (flip2) dre += -2.0f * (float)s_state[j2] * a2;
    }
    // ---- acceptance (samplers/metropolis_hastings.py:106-110); a
    // candidate whose log psi is not finite is reported (the reference
    // raises, :103-105) and never accepted
    const bool finite = isfinite(dre);
    if (!finite && a.nonfinite && lane == 0) a.nonfinite[0] = 1;
    const float lr = fminf(2.0f * dre, 0.0f);
    const bool accept = finite && (double)expf(lr) > u;
    if (accept) {
      if (flip1) {
#pragma unroll
        for (int k = 0; k < HPL; ++k) {
          tr[k] = nr[k];
          ti[k] = ni[k];
          if (lane + 64 * k < M) lc[k] = nl[k];
        }
        __syncthreads();
        if (lane == 0) {
          if (flip2) {
            const int8_t t = s_state[j1];
            s_state[j1] = s_state[j2];
            s_state[j2] = t;
          } else {
            s_state[j1] = (int8_t)-s_state[j1];
          }
        }
        __syncthreads();
      }
      if (step >= a.n_burn) ++acc;
    }
    // ---- record
    if (step >= a.n_burn) {
      const int64_t rs = step - a.n_burn;
      if ((rs + 1) % (a.unused + 1) == 0) {
        const int64_t rec = rs / (a.unused + 1);
        for (int j = lane; j < N; j += 64) a.out[(ch * a.n_record + rec) * N + j] = s_state[j];
      }
    }
  }
  for (int j = lane; j < N; j += 64) a.state[ch * N + j] = s_state[j];
  // final log psi from theta
  float lr = 0.0f, li = 0.0f;
#pragma unroll
  for (int k = 0; k < HPL; ++k) {
    if (lane + 64 * k < M) {
      lr += lncosh_re(tr[k], ti[k]);
      li += lncosh_im(tr[k], ti[k]);
    }
  }
  for (int j = lane; j < N; j += 64) {
    const float s = (float)s_state[j];
    lr += p.a_re[j] * s;
    li -= p.a_im[j] * s;
  }
  lr = wave_sum(lr);
  li = wave_sum(li);
  if (lane == 0) {
    a.accepts[ch] = acc;
    if (a.logpsi) {
      a.logpsi[2 * ch] = lr;
      a.logpsi[2 * ch + 1] = li;
    }
  }
}

// ---- machine-agnostic MH pieces (any wave function evaluated elsewhere)
__global__ void mh_propose_kernel(const int8_t* __restrict__ state, int64_t chains, int H, int W, int kind,
                                  const int32_t* __restrict__ r_site, const int32_t* __restrict__ r_move,
                                  int8_t* __restrict__ cand) {
  const int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ch >= chains) return;
  const int N = H * W;
  const int8_t* s = state + ch * N;
  int8_t* c = cand + ch * N;
  for (int j = 0; j < N; ++j) c[j] = s[j];
  const int j1 = r_site[ch];
  if (kind == 0) {
    c[j1] = (int8_t)-c[j1];
  } else {
    const int mv = r_move[ch];
    const int dy = mv / 3 - 1, dx = mv % 3 - 1;
    const int y = j1 / W, x = j1 - (j1 / W) * W;
    const int j2 = ((y + dy + H) % H) * W + ((x + dx + W) % W);
    const int8_t t = c[j1];
    c[j1] = c[j2];
    c[j2] = t;
  }
}

__global__ void mh_accept_kernel(int8_t* __restrict__ state, float* __restrict__ logpsi_cur,
                                 const int8_t* __restrict__ cand, const float* __restrict__ logpsi_cand,
                                 const double* __restrict__ u, int64_t chains, int N, int32_t* __restrict__ accepted,
                                 int32_t* __restrict__ nonfinite) {
  const int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ch >= chains) return;
  const bool finite = isfinite(logpsi_cand[2 * ch]) && isfinite(logpsi_cand[2 * ch + 1]);
  if (!finite && nonfinite) nonfinite[0] = 1;
  const float lr = fminf(2.0f * (logpsi_cand[2 * ch] - logpsi_cur[2 * ch]), 0.0f);
  const bool acc = finite && (double)expf(lr) > u[ch];
  accepted[ch] = acc ? 1 : 0;
  if (acc) {
    for (int j = 0; j < N; ++j) state[ch * N + j] = cand[ch * N + j];
    logpsi_cur[2 * ch] = logpsi_cand[2 * ch];
    logpsi_cur[2 * ch + 1] = logpsi_cand[2 * ch + 1];
  }
}

// Independent proposal with the exact Metropolis-Hastings ratio:
// accept iff exp(min(2 Re(logpsi' - logpsi) + log q(x) - log q(x'), 0)) > U.
__global__ void mh_accept_independent_kernel(int8_t* __restrict__ state, float* __restrict__ logpsi_cur,
                                             float* __restrict__ logq_cur, const int8_t* __restrict__ cand,
                                             const float* __restrict__ logpsi_cand, const float* __restrict__ logq_cand,
                                             const double* __restrict__ u, int64_t chains, int N,
                                             int32_t* __restrict__ accepted, int32_t* __restrict__ nonfinite) {
  const int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ch >= chains) return;
  const bool finite = isfinite(logpsi_cand[2 * ch]) && isfinite(logpsi_cand[2 * ch + 1]);
  if (!finite && nonfinite) nonfinite[0] = 1;
  const float lr = fminf(2.0f * (logpsi_cand[2 * ch] - logpsi_cur[2 * ch]) + (logq_cur[ch] - logq_cand[ch]), 0.0f);
  const bool acc = finite && (double)expf(lr) > u[ch];
  accepted[ch] = acc ? 1 : 0;
  if (acc) {
    for (int j = 0; j < N; ++j) state[ch * N + j] = cand[ch * N + j];
    logpsi_cur[2 * ch] = logpsi_cand[2 * ch];
    logpsi_cur[2 * ch + 1] = logpsi_cand[2 * ch + 1];
    logq_cur[ch] = logq_cand[ch];
  }
}

// Per-sample O = d log psi / d(parameter) of the RBM (the Jacobian the
// reference's ComplexValuesOptimizer takes of complex layers,
// optimizers/complex_values_optimizer.py:60-76, written over the stored real
// and imaginary parts): with theta_h = b_h + sum_j W_jh s_j and
// T_h = tanh(theta_h), d/dW_re[j,h] = s_j T_h, d/dW_im[j,h] = -i s_j T_h,
// d/db_re[h] = T_h, d/db_im[h] = -i T_h, d/da_re[j] = s_j, d/da_im[j] = -i s_j
// (every complex parameter is used as real - i*imag, base_layer.py:30-35).
// One workgroup per sample; rows in the flat order W_re, W_im, b_re, b_im,
// a_re, a_im.
__global__ __launch_bounds__(256) void rbm_grad_kernel(RbmParams p, const int8_t* __restrict__ spins, int64_t batch,
                                                       float* __restrict__ o_re, float* __restrict__ o_im,
                                                       int64_t ld) {
  extern __shared__ float s_t[];  // T_h (re, im), then the spins
  const int N = p.N, M = p.M;
  const int64_t b = blockIdx.x;
  if (b >= batch) return;
  float* tre = s_t;
  float* tim = s_t + M;
  float* sv = s_t + 2 * M;
  for (int j = threadIdx.x; j < N; j += blockDim.x) sv[j] = (float)spins[b * N + j];
  __syncthreads();
  for (int h = threadIdx.x; h < M; h += blockDim.x) {
    float x = p.b_re[h], y = -p.b_im[h];
    for (int j = 0; j < N; ++j) {
      x += p.w_re[(int64_t)j * M + h] * sv[j];
      y -= p.w_im[(int64_t)j * M + h] * sv[j];
    }
    // tanh(x + iy) = (sinh 2x + i sin 2y) / (cosh 2x + cos 2y), written with
    // e = exp(-2|x|) so that it cannot overflow
    const float e = expf(-2.0f * fabsf(x));
    float sn, cs;
    sincosf(2.0f * y, &sn, &cs);
    const float den = 1.0f + e * e + 2.0f * e * cs;
    tre[h] = copysignf((1.0f - e * e) / den, x);
    tim[h] = 2.0f * e * sn / den;
  }
  __syncthreads();
  float* ore = o_re + b * ld;
  float* oim = o_im + b * ld;
  const int64_t NM = (int64_t)N * M;
  for (int64_t q = threadIdx.x; q < NM; q += blockDim.x) {
    const int j = (int)(q / M), h = (int)(q - (q / M) * M);
    const float sj = sv[j];
    ore[q] = sj * tre[h];
    oim[q] = sj * tim[h];
    ore[NM + q] = sj * tim[h];
    oim[NM + q] = -sj * tre[h];
  }
  for (int h = threadIdx.x; h < M; h += blockDim.x) {
    ore[2 * NM + h] = tre[h];
    oim[2 * NM + h] = tim[h];
    ore[2 * NM + M + h] = tim[h];
    oim[2 * NM + M + h] = -tre[h];
  }
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    ore[2 * NM + 2 * M + j] = sv[j];
    oim[2 * NM + 2 * M + j] = 0.0f;
    ore[2 * NM + 2 * M + N + j] = 0.0f;
    oim[2 * NM + 2 * M + N + j] = -sv[j];
  }
}

}  // namespace fk

using namespace fk;

extern "C" {

int fk_rbm_log_psi(const float* w_re, const float* w_im, const float* b_re, const float* b_im, const float* a_re,
                   const float* a_im, int32_t n_visible, int32_t n_hidden, const int8_t* spins,
                   const int32_t* pairs, const int64_t* m_dev, int64_t m_count, int32_t conn_kind,
                   const int32_t* conn_a, const int32_t* conn_b, float* logpsi, void* stream) {
  FK_REQUIRE(n_visible > 0 && n_hidden > 0, "fk_rbm_log_psi: bad sizes");
  if (m_count <= 0) return FK_OK;
  FK_REQUIRE(w_re && w_im && b_re && b_im && a_re && a_im && spins && logpsi, "fk_rbm_log_psi: null pointer");
  FK_REQUIRE(!pairs || (conn_a && conn_b), "fk_rbm_log_psi: pairs need a connection table");
  RbmParams p{w_re, w_im, b_re, b_im, a_re, a_im, n_visible, n_hidden};
  const int per_block = 4;
  hipLaunchKernelGGL(rbm_log_psi_kernel, dim3(div_up(m_count, per_block)), dim3(64 * per_block), 0,
                     (hipStream_t)stream, p, spins, pairs, m_dev, m_count, conn_kind, conn_a, conn_b, logpsi);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_mh_rbm(const float* w_re, const float* w_im, const float* b_re, const float* b_im, const float* a_re,
              const float* a_im, int32_t n_visible, int32_t n_hidden, int32_t H, int32_t W, int32_t kind,
              int64_t chains, int64_t n_burn, int64_t n_record, int64_t unused, int8_t* state, int8_t* out,
              int64_t* accepts, float* logpsi, const int32_t* r_site, const int32_t* r_move, const double* r_u,
              uint64_t seed, int32_t* nonfinite, void* stream) {
  FK_REQUIRE(n_visible == H * W && n_visible > 0 && n_hidden > 0, "fk_mh_rbm: bad sizes");
  FK_REQUIRE(kind == 0 || kind == 1, "fk_mh_rbm: kind must be 0 (local) or 1 (exchange)");
  FK_REQUIRE(chains >= 0 && n_burn >= 0 && n_record >= 0 && unused >= 0, "fk_mh_rbm: bad counts");
  if (chains == 0) return FK_OK;
  FK_REQUIRE(w_re && w_im && b_re && b_im && a_re && a_im && state && accepts, "fk_mh_rbm: null pointer");
  FK_REQUIRE(n_record == 0 || out, "fk_mh_rbm: null output");
  FK_REQUIRE(!r_u || (r_site && (kind == 0 || r_move)), "fk_mh_rbm: injected randomness incomplete");
  FK_REQUIRE(chains <= 0x7fffffff, "fk_mh_rbm: too many chains");
  MhArgs a;
  a.p = RbmParams{w_re, w_im, b_re, b_im, a_re, a_im, n_visible, n_hidden};
  a.H = H;
  a.W = W;
  a.kind = kind;
  a.chains = chains;
  a.n_burn = n_burn;
  a.n_record = n_record;
  a.unused = unused;
  a.state = state;
  a.out = out;
  a.accepts = accepts;
  a.logpsi = logpsi;
  a.r_site = r_site;
  a.r_move = r_move;
  a.r_u = r_u;
  a.seed = seed;
  a.nonfinite = nonfinite;
  const size_t lds = (size_t)((n_visible + 15) & ~15);
  hipStream_t s = (hipStream_t)stream;
  const int hpl = (n_hidden + 63) / 64;
  if (hpl <= 1) hipLaunchKernelGGL(mh_rbm_kernel<1>, dim3((unsigned)chains), dim3(64), lds, s, a);
  else if (hpl <= 2) hipLaunchKernelGGL(mh_rbm_kernel<2>, dim3((unsigned)chains), dim3(64), lds, s, a);
  else if (hpl <= 4) hipLaunchKernelGGL(mh_rbm_kernel<4>, dim3((unsigned)chains), dim3(64), lds, s, a);
  else if (hpl <= 8) hipLaunchKernelGGL(mh_rbm_kernel<8>, dim3((unsigned)chains), dim3(64), lds, s, a);
  else {
    set_error("fk_mh_rbm: at most 512 hidden units");
    return FK_ERR_UNSUPPORTED;
  }
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_rbm_grad(const float* w_re, const float* w_im, const float* b_re, const float* b_im, const float* a_re,
                const float* a_im, int32_t n_visible, int32_t n_hidden, const int8_t* spins, int64_t batch,
                float* o_re, float* o_im, int64_t row_stride, void* stream) {
  FK_REQUIRE(n_visible > 0 && n_hidden > 0 && batch >= 0, "fk_rbm_grad: bad sizes");
  if (batch == 0) return FK_OK;
  FK_REQUIRE(w_re && w_im && b_re && b_im && a_re && a_im && spins && o_re && o_im, "fk_rbm_grad: null pointer");
  const int64_t P = 2 * (int64_t)n_visible * n_hidden + 2 * (int64_t)n_hidden + 2 * (int64_t)n_visible;
  FK_REQUIRE(row_stride == 0 || row_stride >= P, "fk_rbm_grad: row stride shorter than the parameters");
  FK_REQUIRE(batch <= 0x7fffffff, "fk_rbm_grad: batch too large");
  const size_t lds = (size_t)(2 * n_hidden + n_visible) * sizeof(float);
  FK_REQUIRE(lds <= 64 * 1024, "fk_rbm_grad: too many units");
  RbmParams p{w_re, w_im, b_re, b_im, a_re, a_im, n_visible, n_hidden};
  hipLaunchKernelGGL(rbm_grad_kernel, dim3((unsigned)batch), dim3(256), lds, (hipStream_t)stream, p, spins, batch,
                     o_re, o_im, row_stride > 0 ? row_stride : P);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_mh_propose(const int8_t* state, int64_t chains, int32_t H, int32_t W, int32_t kind, const int32_t* r_site,
                  const int32_t* r_move, int8_t* candidates, void* stream) {
  FK_REQUIRE(kind == 0 || kind == 1, "fk_mh_propose: bad kind");
  if (chains <= 0) return FK_OK;
  FK_REQUIRE(state && r_site && candidates && (kind == 0 || r_move), "fk_mh_propose: null pointer");
  hipLaunchKernelGGL(mh_propose_kernel, dim3(div_up(chains, 256)), dim3(256), 0, (hipStream_t)stream, state, chains,
                     H, W, kind, r_site, r_move, candidates);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_mh_accept(int8_t* state, float* logpsi_cur, const int8_t* candidates, const float* logpsi_cand,
                 const double* uniforms, int64_t chains, int32_t n_sites, int32_t* accepted, int32_t* nonfinite,
                 void* stream) {
  if (chains <= 0) return FK_OK;
  FK_REQUIRE(state && logpsi_cur && candidates && logpsi_cand && uniforms && accepted, "fk_mh_accept: null pointer");
  hipLaunchKernelGGL(mh_accept_kernel, dim3(div_up(chains, 256)), dim3(256), 0, (hipStream_t)stream, state,
                     logpsi_cur, candidates, logpsi_cand, uniforms, chains, n_sites, accepted, nonfinite);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_mh_accept_independent(int8_t* state, float* logpsi_cur, float* logq_cur, const int8_t* candidates,
                             const float* logpsi_cand, const float* logq_cand, const double* uniforms, int64_t chains,
                             int32_t n_sites, int32_t* accepted, int32_t* nonfinite, void* stream) {
  if (chains <= 0) return FK_OK;
  FK_REQUIRE(state && logpsi_cur && logq_cur && candidates && logpsi_cand && logq_cand && uniforms && accepted,
             "fk_mh_accept_independent: null pointer");
  hipLaunchKernelGGL(mh_accept_independent_kernel, dim3(div_up(chains, 256)), dim3(256), 0, (hipStream_t)stream,
                     state, logpsi_cur, logq_cur, candidates, logpsi_cand, logq_cand, uniforms, chains, n_sites,
                     accepted, nonfinite);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
