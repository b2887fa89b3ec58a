// Stochastic reconfiguration on the implicit S = Re[(O-Obar)^H (O-Obar)]/B +
// shift*I (optimizers/stochastic_reconfiguration/optimizer.py:69-92 and its
// force :94-109), never forming S: O (B, P) is stored as two fp32 arrays
// (re, im), the two matvecs read it once each (HBM-bound), products are FP32
// and every reduction across a row or a vector is FP64.
#include "fk_common.h"

#include <cmath>

namespace fk {

// mean over rows of O_re and O_im (column means Obar)
__global__ void o_mean_kernel(const float* __restrict__ ore, const float* __restrict__ oim, int64_t B, int64_t P,
                              float* __restrict__ mre, float* __restrict__ mim) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  double sr = 0.0, si = 0.0;
  for (int64_t b = 0; b < B; ++b) {
    sr += ore[b * P + k];
    si += oim[b * P + k];
  }
  mre[k] = (float)(sr / (double)B);
  mim[k] = (float)(si / (double)B);
}

__global__ __launch_bounds__(1024) void dot_f64_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                       int64_t n, double* __restrict__ out) {
  __shared__ double s[1024];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void sum_f64_kernel(const double* __restrict__ in, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += in[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// u_b = sum_k (O_bk - Obar_k) v_k = O_b.v - Obar.v (complex).  A workgroup
// covers ROWS rows so each v_k is read once per ROWS rows; ov_c = Obar.v
// (re, im) in FP64 is subtracted at the end.  With P a multiple of 4 the rows
// are streamed as float4.
constexpr int OV_ROWS = 4;
__global__ __launch_bounds__(256) void o_times_v_kernel(const float* __restrict__ ore, const float* __restrict__ oim,
                                                        const double* __restrict__ ov_c, const float* __restrict__ v,
                                                        int64_t B, int64_t P, float* __restrict__ u) {
  __shared__ double s_r[OV_ROWS][256], s_i[OV_ROWS][256];
  const int64_t b0 = (int64_t)blockIdx.x * OV_ROWS;
  double dr[OV_ROWS], di[OV_ROWS];
  float ar[OV_ROWS], ai[OV_ROWS];
#pragma unroll
  for (int r = 0; r < OV_ROWS; ++r) dr[r] = di[r] = 0.0, ar[r] = ai[r] = 0.0f;
  int cnt = 0;
  if ((P & 3) == 0) {
    const int64_t P4 = P >> 2;
    const f32x4* v4 = reinterpret_cast<const f32x4*>(v);
    for (int64_t k = threadIdx.x; k < P4; k += blockDim.x) {
      const f32x4 vk = v4[k];
#pragma unroll
      for (int r = 0; r < OV_ROWS; ++r) {
        if (b0 + r < B) {
          const f32x4 x = reinterpret_cast<const f32x4*>(ore + (b0 + r) * P)[k];
          const f32x4 y = reinterpret_cast<const f32x4*>(oim + (b0 + r) * P)[k];
          ar[r] += x.x * vk.x + x.y * vk.y + x.z * vk.z + x.w * vk.w;
          ai[r] += y.x * vk.x + y.y * vk.y + y.z * vk.z + y.w * vk.w;
        }
      }
      if (++cnt == 16) {  // fold fp32 partials into fp64 every 64 terms
#pragma unroll
        for (int r = 0; r < OV_ROWS; ++r) dr[r] += ar[r], di[r] += ai[r], ar[r] = ai[r] = 0.0f;
        cnt = 0;
      }
    }
  } else {
    for (int64_t k = threadIdx.x; k < P; k += blockDim.x) {
      const float vk = v[k];
#pragma unroll
      for (int r = 0; r < OV_ROWS; ++r) {
        if (b0 + r < B) {
          ar[r] += ore[(b0 + r) * P + k] * vk;
          ai[r] += oim[(b0 + r) * P + k] * vk;
        }
      }
      if (++cnt == 64) {
#pragma unroll
        for (int r = 0; r < OV_ROWS; ++r) dr[r] += ar[r], di[r] += ai[r], ar[r] = ai[r] = 0.0f;
        cnt = 0;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < OV_ROWS; ++r) {
    s_r[r][threadIdx.x] = dr[r] + ar[r];
    s_i[r][threadIdx.x] = di[r] + ai[r];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int r = 0; r < OV_ROWS; ++r) {
        s_r[r][threadIdx.x] += s_r[r][threadIdx.x + s];
        s_i[r][threadIdx.x] += s_i[r][threadIdx.x + s];
      }
    }
    __syncthreads();
  }
  if ((int)threadIdx.x < OV_ROWS && b0 + threadIdx.x < B) {
    const int r = threadIdx.x;
    u[2 * (b0 + r)] = (float)(s_r[r][0] - ov_c[0]);
    u[2 * (b0 + r) + 1] = (float)(s_i[r][0] - ov_c[1]);
  }
}

__global__ __launch_bounds__(1024) void sum_u_kernel(const float* __restrict__ u, int64_t B, double* __restrict__ out) {
  __shared__ double s_r[1024], s_i[1024];
  double r = 0.0, i = 0.0;
  for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
    r += u[2 * b];
    i += u[2 * b + 1];
  }
  s_r[threadIdx.x] = r;
  s_i[threadIdx.x] = i;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_r[threadIdx.x] += s_r[threadIdx.x + s];
      s_i[threadIdx.x] += s_i[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_r[0];
    out[1] = s_i[0];
  }
}

// u_b -= mean_b u (complex pairs; sums from sum_u_kernel in FP64).
__global__ void center_pairs_kernel(float* __restrict__ u, int64_t B, const double* __restrict__ sums) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  u[2 * b] = (float)((double)u[2 * b] - sums[0] / (double)B);
  u[2 * b + 1] = (float)((double)u[2 * b + 1] - sums[1] / (double)B);
}

// out_k = scale * Re[sum_b conj(O_bk - Obar_k) u_b] + shift * v_k; each
// thread owns 4 consecutive columns (float4 rows) when P is a multiple of 4.
__global__ void oh_times_u_kernel(const float* __restrict__ ore, const float* __restrict__ oim,
                                  const float* __restrict__ mre, const float* __restrict__ mim,
                                  const float* __restrict__ u, const double* __restrict__ usum, int64_t B, int64_t P,
                                  double scale, double shift, const float* __restrict__ v, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((P & 3) == 0) {
    const int64_t P4 = P >> 2;
    if (t >= P4) return;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    f32x4 part = {0.f, 0.f, 0.f, 0.f};
    int cnt = 0;
#pragma unroll 8
    for (int64_t b = 0; b < B; ++b) {
      const f32x4 x = reinterpret_cast<const f32x4*>(ore + b * P)[t];
      const f32x4 y = reinterpret_cast<const f32x4*>(oim + b * P)[t];
      const float ur = u[2 * b], ui = u[2 * b + 1];
      part += x * ur + y * ui;
      if (++cnt == 64) {
        acc[0] += part.x, acc[1] += part.y, acc[2] += part.z, acc[3] += part.w;
        part = f32x4{0.f, 0.f, 0.f, 0.f};
        cnt = 0;
      }
    }
    acc[0] += part.x, acc[1] += part.y, acc[2] += part.z, acc[3] += part.w;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t k = 4 * t + c;
      double r = scale * (acc[c] - ((double)mre[k] * usum[0] + (double)mim[k] * usum[1]));
      if (v) r += shift * (double)v[k];
      out[k] = (float)r;
    }
    return;
  }
  const int64_t k = t;
  if (k >= P) return;
  double acc = 0.0;
  float part = 0.0f;
  int cnt = 0;
  for (int64_t b = 0; b < B; ++b) {
    part += ore[b * P + k] * u[2 * b] + oim[b * P + k] * u[2 * b + 1];
    if (++cnt == 64) {
      acc += part;
      part = 0.0f;
      cnt = 0;
    }
  }
  acc += part;
  acc -= (double)mre[k] * usum[0] + (double)mim[k] * usum[1];
  double r = scale * acc;
  if (v) r += shift * (double)v[k];
  out[k] = (float)r;
}

// S v with O read from HBM once per product.  The rows go in blocks of R
// (<= SV_MAXR); launch k runs two independent jobs side by side:
//   job A, block k: partial dots O_b . v of the block's rows over SV_NA column
//     chunks (one workgroup per chunk, FP64 partials by block parity);
//   job B, block k-1: u_b = sum of its partials - Obar.v, then
//     acc += sum_b Re[conj(O_b) u_b] over every column (FP32 over the block's
//     rows, FP64 across blocks).
// Job B re-reads rows that job A of the previous launch streamed in, so with
// two blocks in flight (2 R rows x 2 parts x 4 P bytes) inside the 256 MB
// last-level cache the second read does not reach HBM.  (Measured at B = 8192,
// P = 849,156: 20.0 ms at R = 16 against 18.9 ms for the two HBM passes --
// reads served by the last-level cache come no faster than HBM streams, so
// the two-pass form stays the default.)  The kernel boundary
// between the launches orders job A's partials before job B's reads; every
// job-B workgroup reduces the partials in the same fixed order, so all of
// them (and the u they store) agree bit for bit.
constexpr int SV_MAXR = 32;
constexpr int SV_NA = 256;
struct SvArgs {
  const float* ore;
  const float* oim;
  const float* v;
  const double* ov_c;  // Obar.v (re, im)
  double* part;        // [2][SV_MAXR][SV_NA][2] partial dots, by block parity
  float* u;            // (B, 2): u of every row (written by job-B workgroup 0)
  double* acc;         // (P): running sum_b Re[conj(O_b) u_b]
  int64_t B, P4;
  int R, nB;        // rows per block, workgroups of job B
  int64_t ka, kb;  // row blocks of job A and job B (-1: none)
};

__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s, 64);
  return __shfl(x, 0, 64);  // lane 0's order, the same in every wave
}

constexpr int SV_T = 512;  // threads per workgroup (8 waves)
constexpr int SV_RW = SV_MAXR / (SV_T / 64);  // rows per wave in job A
__global__ __launch_bounds__(SV_T, 4) void sv_rows_kernel(SvArgs a) {
  __shared__ float s_u[SV_MAXR][2];
  constexpr int NW = SV_T / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nA = a.ka >= 0 ? SV_NA : 0;
  const f32x4* ore4 = reinterpret_cast<const f32x4*>(a.ore);
  const f32x4* oim4 = reinterpret_cast<const f32x4*>(a.oim);
  if ((int)blockIdx.x < nA) {
    // job A: rows wave, wave + NW, ... of block ka over this chunk's
    // columns, two columns per lane in flight
    const int ch = blockIdx.x;
    const int64_t per = (a.P4 + SV_NA - 1) / SV_NA;
    const int64_t c0 = ch * per, c1 = c0 + per < a.P4 ? c0 + per : a.P4;
    const int64_t b0 = a.ka * a.R;
    int nr = 0;  // rows of this wave
    for (int rr = wave; rr < a.R && b0 + rr < a.B; rr += NW) ++nr;
    float pr[SV_RW], pi[SV_RW];
#pragma unroll
    for (int j = 0; j < SV_RW; ++j) pr[j] = pi[j] = 0.0f;
    const f32x4* v4 = reinterpret_cast<const f32x4*>(a.v);
    constexpr int U = 2;
    for (int64_t cb = c0; cb < c1; cb += 64 * U) {
      f32x4 vk[U], x[U][SV_RW], y[U][SV_RW];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        const int64_t c = cb + 64 * q + lane;
        const bool in = c < c1;
        vk[q] = in ? v4[c] : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < SV_RW; ++j) {
          const int64_t row = (b0 + wave + NW * j) * a.P4 + c;
          x[q][j] = (in && j < nr) ? ore4[row] : f32x4{0.f, 0.f, 0.f, 0.f};
          y[q][j] = (in && j < nr) ? oim4[row] : f32x4{0.f, 0.f, 0.f, 0.f};
        }
      }
#pragma unroll
      for (int q = 0; q < U; ++q)
#pragma unroll
        for (int j = 0; j < SV_RW; ++j) {
          pr[j] += x[q][j].x * vk[q].x + x[q][j].y * vk[q].y + x[q][j].z * vk[q].z + x[q][j].w * vk[q].w;
          pi[j] += y[q][j].x * vk[q].x + y[q][j].y * vk[q].y + y[q][j].z * vk[q].z + y[q][j].w * vk[q].w;
        }
    }
    double* part = a.part + (int64_t)(a.ka & 1) * SV_MAXR * SV_NA * 2;
#pragma unroll
    for (int j = 0; j < SV_RW; ++j) {
      if (j < nr) {
        const double sr = wave_sum_f64((double)pr[j]), si = wave_sum_f64((double)pi[j]);
        if (lane == 0) {
          const int rr = wave + NW * j;
          part[((int64_t)rr * SV_NA + ch) * 2] = sr;
          part[((int64_t)rr * SV_NA + ch) * 2 + 1] = si;
        }
      }
    }
    return;
  }
  if (a.kb < 0) return;
  // job B: u of block kb, then float4 columns bb * SV_T + t, bb * SV_T + t
  // + nB * SV_T, ... with four rows in flight
  const int64_t bb = (int64_t)blockIdx.x - nA;
  const int64_t b0 = a.kb * a.R;
  const int nrow = (int)(b0 + a.R <= a.B ? a.R : a.B - b0);
  const double* part = a.part + (int64_t)(a.kb & 1) * SV_MAXR * SV_NA * 2;
  for (int rr = wave; rr < nrow; rr += NW) {
    double sr = 0.0, si = 0.0;
    for (int ch = lane; ch < SV_NA; ch += 64) {
      sr += part[((int64_t)rr * SV_NA + ch) * 2];
      si += part[((int64_t)rr * SV_NA + ch) * 2 + 1];
    }
    sr = wave_sum_f64(sr);
    si = wave_sum_f64(si);
    if (lane == 0) {
      const float ur = (float)(sr - a.ov_c[0]), ui = (float)(si - a.ov_c[1]);
      s_u[rr][0] = ur;
      s_u[rr][1] = ui;
      if (bb == 0) {
        a.u[2 * (b0 + rr)] = ur;
        a.u[2 * (b0 + rr) + 1] = ui;
      }
    }
  }
  __syncthreads();
  for (int64_t c = bb * SV_T + threadIdx.x; c < a.P4; c += (int64_t)a.nB * SV_T) {
    f32x4 s = {0.f, 0.f, 0.f, 0.f};
    constexpr int UR = 4;
    for (int r0 = 0; r0 < nrow; r0 += UR) {
      f32x4 x[UR], y[UR];
#pragma unroll
      for (int q = 0; q < UR; ++q) {
        const int64_t row = (b0 + r0 + q) * a.P4 + c;
        x[q] = r0 + q < nrow ? ore4[row] : f32x4{0.f, 0.f, 0.f, 0.f};
        y[q] = r0 + q < nrow ? oim4[row] : f32x4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int q = 0; q < UR; ++q)
        if (r0 + q < nrow) s += x[q] * s_u[r0 + q][0] + y[q] * s_u[r0 + q][1];
    }
    double2* acc = reinterpret_cast<double2*>(a.acc) + 2 * c;
    if (a.kb == 0) {
      acc[0] = make_double2((double)s.x, (double)s.y);
      acc[1] = make_double2((double)s.z, (double)s.w);
    } else {
      const double2 p0 = acc[0], p1 = acc[1];
      acc[0] = make_double2(p0.x + (double)s.x, p0.y + (double)s.y);
      acc[1] = make_double2(p1.x + (double)s.z, p1.y + (double)s.w);
    }
  }
}

// out_k = scale * (acc_k - Re[conj(Obar_k) sum_b u_b]) + shift * v_k
__global__ void sv_finish_kernel(const double* __restrict__ acc, const float* __restrict__ mre,
                                 const float* __restrict__ mim, const double* __restrict__ usum, int64_t P,
                                 double scale, double shift, const float* __restrict__ v, float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  double r = scale * (acc[k] - ((double)mre[k] * usum[0] + (double)mim[k] * usum[1]));
  if (v) r += shift * (double)v[k];
  out[k] = (float)r;
}

// y = a*x + y (fp32 vectors, fp64 coefficient)
__global__ void axpy_kernel(double alpha, const float* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)((double)y[i] + alpha * (double)x[i]);
}
// p = r + beta*p
__global__ void xpby_kernel(const float* __restrict__ r, double beta, float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (float)((double)r[i] + beta * (double)p[i]);
}

// Device-side conjugate gradient (linear_equations.py:87-131, no
// preconditioner) for a host loop with no host round trip.  State in
// st[CG_N] doubles: gamma = r.r, p.z, gamma_new, tol_abs (tol*||b|| with both
// factors rounded to float32, as the reference's float norms), done flag,
// iterations, max_iter.  A step after convergence (or after max_iter) leaves
// x, r, p unchanged, so a fixed number of launched steps gives the same
// result as the reference's early exit.
enum { CG_GAMMA = 0, CG_PZ = 1, CG_GNEW = 2, CG_TOL = 3, CG_DONE = 4, CG_ITER = 5, CG_MAXIT = 6, CG_N = 8 };

__global__ void cg_init_kernel(const float* __restrict__ b, float* __restrict__ x, float* __restrict__ r,
                               float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = b[i];
  x[i] = 0.0f;
  r[i] = v;
  p[i] = v;
}

__global__ void cg_init_scalars_kernel(double* st, double tol, double max_iter) {
  const float nb = (float)sqrt(st[CG_GAMMA]);
  st[CG_TOL] = (double)((float)tol * nb);
  st[CG_ITER] = 0.0;
  st[CG_MAXIT] = max_iter;
  st[CG_DONE] = (max_iter <= 0.0 || !((double)nb > st[CG_TOL])) ? 1.0 : 0.0;
}

__global__ void cg_update_xr_kernel(const double* __restrict__ st, const float* __restrict__ p,
                                    const float* __restrict__ z, float* __restrict__ x, float* __restrict__ r,
                                    int64_t n) {
  if (st[CG_DONE] != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = st[CG_GAMMA] / st[CG_PZ];
  x[i] = (float)((double)x[i] + alpha * (double)p[i]);
  r[i] = (float)((double)r[i] - alpha * (double)z[i]);
}

__global__ void cg_update_p_kernel(const double* __restrict__ st, const float* __restrict__ r, float* __restrict__ p,
                                   int64_t n) {
  if (st[CG_DONE] != 0.0) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = st[CG_GNEW] / st[CG_GAMMA];
  p[i] = (float)((double)r[i] + beta * (double)p[i]);
}

__global__ void cg_finish_step_kernel(double* st) {
  if (st[CG_DONE] != 0.0) return;
  st[CG_GAMMA] = st[CG_GNEW];
  st[CG_ITER] += 1.0;
  const float nr = (float)sqrt(st[CG_GAMMA]);
  if (!((double)nr > st[CG_TOL]) || st[CG_ITER] >= st[CG_MAXIT]) st[CG_DONE] = 1.0;
}

// Keras/TensorFlow Adam (tf.keras Adam, the optimizer of
// examples/basic_autoregressive_2d.py:24): with lr_t = lr*sqrt(1-b2^t)/(1-b1^t)
// given by the host, m += (g - m)(1 - b1), v += (g^2 - v)(1 - b2),
// p -= lr_t * m / (sqrt(v) + eps), all in FP32 as TF's ApplyAdam on float32
// variables (epsilon added to the uncorrected sqrt(v)).
__global__ void adam_step_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                 const float* __restrict__ g, int64_t n, float lr_t, float one_m_b1, float one_m_b2,
                                 float eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  float mi = m[i], vi = v[i];
  mi += (gi - mi) * one_m_b1;
  vi += (gi * gi - vi) * one_m_b2;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

}  // namespace fk

using namespace fk;

extern "C" {
int fk_adam_step(float* params, float* m, float* v, const float* grad, int64_t n, double lr, double beta_1,
                 double beta_2, double epsilon, int64_t t, void* stream) {
  FK_REQUIRE(params && m && v && grad, "fk_adam_step: null pointer");
  FK_REQUIRE(t >= 1 && beta_1 >= 0 && beta_1 < 1 && beta_2 >= 0 && beta_2 < 1, "fk_adam_step: bad arguments");
  if (n <= 0) return FK_OK;
  const float lr_t = (float)(lr * sqrt(1.0 - pow(beta_2, (double)t)) / (1.0 - pow(beta_1, (double)t)));
  hipLaunchKernelGGL(adam_step_kernel, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, params, m, v, grad,
                     n, lr_t, (float)(1.0 - beta_1), (float)(1.0 - beta_2), (float)epsilon);
  FK_CHECK_LAUNCH();
  return FK_OK;
}


int fk_o_mean(const float* o_re, const float* o_im, int64_t batch, int64_t n_params, float* mean_re, float* mean_im,
              void* stream) {
  FK_REQUIRE(o_re && o_im && mean_re && mean_im, "fk_o_mean: null pointer");
  FK_REQUIRE(batch > 0 && n_params > 0, "fk_o_mean: empty");
  hipLaunchKernelGGL(o_mean_kernel, dim3(div_up(n_params, 256)), dim3(256), 0, (hipStream_t)stream, o_re, o_im, batch,
                     n_params, mean_re, mean_im);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_o_times_v(const float* o_re, const float* o_im, const float* mean_re, const float* mean_im, const float* v,
                 int64_t batch, int64_t n_params, float* u, double* scratch, void* stream) {
  FK_REQUIRE(o_re && o_im && mean_re && mean_im && v && u && scratch, "fk_o_times_v: null pointer");
  FK_REQUIRE(batch > 0 && n_params > 0, "fk_o_times_v: bad sizes");
  hipStream_t s = (hipStream_t)stream;
  // scratch: [0..255] partials, [256] Obar_re.v, [257] Obar_im.v
  hipLaunchKernelGGL(dot_f64_kernel, dim3(256), dim3(1024), 0, s, mean_re, v, n_params, scratch);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, s, scratch, 256, scratch + 256);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(dot_f64_kernel, dim3(256), dim3(1024), 0, s, mean_im, v, n_params, scratch);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, s, scratch, 256, scratch + 257);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(o_times_v_kernel, dim3(div_up(batch, OV_ROWS)), dim3(256), 0, s, o_re, o_im, scratch + 256, v,
                     batch, n_params, u);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_oh_times_u(const float* o_re, const float* o_im, const float* mean_re, const float* mean_im, const float* u,
                  int64_t batch, int64_t n_params, double scale, double shift, const float* v, float* out,
                  double* scratch2, void* stream) {
  FK_REQUIRE(o_re && o_im && mean_re && mean_im && u && out && scratch2, "fk_oh_times_u: null pointer");
  FK_REQUIRE(batch > 0 && n_params > 0, "fk_oh_times_u: bad sizes");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sum_u_kernel, dim3(1), dim3(1024), 0, s, u, batch, scratch2);
  FK_CHECK_LAUNCH();
  const int64_t cols = (n_params & 3) == 0 ? n_params / 4 : n_params;
  hipLaunchKernelGGL(oh_times_u_kernel, dim3(div_up(cols, 256)), dim3(256), 0, s, o_re, o_im, mean_re, mean_im, u,
                     scratch2, batch, n_params, scale, shift, v, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_sr_scratch_doubles(void) { return 260; }

int64_t fk_sr_matvec_workspace_bytes(int64_t batch, int64_t n_params) {
  (void)batch;
  return (int64_t)(2 * SV_MAXR * SV_NA * 2) * 8 + ((n_params + 63) / 64) * 64 * 8;
}

int fk_sr_matvec(const float* o_re, const float* o_im, const float* mean_re, const float* mean_im, const float* v,
                 int64_t batch, int64_t n_params, double scale, double shift, float* out, float* u,
                 int32_t block_rows, void* workspace, int64_t workspace_bytes, double* scratch, void* stream) {
  FK_REQUIRE(o_re && o_im && mean_re && mean_im && v && out && u && scratch, "fk_sr_matvec: null pointer");
  FK_REQUIRE(batch > 0 && n_params > 0, "fk_sr_matvec: bad sizes");
  FK_REQUIRE(block_rows >= 0 && block_rows <= SV_MAXR, "fk_sr_matvec: block_rows must be in [0, %d]", SV_MAXR);
  if (block_rows == 0 || (n_params & 3)) {  // the two-pass products
    const int rc = fk_o_times_v(o_re, o_im, mean_re, mean_im, v, batch, n_params, u, scratch, stream);
    if (rc != FK_OK) return rc;
    return fk_oh_times_u(o_re, o_im, mean_re, mean_im, u, batch, n_params, scale, shift, v, out, scratch, stream);
  }
  FK_REQUIRE(workspace && workspace_bytes >= fk_sr_matvec_workspace_bytes(batch, n_params),
             "fk_sr_matvec: workspace of %lld bytes, need %lld", (long long)workspace_bytes,
             (long long)fk_sr_matvec_workspace_bytes(batch, n_params));
  FK_REQUIRE(((uintptr_t)workspace & 15) == 0, "fk_sr_matvec: workspace must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  // scratch: [0..255] dot partials, [256..257] Obar.v, [258..259] sum_b u_b
  for (int part = 0; part < 2; ++part) {
    hipLaunchKernelGGL(dot_f64_kernel, dim3(256), dim3(1024), 0, s, part ? mean_im : mean_re, v, n_params, scratch);
    FK_CHECK_LAUNCH();
    hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, s, scratch, 256, scratch + 256 + part);
    FK_CHECK_LAUNCH();
  }
  SvArgs a{};
  a.ore = o_re;
  a.oim = o_im;
  a.v = v;
  a.ov_c = scratch + 256;
  a.part = reinterpret_cast<double*>(workspace);
  a.acc = a.part + 2 * SV_MAXR * SV_NA * 2;
  a.u = u;
  a.B = batch;
  a.P4 = n_params / 4;
  a.R = block_rows;
  const int64_t nblk = (batch + a.R - 1) / a.R;
  const int nB = div_up(a.P4, SV_T) < SV_NA ? div_up(a.P4, SV_T) : SV_NA;
  a.nB = nB;
  for (int64_t k = 0; k <= nblk; ++k) {
    a.ka = k < nblk ? k : -1;
    a.kb = k - 1;
    const int grid = (a.ka >= 0 ? SV_NA : 0) + (a.kb >= 0 ? nB : 0);
    hipLaunchKernelGGL(sv_rows_kernel, dim3(grid), dim3(SV_T), 0, s, a);
    FK_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(sum_u_kernel, dim3(1), dim3(1024), 0, s, u, batch, scratch + 258);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(sv_finish_kernel, dim3(div_up(n_params, 256)), dim3(256), 0, s, a.acc, mean_re, mean_im,
                     scratch + 258, n_params, scale, shift, v, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_dot_f64(const float* a, const float* b, int64_t n, double* scratch, double* out, void* stream) {
  FK_REQUIRE(a && b && scratch && out, "fk_dot_f64: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const int blocks = 256;
  hipLaunchKernelGGL(dot_f64_kernel, dim3(blocks), dim3(1024), 0, s, a, b, n, scratch);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, s, scratch, blocks, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_cg_state_doubles(void) { return CG_N; }

int fk_cg_init(const float* rhs, float* x, float* r, float* p, int64_t n, double tol, int64_t max_iter,
               double* state, double* scratch, void* stream) {
  FK_REQUIRE(rhs && x && r && p && state && scratch && n > 0, "fk_cg_init: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(cg_init_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, rhs, x, r, p, n);
  FK_CHECK_LAUNCH();
  const int rc = fk_dot_f64(rhs, rhs, n, scratch, state + CG_GAMMA, stream);
  if (rc != FK_OK) return rc;
  hipLaunchKernelGGL(cg_init_scalars_kernel, dim3(1), dim3(1), 0, s, state, tol, (double)max_iter);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_cg_step(float* x, float* r, float* p, const float* z, int64_t n, double* state, double* scratch,
               void* stream) {
  FK_REQUIRE(x && r && p && z && state && scratch && n > 0, "fk_cg_step: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc = fk_dot_f64(p, z, n, scratch, state + CG_PZ, stream);
  if (rc != FK_OK) return rc;
  hipLaunchKernelGGL(cg_update_xr_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, state, p, z, x, r, n);
  FK_CHECK_LAUNCH();
  rc = fk_dot_f64(r, r, n, scratch, state + CG_GNEW, stream);
  if (rc != FK_OK) return rc;
  hipLaunchKernelGGL(cg_update_p_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, state, r, p, n);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(cg_finish_step_kernel, dim3(1), dim3(1), 0, s, state);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_center_pairs(float* u, int64_t batch, double* scratch, void* stream) {
  FK_REQUIRE(u && scratch && batch > 0, "fk_center_pairs: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(sum_u_kernel, dim3(1), dim3(1024), 0, s, u, batch, scratch);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(center_pairs_kernel, dim3(div_up(batch, 256)), dim3(256), 0, s, u, batch, scratch);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_axpy(double alpha, const float* x, float* y, int64_t n, void* stream) {
  FK_REQUIRE(x && y, "fk_axpy: null pointer");
  if (n <= 0) return FK_OK;
  hipLaunchKernelGGL(axpy_kernel, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, alpha, x, y, n);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_xpby(const float* r, double beta, float* p, int64_t n, void* stream) {
  FK_REQUIRE(r && p, "fk_xpby: null pointer");
  if (n <= 0) return FK_OK;
  hipLaunchKernelGGL(xpby_kernel, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, r, beta, p, n);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
