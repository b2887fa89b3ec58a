// Operators (find_conn), the ragged index map and local-energy reductions.
// Integer/byte and small-FP64 work: HBM/latency bound, no MFMA.
#include "fk_common.h"

#include <algorithm>
#include <cstdarg>
#include <string>

namespace fk {

static thread_local char g_err[1024];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int device_cu_count() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 256;
  return cus;
}

// Connection slot c of sample b follows HeisenbergFindConn.calc_conn_and_mel
// (operators/heisenberg.py:84-121), Ising.find_conn (operators/ising.py:17-46)
// or the J1J2 bond operator (operators/j1j2.py:37-55 through
// operators/netket_operator.py:46-66).
// One workgroup per tile of up to 64 consecutive samples, one lane per
// sample: for each connection slot c a wave writes mel[c][b0..] and
// use[c][b0..] as one contiguous run (the slot-major layout of the dense
// contract), the slots shared out over the waves.
// Diagonal element: FK_DIAG_SUM_F64 accumulates in double in table order per
// sample (exact for the integer/dyadic couplings of Heisenberg and J1J2);
// FK_DIAG_ISING_F32 reproduces Ising.find_conn's float32 arithmetic
// (operators/ising.py:25,30-41, see IsingDiagCursor / pairwise_sum).

// Per-site values of the Ising diagonal, produced in ascending site order:
// mel is a float32 array (ising.py:25) from which each bond term -j*s*s' is
// subtracted in place (:33,37,39), i.e. computed in double and rounded to
// float32 after every term.  The table lists site a's terms in the
// reference's order (down, then right) and is grouped by a.
struct IsingDiagCursor {
  const int2* ab;
  const double* jj;
  const int8_t* row;
  int n_diag, k, site;
  __device__ __forceinline__ float next() {
    float v = 0.0f;
    while (k < n_diag && ab[k].x == site) {
      v = (float)((double)v + jj[k] * (double)row[ab[k].x] * (double)row[ab[k].y]);
      ++k;
    }
    ++site;
    return v;
  }
};

// numpy's float32 pairwise summation of n consecutive values (the inner loop
// of ndarray.sum over contiguous axes, float32 accumulators): fewer than 8
// values are added one by one; up to 128 go to 8 interleaved partial sums
// combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) and then the remainder;
// longer runs split at n/2 rounded down to a multiple of 8.  The values are
// consumed in ascending order, so the cursor produces them on demand.
__device__ __forceinline__ float pairwise_leaf(IsingDiagCursor& c, int n) {
  if (n < 8) {
    float res = 0.0f;
    for (int i = 0; i < n; ++i) res += c.next();
    return res;
  }
  float r[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) r[q] = c.next();
  int i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] += c.next();
  }
  float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res += c.next();
  return res;
}
template <int LEVEL>
__device__ __noinline__ float pairwise_sum(IsingDiagCursor& c, int n) {
  if constexpr (LEVEL == 0) {
    return pairwise_leaf(c, n);  // n <= 128 here (n_sites <= 65536)
  } else {
    if (n <= 128) return pairwise_leaf(c, n);
    int n2 = n / 2;
    n2 -= n2 % 8;
    const float a = pairwise_sum<LEVEL - 1>(c, n2);
    const float b = pairwise_sum<LEVEL - 1>(c, n - n2);
    return a + b;
  }
}

__global__ __launch_bounds__(1024) void find_conn_kernel(
    const int8_t* __restrict__ spins, int64_t batch, int n_sites, int tile, int kind,
    const int32_t* __restrict__ conn_a, const int32_t* __restrict__ conn_b,
    const double* __restrict__ conn_mel, int n_conn,
    const int32_t* __restrict__ diag_a, const int32_t* __restrict__ diag_b,
    const double* __restrict__ diag_j, int n_diag, int diag_rule,
    int8_t* __restrict__ conn_out, double* __restrict__ mel_out,
    uint8_t* __restrict__ use_out, int32_t* __restrict__ n_used_out) {
  // LDS: the diagonal bond table (site pairs, couplings), then the tile's
  // spin rows
  extern __shared__ __align__(16) uint8_t fc_sm[];
  int2* s_dab = reinterpret_cast<int2*>(fc_sm);
  double* s_dj = reinterpret_cast<double*>(fc_sm + 8 * (size_t)n_diag);
  int8_t* s_spin = reinterpret_cast<int8_t*>(fc_sm + 16 * (size_t)n_diag);
  __shared__ int s_cnt[16][64];
  const int64_t b0 = (int64_t)blockIdx.x * tile;
  const int nb = (int)((batch - b0) < tile ? (batch - b0) : tile);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, NW = blockDim.x >> 6;
  for (int k = tid; k < n_diag; k += blockDim.x) {
    s_dab[k] = make_int2(diag_a[k], diag_b[k]);
    s_dj[k] = diag_j[k];
  }
  for (int q = tid; q < nb * n_sites; q += blockDim.x) s_spin[q] = spins[b0 * n_sites + q];
  __syncthreads();
  const bool ok = lane < nb;
  const int8_t* row = s_spin + (ok ? lane : 0) * n_sites;
  const int64_t b = b0 + lane;
  int local = 0;
  // wave 0: the diagonal; the other waves (all of them in a one-wave
  // workgroup): the off-diagonal slots
  if (wave == 0) {
    double d = 0.0;
    if (diag_rule == FK_DIAG_ISING_F32) {
      IsingDiagCursor cur{s_dab, s_dj, row, n_diag, 0, 0};
      d = (double)pairwise_sum<10>(cur, n_sites);
    } else {
#pragma unroll 8
      for (int k = 0; k < n_diag; ++k) {
        const int2 ab = s_dab[k];
        d += s_dj[k] * (double)row[ab.x] * (double)row[ab.y];
      }
    }
    if (ok) {
      mel_out[b] = d;
      use_out[b] = 1;
    }
    local = 1;
  }
  const int w0 = NW > 1 ? 1 : 0, nw = NW > 1 ? NW - 1 : 1;
#pragma unroll 4
  for (int c = 1 + wave - w0; wave >= w0 && c < n_conn; c += nw) {
    const int a = conn_a[c];
    int u = 0;
    double m = 0.0;
    if (kind == FK_CONN_FLIP) {
      u = (a >= 0);
      m = u ? conn_mel[c] : 0.0;
    } else if (a >= 0) {
      const int bb = conn_b[c];
      const int anti = row[a] != row[bb];
      m = anti ? conn_mel[c] : 0.0;
      u = anti && (m != 0.0);
    }
    if (ok) {
      mel_out[(int64_t)c * batch + b] = m;
      use_out[(int64_t)c * batch + b] = (uint8_t)u;
    }
    local += u;
  }
  s_cnt[wave][lane] = local;
  __syncthreads();
  if (wave == 0 && ok && n_used_out) {
    int t = 0;
    for (int w = 0; w < NW; ++w) t += s_cnt[w][lane];
    n_used_out[b] = t;
  }
  if (!conn_out) return;
  // Dense connected configurations, slot-major (n_conn, batch, n_sites): the
  // tile's rows of one slot are one contiguous run.
  const int run = nb * n_sites;
  for (int c = 0; c < n_conn; ++c) {
    const int a = c > 0 ? conn_a[c] : -1;
    const int bb = (a >= 0 && kind != FK_CONN_FLIP) ? conn_b[c] : -1;
    const bool has_mel = c > 0 && conn_mel[c] != 0.0;
    int8_t* dst = conn_out + ((int64_t)c * batch + b0) * n_sites;
    for (int idx = tid; idx < run; idx += blockDim.x) {
      const int l = idx / n_sites, q = idx - l * n_sites;
      const int8_t* r = s_spin + l * n_sites;
      int8_t v = r[q];
      if (c > 0) {
        if (kind == FK_CONN_FLIP) {
          if (q == a) v = (int8_t)-v;
        } else if (a >= 0) {
          if (kind == FK_CONN_SWAP_ZERO_PAD && !(r[a] != r[bb] && has_mel)) {
            v = 0;
          } else {
            if (q == a) v = r[bb];
            else if (q == bb) v = r[a];
          }
        } else if (kind == FK_CONN_SWAP_ZERO_PAD) {
          v = 0;
        }
      }
      dst[idx] = v;
    }
  }
}

__global__ void zero_i64_kernel(int64_t* p) { *p = 0; }

// counts[b] -> offsets[b+1]: a workgroup per 64 samples, lane = sample, the
// slots shared out over the waves (64-byte coalesced reads of use[c][b0..]),
// partial counts summed through LDS.
__global__ __launch_bounds__(256) void count_used_kernel(const uint8_t* __restrict__ use, int64_t batch, int n_conn,
                                                         int64_t* __restrict__ offsets) {
  __shared__ int s_cnt[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, NW = blockDim.x >> 6;
  const int64_t b = (int64_t)blockIdx.x * 64 + lane;
  int t = 0;
  if (b < batch) {
#pragma unroll 4
    for (int c = wave; c < n_conn; c += NW) t += use[(int64_t)c * batch + b] ? 1 : 0;
  }
  s_cnt[wave][lane] = t;
  __syncthreads();
  if (wave == 0 && b < batch) {
    int64_t n = 0;
    for (int w = 0; w < NW; ++w) n += s_cnt[w][lane];
    offsets[b + 1] = n;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) offsets[0] = 0;
}

// In-place inclusive scan of offsets[1..batch] by a single workgroup.
__global__ __launch_bounds__(1024) void scan_offsets_kernel(int64_t* __restrict__ offsets, int64_t batch) {
  __shared__ int64_t s_wave[16];
  __shared__ int64_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < batch; base += blockDim.x) {
    const int64_t i = base + tid;
    int64_t v = (i < batch) ? offsets[i + 1] : 0;
    for (int off = 1; off < 64; off <<= 1) {
      int64_t o = __shfl_up(v, off, 64);
      if (lane >= off) v += o;
    }
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    if (tid == 0) {
      int64_t run = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        int64_t t = s_wave[w];
        s_wave[w] = run;
        run += t;
      }
    }
    __syncthreads();
    v += s_wave[wave] + s_carry;
    if (i < batch) offsets[i + 1] = v;
    __syncthreads();
    if (tid == (int)blockDim.x - 1) s_carry = v;
    __syncthreads();
  }
}

// pairs of sample b at offsets[b]...: a workgroup per 64 samples, lane =
// sample, wave w takes the slot range [w*n_conn/NW, (w+1)*n_conn/NW); each
// wave counts its range, an exclusive prefix over the waves places it, and
// the pairs come out in (sample, slot) order.
__global__ __launch_bounds__(256) void write_pairs_kernel(const uint8_t* __restrict__ use, int64_t batch, int n_conn,
                                                          const int64_t* __restrict__ offsets,
                                                          int32_t* __restrict__ pairs) {
  __shared__ int s_cnt[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, NW = blockDim.x >> 6;
  const int64_t b = (int64_t)blockIdx.x * 64 + lane;
  const int c0 = (int)((int64_t)n_conn * wave / NW), c1 = (int)((int64_t)n_conn * (wave + 1) / NW);
  int t = 0;
  if (b < batch) {
#pragma unroll 4
    for (int c = c0; c < c1; ++c) t += use[(int64_t)c * batch + b] ? 1 : 0;
  }
  s_cnt[wave][lane] = t;
  __syncthreads();
  if (b >= batch) return;
  int64_t k = offsets[b];
  for (int w = 0; w < wave; ++w) k += s_cnt[w][lane];
#pragma unroll 4
  for (int c = c0; c < c1; ++c) {
    if (use[(int64_t)c * batch + b]) {
      pairs[2 * k] = (int32_t)b;
      pairs[2 * k + 1] = c;
      ++k;
    }
  }
}

// Row buckets of the ragged map (the forward's row reuse): bucket_of[c] is
// the bucket of connection slot c.  Counts per (bucket, sample) with one lane
// per sample walking the slots in order (counters in LDS).
constexpr int BK_MAX = 128;
__global__ __launch_bounds__(64) void bucket_count_kernel(const uint8_t* __restrict__ use, int64_t batch, int n_conn,
                                                          const int32_t* __restrict__ bucket_of, int nbk,
                                                          int32_t* __restrict__ counts) {
  __shared__ int s_cnt[BK_MAX][64];
  const int lane = threadIdx.x;
  const int64_t b = (int64_t)blockIdx.x * 64 + lane;
  for (int k = 0; k < nbk; ++k) s_cnt[k][lane] = 0;
  if (b < batch)
    for (int c = 0; c < n_conn; ++c)
      if (use[(int64_t)c * batch + b]) s_cnt[bucket_of[c]][lane] += 1;
  if (b < batch)
    for (int k = 0; k < nbk; ++k) counts[(int64_t)k * batch + b] = s_cnt[k][lane];
}

// In-place exclusive scan of counts[k][0..batch) per bucket k (a workgroup
// per bucket); totals[k] = the bucket's size.
__global__ __launch_bounds__(1024) void bucket_scan_kernel(int32_t* __restrict__ counts, int64_t batch,
                                                           int64_t* __restrict__ totals) {
  __shared__ int64_t s_wave[16];
  __shared__ int64_t s_carry;
  int32_t* row = counts + (int64_t)blockIdx.x * batch;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < batch; base += blockDim.x) {
    const int64_t i = base + tid;
    const int64_t x = i < batch ? row[i] : 0;
    int64_t v = x;
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t o = __shfl_up(v, off, 64);
      if (lane >= off) v += o;
    }
    if (lane == 63) s_wave[wave] = v;
    __syncthreads();
    if (tid == 0) {
      int64_t run = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        const int64_t t = s_wave[w];
        s_wave[w] = run;
        run += t;
      }
    }
    __syncthreads();
    v += s_wave[wave] + s_carry;
    if (i < batch) row[i] = (int32_t)(v - x);
    __syncthreads();
    if (tid == (int)blockDim.x - 1) s_carry = v;
    __syncthreads();
  }
  if (tid == 0) totals[blockIdx.x] = s_carry;
}

__global__ void bucket_offsets_kernel(const int64_t* __restrict__ totals, int nbk, int64_t* __restrict__ boff) {
  int64_t run = 0;
  for (int k = 0; k < nbk; ++k) {
    boff[k] = run;
    run += totals[k];
  }
  boff[nbk] = run;
}

// pairs in (bucket, sample, slot) order and, for each, its index in the
// sample-major order of fk_compact_connections.
__global__ __launch_bounds__(64) void bucket_write_kernel(const uint8_t* __restrict__ use, int64_t batch, int n_conn,
                                                          const int32_t* __restrict__ bucket_of, int nbk,
                                                          const int32_t* __restrict__ starts,
                                                          const int64_t* __restrict__ boff,
                                                          const int64_t* __restrict__ offsets,
                                                          int32_t* __restrict__ pairs, int64_t* __restrict__ out_index) {
  __shared__ int s_run[BK_MAX][64];
  const int lane = threadIdx.x;
  const int64_t b = (int64_t)blockIdx.x * 64 + lane;
  for (int k = 0; k < nbk; ++k) s_run[k][lane] = 0;
  if (b >= batch) return;
  int64_t kk = offsets[b];
  for (int c = 0; c < n_conn; ++c) {
    if (!use[(int64_t)c * batch + b]) continue;
    const int bk = bucket_of[c];
    const int64_t pos = boff[bk] + starts[(int64_t)bk * batch + b] + s_run[bk][lane]++;
    pairs[2 * pos] = (int32_t)b;
    pairs[2 * pos + 1] = c;
    out_index[pos] = kk++;
  }
}

__device__ __forceinline__ void cexp_diff(float dr, float di, double& er, double& ei) {
  // exp of a complex64 difference, evaluated in float like the reference's
  // numpy complex64 arithmetic (observables/monte_carlo/operator.py:31).
  float m = expf(dr);
  float s, c;
  sincosf(di, &s, &c);
  er = (double)(m * c);
  ei = (double)(m * s);
}

// One wave per sample, lanes over its used connections (k0 is slot 0, the
// sample itself); the FP64 partial sums are reduced across the wave in a
// fixed order.
__global__ __launch_bounds__(256) void local_energy_kernel(const int64_t* __restrict__ offsets,
                                                           const int32_t* __restrict__ pairs,
                                                           const float* __restrict__ logpsi,
                                                           const double* __restrict__ mel,
                                                           const double* __restrict__ mel_im, int64_t batch,
                                                           double* __restrict__ eloc) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= batch) return;
  const int64_t k0 = offsets[b], k1 = offsets[b + 1];
  double accr = 0.0, acci = 0.0;
  if (k1 > k0) {
    const float r0 = logpsi[2 * k0], i0 = logpsi[2 * k0 + 1];
    for (int64_t k = k0 + lane; k < k1; k += 64) {
      const int c = pairs[2 * k + 1];
      double er, ei;
      cexp_diff(logpsi[2 * k] - r0, logpsi[2 * k + 1] - i0, er, ei);
      const double m = mel[(int64_t)c * batch + b];
      const double mi = mel_im ? mel_im[(int64_t)c * batch + b] : 0.0;
      accr += m * er - mi * ei;
      acci += m * ei + mi * er;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    accr += __shfl_xor(accr, off, 64);
    acci += __shfl_xor(acci, off, 64);
  }
  if (lane == 0) {
    eloc[2 * b] = accr;
    eloc[2 * b + 1] = acci;
  }
}

__global__ void local_energy_dense_kernel(const float* __restrict__ logpsi, const double* __restrict__ mel,
                                          const double* __restrict__ mel_im, int n_conn, int64_t batch,
                                          double* __restrict__ eloc) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const float r0 = logpsi[2 * b], i0 = logpsi[2 * b + 1];
  double accr = 0.0, acci = 0.0;
  for (int c = 0; c < n_conn; ++c) {
    const int64_t k = (int64_t)c * batch + b;
    double er, ei;
    cexp_diff(logpsi[2 * k] - r0, logpsi[2 * k + 1] - i0, er, ei);
    const double mi = mel_im ? mel_im[k] : 0.0;
    accr += mel[k] * er - mi * ei;
    acci += mel[k] * ei + mi * er;
  }
  eloc[2 * b] = accr;
  eloc[2 * b + 1] = acci;
}

// Magnetisation per sample (observables/monte_carlo/sigma_z.py:7-16): one
// wave per sample sums its int8 spins exactly in integers; the value (or its
// absolute value) divided by n_sites is written as a (re, 0) pair.
__global__ __launch_bounds__(256) void sigma_z_kernel(const int8_t* __restrict__ spins, int64_t batch, int n_sites,
                                                      int absolute, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= batch) return;
  const int8_t* row = spins + b * n_sites;
  int acc = 0;
  for (int q = lane; q < n_sites; q += 64) acc += row[q];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) {
    out[2 * b] = (double)(absolute ? abs(acc) : acc) / (double)n_sites;
    out[2 * b + 1] = 0.0;
  }
}

// Two-pass FP64 mean / population variance of Re (numpy.mean / numpy.var in
// observables/monte_carlo/observable.py:12-13), fixed reduction order.
__global__ __launch_bounds__(1024) void energy_moments_kernel(const double* __restrict__ eloc, int64_t batch,
                                                              double* __restrict__ out) {
  __shared__ double s_r[1024], s_i[1024];
  const int tid = threadIdx.x;
  double r = 0.0, im = 0.0;
  for (int64_t b = tid; b < batch; b += blockDim.x) {
    r += eloc[2 * b];
    im += eloc[2 * b + 1];
  }
  s_r[tid] = r;
  s_i[tid] = im;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) {
      s_r[tid] += s_r[tid + s];
      s_i[tid] += s_i[tid + s];
    }
    __syncthreads();
  }
  const double mr = s_r[0] / (double)batch, mi = s_i[0] / (double)batch;
  __syncthreads();
  double v = 0.0;
  for (int64_t b = tid; b < batch; b += blockDim.x) {
    const double d = eloc[2 * b] - mr;
    v += d * d;
  }
  s_r[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) s_r[tid] += s_r[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = mr;
    out[1] = mi;
    out[2] = s_r[0] / (double)batch;
    out[3] = (double)batch;
  }
}

__global__ void ar_draw_site_kernel(const float* __restrict__ cond_logp, const double* __restrict__ u,
                                    int64_t batch, int n_sites, int site, int8_t* __restrict__ spins) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const float lp = cond_logp[(b * n_sites + site) * 2];
  // exp of the float32 log-probability compared with the float64 uniform,
  // strict '>' (deepar/samplers/autoregressive.py:42-44).
  const double p = (double)expf(lp);
  spins[b * n_sites + site] = (p > u[b * n_sites + site]) ? (int8_t)1 : (int8_t)-1;
}

}  // namespace fk

using namespace fk;

extern "C" {

int fk_version(void) { return 1; }

const char* fk_last_error(void) { return g_err; }

int fk_find_conn(const int8_t* spins, int64_t batch, int32_t n_sites, int32_t kind,
                 const int32_t* conn_a, const int32_t* conn_b, const double* conn_mel, int32_t n_conn,
                 const int32_t* diag_a, const int32_t* diag_b, const double* diag_j, int32_t n_diag,
                 int32_t diag_rule, int8_t* conn_out, double* mel_out, uint8_t* use_out, int32_t* n_used_out,
                 void* stream) {
  FK_REQUIRE(batch >= 0 && n_sites > 0 && n_sites <= 65536 && n_conn >= 1, "fk_find_conn: bad sizes");
  if (batch == 0) return FK_OK;
  FK_REQUIRE(spins && conn_a && conn_b && conn_mel && mel_out && use_out, "fk_find_conn: null pointer");
  FK_REQUIRE(kind == FK_CONN_SWAP || kind == FK_CONN_SWAP_ZERO_PAD || kind == FK_CONN_FLIP,
             "fk_find_conn: bad kind %d", kind);
  FK_REQUIRE(n_diag == 0 || (diag_a && diag_b && diag_j), "fk_find_conn: null diagonal table");
  FK_REQUIRE(diag_rule == FK_DIAG_SUM_F64 || diag_rule == FK_DIAG_ISING_F32, "fk_find_conn: bad diag_rule %d",
             diag_rule);
  if (batch == 0) return FK_OK;
  FK_REQUIRE(batch <= 0x7fffffff, "fk_find_conn: batch too large");
  // tile of samples per workgroup: the diagonal table and their spin rows
  // fill at most 64 KB of LDS
  FK_REQUIRE(n_diag <= 2048 && 16 * (int64_t)n_diag + n_sites <= 65536, "fk_find_conn: table too large for LDS");
  const int tile = (int)std::min<int64_t>(64, (65536 - 16 * (int64_t)n_diag) / n_sites);
  const size_t lds = 16 * (size_t)n_diag + (size_t)tile * n_sites;
  hipLaunchKernelGGL(find_conn_kernel, dim3((unsigned)div_up(batch, tile)), dim3(1024), lds, (hipStream_t)stream,
                     spins, batch, n_sites, tile, kind, conn_a, conn_b, conn_mel, n_conn, diag_a, diag_b, diag_j,
                     n_diag, diag_rule, conn_out, mel_out, use_out, n_used_out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_compact_scratch_bytes(int64_t batch) {
  (void)batch;
  return 0;
}

int fk_compact_connections(const uint8_t* use, int64_t batch, int32_t n_conn, int64_t* offsets, int32_t* pairs,
                           void* scratch, void* stream) {
  (void)scratch;
  FK_REQUIRE(batch >= 0 && n_conn >= 1, "fk_compact_connections: bad sizes");
  FK_REQUIRE(offsets, "fk_compact_connections: null offsets");
  hipStream_t s = (hipStream_t)stream;
  if (batch == 0) {
    hipLaunchKernelGGL(zero_i64_kernel, dim3(1), dim3(1), 0, s, offsets);
    FK_CHECK_LAUNCH();
    return FK_OK;
  }
  FK_REQUIRE(use && pairs, "fk_compact_connections: null pointer");
  const int tb = 256;
  hipLaunchKernelGGL(count_used_kernel, dim3(div_up(batch, 64)), dim3(tb), 0, s, use, batch, n_conn, offsets);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, offsets, batch);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(write_pairs_kernel, dim3(div_up(batch, 64)), dim3(tb), 0, s, use, batch, n_conn, offsets,
                     pairs);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int64_t fk_bucket_scratch_bytes(int64_t batch, int32_t n_buckets) {
  if (batch < 0 || n_buckets < 1) return -1;
  return (int64_t)n_buckets * batch * 4 + ((int64_t)n_buckets * 8 + 15) / 16 * 16;
}

int fk_bucket_connections(const uint8_t* use, int64_t batch, int32_t n_conn, const int32_t* bucket_of,
                          int32_t n_buckets, const int64_t* offsets, int32_t* pairs, int64_t* out_index,
                          int64_t* bucket_offsets, void* scratch, void* stream) {
  FK_REQUIRE(batch >= 1 && n_conn >= 1 && n_buckets >= 1 && n_buckets <= BK_MAX,
             "fk_bucket_connections: bad sizes (1 <= n_buckets <= %d)", BK_MAX);
  FK_REQUIRE(use && bucket_of && offsets && pairs && out_index && bucket_offsets && scratch,
             "fk_bucket_connections: null pointer");
  FK_REQUIRE(batch <= 0x7fffffff, "fk_bucket_connections: batch too large");
  hipStream_t s = (hipStream_t)stream;
  int32_t* counts = reinterpret_cast<int32_t*>(scratch);
  int64_t* totals = reinterpret_cast<int64_t*>(reinterpret_cast<uint8_t*>(scratch) +
                                               ((int64_t)n_buckets * batch * 4 + 15) / 16 * 16);
  hipLaunchKernelGGL(bucket_count_kernel, dim3(div_up(batch, 64)), dim3(64), 0, s, use, batch, n_conn, bucket_of,
                     n_buckets, counts);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(bucket_scan_kernel, dim3(n_buckets), dim3(1024), 0, s, counts, batch, totals);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(bucket_offsets_kernel, dim3(1), dim3(1), 0, s, totals, n_buckets, bucket_offsets);
  FK_CHECK_LAUNCH();
  hipLaunchKernelGGL(bucket_write_kernel, dim3(div_up(batch, 64)), dim3(64), 0, s, use, batch, n_conn, bucket_of,
                     n_buckets, counts, bucket_offsets, offsets, pairs, out_index);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_local_energy(const int64_t* offsets, const int32_t* pairs, const float* logpsi, const double* mel,
                    const double* mel_im, int64_t batch, double* eloc_out, void* stream) {
  FK_REQUIRE(offsets && pairs && logpsi && mel && eloc_out, "fk_local_energy: null pointer");
  if (batch <= 0) return FK_OK;
  hipLaunchKernelGGL(local_energy_kernel, dim3(div_up(batch, 4)), dim3(256), 0, (hipStream_t)stream, offsets,
                     pairs, logpsi, mel, mel_im, batch, eloc_out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_local_energy_dense(const float* logpsi, const double* mel, const double* mel_im, int32_t n_conn,
                          int64_t batch, double* eloc_out, void* stream) {
  FK_REQUIRE(logpsi && mel && eloc_out, "fk_local_energy_dense: null pointer");
  if (batch <= 0) return FK_OK;
  hipLaunchKernelGGL(local_energy_dense_kernel, dim3(div_up(batch, 256)), dim3(256), 0, (hipStream_t)stream,
                     logpsi, mel, mel_im, n_conn, batch, eloc_out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_sigma_z(const int8_t* spins, int64_t batch, int32_t n_sites, int32_t absolute, double* out, void* stream) {
  FK_REQUIRE(spins && out, "fk_sigma_z: null pointer");
  FK_REQUIRE(batch >= 0 && n_sites > 0, "fk_sigma_z: bad sizes");
  if (batch == 0) return FK_OK;
  hipLaunchKernelGGL(sigma_z_kernel, dim3(div_up(batch, 4)), dim3(256), 0, (hipStream_t)stream, spins, batch,
                     n_sites, absolute, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_energy_moments(const double* eloc, int64_t batch, double* out, void* stream) {
  FK_REQUIRE(eloc && out, "fk_energy_moments: null pointer");
  FK_REQUIRE(batch > 0, "fk_energy_moments: empty batch");
  hipLaunchKernelGGL(energy_moments_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, eloc, batch, out);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_ar_draw_site(const float* cond_logp, const double* uniforms, int64_t batch, int32_t n_sites, int32_t site,
                    int8_t* spins, void* stream) {
  FK_REQUIRE(cond_logp && uniforms && spins, "fk_ar_draw_site: null pointer");
  FK_REQUIRE(site >= 0 && site < n_sites, "fk_ar_draw_site: bad site");
  if (batch <= 0) return FK_OK;
  hipLaunchKernelGGL(ar_draw_site_kernel, dim3(div_up(batch, 256)), dim3(256), 0, (hipStream_t)stream, cond_logp,
                     uniforms, batch, n_sites, site, spins);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // extern "C"
