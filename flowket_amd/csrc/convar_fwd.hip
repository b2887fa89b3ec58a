// Fused ConvNetAutoregressive2D forward (machines/conv_net_autoregressive_2D.py
// :33-74 + machines/abstract_machine.py:31-71) for gfx950.
//
// One workgroup evaluates a tile of T configurations through the whole network
// with every activation resident in LDS (buffers V0, H0, V1, H1, CAT of
// T*H*W pixels x C channels).  Each convolution is an implicit GEMM on MFMA:
// A = weights (16 output channels x k input slots), B = activations (k slots
// x 16 pixels) read from LDS at the tap offset, D = 16 channels x 16 pixels.
// The 1x1 convolutions (x, y, head) take their B operand straight from the
// registers of the preceding convolution's D fragment: input slots are
// numbered so that a D fragment *is* a B fragment (see convar_layout.h).
// Zero padding, DownShift and RightShift are index arithmetic (tap validity
// masks and shifted write targets), never materialised.
// HBM traffic: H*W int8 spins in, 8 bytes out per configuration.
#include <map>
#include <mutex>

#include "fk_common.h"
#include "convar_layout.h"
#include "convar_dev.h"

namespace fk {

// ---------------------------------------------------------------------------
// all-zero pixels after each activation buffer (one per bank placement of
// the 8-pixel swizzle period)
constexpr int ZERO_PIX = 8;

// LDS addressing.  A pixel holds C channels as 16-byte chunks; chunk indices
// are XOR-swizzled with the pixel index so that 16 lanes reading the same
// chunk of 16 consecutive pixels hit distinct banks.
template <int PREC, int C>
struct Lds {
  static constexpr int ESZ = Prec<PREC>::ESZ;
  static constexpr int PIXB = C * ESZ;
  static constexpr int CPP = PIXB / 16;  // chunks per pixel
  static constexpr int ROWB = 128;       // bytes per LDS bank row (32 banks x 4 B)
  static constexpr int PPR = PIXB >= ROWB ? 1 : ROWB / PIXB;
  static constexpr bool SWZ = (CPP & (CPP - 1)) == 0;
  __device__ static __forceinline__ int chunk(int p, int c) {
    const int sw = SWZ ? ((p / PPR) & (CPP - 1)) : 0;
    return p * PIXB + ((c ^ sw) << 4);
  }
  // byte offset of the 4-channel quad q (channels 4q..4q+3) of pixel p
  __device__ static __forceinline__ int quad(int p, int q) {
    if constexpr (Prec<PREC>::H16) {
      const int ks = q >> 3, w = q & 7;
      return chunk(p, 4 * ks + (w & 3)) + ((w >> 2) << 3);
    } else {
      return chunk(p, (q & 3) * (C / 16) + (q >> 2));
    }
  }
  __device__ static __forceinline__ void store_quad(uint8_t* buf, int p, int q, f32x4 v) {
    if constexpr (Prec<PREC>::H16) {
      *reinterpret_cast<typename Prec<PREC>::quad*>(buf + quad(p, q)) = to_quad<PREC>(v);
    } else {
      *reinterpret_cast<f32x4*>(buf + quad(p, q)) = v;
    }
  }
  __device__ static __forceinline__ f32x4 load_quad(const uint8_t* buf, int p, int q) {
    if constexpr (Prec<PREC>::H16) {
      return from_quad<PREC>(*reinterpret_cast<const typename Prec<PREC>::quad*>(buf + quad(p, q)));
    } else {
      return *reinterpret_cast<const f32x4*>(buf + quad(p, q));
    }
  }
  // Chunk of k-step 0 for lane group g (the address that precomputed tap
  // offsets point at); further k-steps / chunks are reached by XOR.
  __device__ static __forceinline__ int base_chunk(int g) {
    return Prec<PREC>::H16 ? g : g * (C / 16);
  }
  template <int KS>
  __device__ static __forceinline__ void load_bfrags_at(const uint8_t* buf, uint32_t off,
                                                        typename Prec<PREC>::frag (&b)[KS]) {
    if constexpr (Prec<PREC>::H16) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        b[ks] = *reinterpret_cast<const typename Prec<PREC>::frag*>(buf + (off ^ (64u * ks)));
    } else {
#pragma unroll
      for (int j = 0; j < KS / 4; ++j) {
        f32x4 v = *reinterpret_cast<const f32x4*>(buf + (off ^ (16u * j)));
        b[4 * j + 0] = v.x;
        b[4 * j + 1] = v.y;
        b[4 * j + 2] = v.z;
        b[4 * j + 3] = v.w;
      }
    }
  }
  // byte offset of M-tile mt's quad of this lane from the lane's own
  // k-step-0 chunk offset
  __device__ static __forceinline__ uint32_t self_quad(uint32_t off, int mt) {
    if constexpr (Prec<PREC>::H16) return (off ^ (64u * (mt >> 1))) + 8u * (mt & 1);
    else return off ^ (16u * mt);
  }
  __device__ static __forceinline__ void store_quad_at(uint8_t* buf, uint32_t off, f32x4 v) {
    if constexpr (Prec<PREC>::H16) {
      *reinterpret_cast<typename Prec<PREC>::quad*>(buf + off) = to_quad<PREC>(v);
    } else {
      *reinterpret_cast<f32x4*>(buf + off) = v;
    }
  }
  __device__ static __forceinline__ f32x4 load_quad_at(const uint8_t* buf, uint32_t off) {
    if constexpr (Prec<PREC>::H16) {
      return from_quad<PREC>(*reinterpret_cast<const typename Prec<PREC>::quad*>(buf + off));
    } else {
      return *reinterpret_cast<const f32x4*>(buf + off);
    }
  }
  // 16-bit precisions: the quads of M-tiles 2j and 2j+1 share the 16-byte
  // chunk at off ^ 64j, so they are written and read as one access.
  __device__ static __forceinline__ void store_pair_at(uint8_t* buf, uint32_t off, f32x4 lo, f32x4 hi) {
    using Q = typename Prec<PREC>::quad;
    const Q a = to_quad<PREC>(lo), b = to_quad<PREC>(hi);
    typename Prec<PREC>::frag f;
    f[0] = a[0];
    f[1] = a[1];
    f[2] = a[2];
    f[3] = a[3];
    f[4] = b[0];
    f[5] = b[1];
    f[6] = b[2];
    f[7] = b[3];
    *reinterpret_cast<typename Prec<PREC>::frag*>(buf + off) = f;
  }
  __device__ static __forceinline__ void load_pair_at(const uint8_t* buf, uint32_t off, f32x4& lo, f32x4& hi) {
    const typename Prec<PREC>::frag f = *reinterpret_cast<const typename Prec<PREC>::frag*>(buf + off);
    lo = f32x4{(float)f[0], (float)f[1], (float)f[2], (float)f[3]};
    hi = f32x4{(float)f[4], (float)f[5], (float)f[6], (float)f[7]};
  }
  // All B fragments (one per k-step) of pixel p for lane group g.
  template <int KS>
  __device__ static __forceinline__ void load_bfrags(const uint8_t* buf, int p, int g,
                                                     typename Prec<PREC>::frag (&b)[KS]) {
    if constexpr (Prec<PREC>::H16) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        b[ks] = *reinterpret_cast<const typename Prec<PREC>::frag*>(buf + chunk(p, 4 * ks + g));
    } else {
#pragma unroll
      for (int j = 0; j < KS / 4; ++j) {
        f32x4 v = *reinterpret_cast<const f32x4*>(buf + chunk(p, g * (C / 16) + j));
        b[4 * j + 0] = v.x;
        b[4 * j + 1] = v.y;
        b[4 * j + 2] = v.z;
        b[4 * j + 3] = v.w;
      }
    }
  }
};

struct FwdArgs {
  const void* w;
  const float* aux;
  int H, W, HW, nblocks, T, npix;
  const int8_t* spins;
  const int32_t* pairs;
  const int64_t* m_dev;
  int64_t m_count;
  int conn_kind;
  const int32_t* conn_a;
  const int32_t* conn_b;
  float* logpsi;
  float* cond;
};

// Per-tile LDS byte offsets (from the buffer start) of the 12 causal taps
// k = (dr+2)*4 + (dc+2) of this lane's pixel, two 16-bit offsets per
// register: every activation buffer is below 64 KB.
template <int MAXT>
struct TapOff {
  uint32_t v[MAXT][6];
  __device__ __forceinline__ uint32_t get(int u, int k) const {
    return (k & 1) ? (v[u][k >> 1] >> 16) : (v[u][k >> 1] & 0xffffu);
  }
};

// acc[u][mt] += sum over taps of W[tap][mt] x B(pixel + tap offset).
// toff.get(u, k) is this lane's byte offset of k-step 0 at tap k of tile u,
// pointing at the all-zero pixel for padding taps -- tiles past
// the wave's count hold only such offsets, so every tile is computed without
// a branch (their epilogues are skipped).
// Software pipeline: the weights (read from L2) are requested WD taps ahead
// and the B fragments of all tiles BD taps ahead (depths bounded by the
// registers they take); scheduling fences keep the compiler from sinking
// those requests next to their first use.
template <int PREC, int C, int MAXT, int KIND, int MTO>
__device__ __forceinline__ void conv_pass(f32x4 (&acc)[MAXT][MTO], const uint8_t* buf,
                                          const typename Prec<PREC>::frag* __restrict__ wg,
                                          const TapOff<MAXT>& toff, int lane) {
  using P = Prec<PREC>;
  using frag = typename P::frag;
  using L = Lds<PREC, C>;
  constexpr int KS = C / P::KW;
  constexpr int NTAP = Taps<KIND>::N;
  // prefetch depth bounded by the registers a tap's fragments take
  constexpr int FR = (int)(sizeof(frag) / 4);
  constexpr int WREGS = MTO * KS * FR, BREGS = MAXT * KS * FR;
  constexpr int WD = WREGS <= 16 ? 2 : (WREGS <= 32 ? 1 : 0);  // weight prefetch distance (taps)
  constexpr int BD = BREGS <= 32 ? 1 : 0;                       // B prefetch distance (taps)
  frag w[WD + 1][MTO][KS];
  frag b[BD + 1][MAXT][KS];
  auto load_w = [&](int tap, frag (&dst)[MTO][KS]) {
#pragma unroll
    for (int mt = 0; mt < MTO; ++mt)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) dst[mt][ks] = wg[((tap * MTO + mt) * KS + ks) * 64 + lane];
  };
  auto load_b = [&](int tap, frag (&dst)[MAXT][KS]) {
#pragma unroll
    for (int u = 0; u < MAXT; ++u) L::template load_bfrags_at<KS>(buf, toff.get(u, Taps<KIND>::bit(tap)), dst[u]);
  };
#pragma unroll
  for (int t = 0; t < WD && t < NTAP; ++t) load_w(t, w[t]);
  if constexpr (BD > 0) load_b(0, b[0]);
#pragma unroll
  for (int tap = 0; tap < NTAP; ++tap) {
    if (tap + WD < NTAP) load_w(tap + WD, w[(tap + WD) % (WD + 1)]);
    if constexpr (BD > 0) {
      if (tap + 1 < NTAP) load_b(tap + 1, b[(tap + 1) % (BD + 1)]);
    } else {
      load_b(tap, b[0]);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < MAXT; ++u)
#pragma unroll
      for (int mt = 0; mt < MTO; ++mt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
          acc[u][mt] = P::mfma(w[tap % (WD + 1)][mt][ks], b[tap % (BD + 1)][u][ks], acc[u][mt]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

__device__ __forceinline__ f32x4 bias4(const float* b) { return *reinterpret_cast<const f32x4*>(b); }

// Register-resident weights of one convolution (all taps) for the tile-outer
// schedule.
template <int PREC, int C, int KIND, int MTO>
struct ConvW {
  static constexpr int KS = C / Prec<PREC>::KW;
  static constexpr int N = Taps<KIND>::N;
  typename Prec<PREC>::frag w[N][MTO][KS];
  __device__ __forceinline__ void load(const typename Prec<PREC>::frag* __restrict__ wg, int lane) {
#pragma unroll
    for (int t = 0; t < N; ++t)
#pragma unroll
      for (int mt = 0; mt < MTO; ++mt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) w[t][mt][ks] = wg[((t * MTO + mt) * KS + ks) * 64 + lane];
  }
};

// One 16-pixel tile of a convolution with register-resident weights: every
// tap's B fragments are read from LDS first, then the MFMAs run back to back.
template <int PREC, int C, int KIND, int MTO, int MAXT>
__device__ __forceinline__ void conv_tile(f32x4 (&acc)[MTO], const uint8_t* buf, const ConvW<PREC, C, KIND, MTO>& cw,
                                          const TapOff<MAXT>& to, int u) {
  using P = Prec<PREC>;
  using L = Lds<PREC, C>;
  constexpr int KS = C / P::KW;
  constexpr int N = Taps<KIND>::N;
  typename P::frag b[N][KS];
#pragma unroll
  for (int t = 0; t < N; ++t) L::template load_bfrags_at<KS>(buf, to.get(u, Taps<KIND>::bit(t)), b[t]);
#pragma unroll
  for (int t = 0; t < N; ++t)
#pragma unroll
    for (int mt = 0; mt < MTO; ++mt)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) acc[mt] = P::mfma(cw.w[t][mt][ks], b[t][ks], acc[mt]);
}

// SCHED 0: taps outer -- each tap's weights are loaded once and serve all the
// wave's tiles (accumulators of every tile live in registers).
// SCHED 1: tiles outer -- a 3x3 convolution's weights stay in registers for
// the whole phase and the wave walks its tiles one by one (fits bf16 C=32 and
// fp32 C=16).
template <int PREC, int C, int MAXT, int ACT, int NWAVE, int SCHED>
__global__ __launch_bounds__(64 * NWAVE) void convar_fwd_kernel(FwdArgs a) {
  using P = Prec<PREC>;
  using frag = typename P::frag;
  using L = Lds<PREC, C>;
  constexpr int MT = C / 16;
  constexpr int KS = C / P::KW;
  constexpr int MTH = (C / 2 + 15) / 16;
  constexpr int PIXB = L::PIXB;
  constexpr bool TILE_OUTER = SCHED == 1;
  const ConvarLayout lay(C, PREC);

  extern __shared__ __align__(16) uint8_t lds[];
  // npix: pixels covered by the 16-pixel tiles; NPX = T*HW real pixels.
  // Each buffer holds the NPX pixels plus ZERO_PIX all-zero pixels that padding
  // taps and the tile lanes past NPX read: a lane whose tap target q falls
  // outside reads zero pixel NPX + ((q - NPX) & 7), which has the bank
  // placement q itself would have, so border tiles stay conflict-free.
  const int npix = a.npix;
  const int NPX = a.T * a.HW;
  const int BUFB = (NPX + ZERO_PIX) * PIXB;
  uint8_t* V0 = lds;
  uint8_t* H0 = V0 + BUFB;
  uint8_t* V1 = H0 + BUFB;
  uint8_t* H1 = V1 + BUFB;
  uint8_t* CAT = H1 + BUFB;
  int8_t* sp = reinterpret_cast<int8_t*>(CAT + BUFB);
  // per-site log-probabilities of the final block: V1 is neither read nor
  // written by the final block, so they reuse its space
  float2* part = reinterpret_cast<float2*>(V1);

  const int lane = threadIdx.x & 63;
  // wave index and tile count are wave-uniform: keep them in SGPRs so tile
  // predicates are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int NW = blockDim.x >> 6;
  const int n16 = lane & 15, g = lane >> 4;
  const int H = a.H, W = a.W, HW = a.HW, T = a.T;
  const int ntiles = npix >> 4;
  const int nt = (ntiles > wave) ? (ntiles - wave + NW - 1) / NW : 0;

  int pix[MAXT], pi[MAXT], pj[MAXT];
  uint32_t msk[MAXT];
  TapOff<MAXT> toff;
  bool inp[MAXT];
#pragma unroll
  for (int u = 0; u < MAXT; ++u) {
    const int p = (wave + u * NW) * 16 + n16;
    const bool in = (u < nt) && (p < T * HW);
    const int t = in ? p / HW : 0;
    const int rem = in ? p - t * HW : 0;
    const int i = rem / W, j = rem - (rem / W) * W;
    uint32_t m = 0;
    if (in) {
#pragma unroll
      for (int dr = -2; dr <= 0; ++dr)
#pragma unroll
        for (int dc = -2; dc <= 1; ++dc)
          if (i + dr >= 0 && j + dc >= 0 && j + dc < W) m |= 1u << ((dr + 2) * 4 + (dc + 2));
    }
    pix[u] = p;
    pi[u] = i;
    pj[u] = j;
    msk[u] = m;
    inp[u] = in;
#pragma unroll
    for (int k2 = 0; k2 < 6; ++k2) {
      uint32_t two = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = 2 * k2 + h;
        const int dr = k / 4 - 2, dc = k % 4 - 2;
        const int q = p + dr * W + dc;
        const int pp = ((m >> k) & 1u) ? q : NPX + ((q - NPX) & (ZERO_PIX - 1));
        two |= (uint32_t)L::chunk(pp, L::base_chunk(g)) << (16 * h);
      }
      toff.v[u][k2] = two;
    }
  }
  for (int z = threadIdx.x; z < 5 * ZERO_PIX * (PIXB / 4); z += blockDim.x) {
    const int buf = z / (ZERO_PIX * PIXB / 4), w4 = z - buf * (ZERO_PIX * PIXB / 4);
    reinterpret_cast<uint32_t*>(lds + buf * BUFB + NPX * PIXB)[w4] = 0u;
  }

  int64_t M = a.m_dev ? *a.m_dev : a.m_count;
  if (M > a.m_count) M = a.m_count;
  const int64_t ngroups = (M + T - 1) / T;
  const int nb = a.nblocks;
  const frag* wall = reinterpret_cast<const frag*>(a.w);

  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    // ---- configurations -> LDS (int8 spins; connection applied on the fly)
    for (int p = threadIdx.x; p < NPX; p += blockDim.x) {
      int8_t s = 0;
      const int t = p / HW;
      const int64_t m = grp * T + t;
      if (m < M) {
        const int q = p - t * HW;
        int64_t b = m;
        int c = 0;
        if (a.pairs) {
          b = a.pairs[2 * m];
          c = a.pairs[2 * m + 1];
        }
        const int8_t* row = a.spins + b * HW;
        int8_t v = row[q];
        if (c > 0) {
          const int ca = a.conn_a[c];
          if (a.conn_kind == FK_CONN_FLIP) {
            if (q == ca) v = (int8_t)-v;
          } else if (ca >= 0) {
            const int cb = a.conn_b[c];
            if (q == ca) v = row[cb];
            else if (q == cb) v = row[ca];
          }
        }
        s = v;
      }
      sp[p] = s;
    }
    __syncthreads();

    for (int blk = 0; blk < nb; ++blk) {
      const int kind = block_kind(blk, nb);
      const uint8_t* Vin = (kind == BK_B) ? V1 : V0;
      const uint8_t* Hin = (kind == BK_B) ? H1 : H0;
      const float* bias = a.aux + lay.aux_bias(blk);
      const float* by = bias + 2 * C + C / 2;
      const float* bx = bias + 2 * C;

      frag wy[MTH][KS], wx[MTH][KS];
#pragma unroll
      for (int mh = 0; mh < MTH; ++mh)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          wy[mh][ks] = wall[(lay.off_y(blk) + mh * KS + ks) * 64 + lane];
          wx[mh][ks] = wall[(lay.off_x(blk) + mh * KS + ks) * 64 + lane];
        }

      // -- block-0 convolutions on the 1-channel input (VALU)
      auto valu_v = [&](int u, f32x4 (&acc)[MT]) {
        const float* kv = a.aux + lay.aux_kv0(nb);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
          const bool valid = (msk[u] >> Taps<0>::bit(tap)) & 1u;
          const float s = valid ? (float)sp[pix[u] + Taps<0>::dr(tap) * W + Taps<0>::dc(tap)] : 0.0f;
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            const f32x4 k = *reinterpret_cast<const f32x4*>(kv + tap * C + 16 * mt + 4 * g);
            acc[mt] += k * s;
          }
        }
      };
      auto valu_h = [&](int u, f32x4 (&acc)[MT]) {
        const float* kh = a.aux + lay.aux_kh0(nb);
#pragma unroll
        for (int tap = 0; tap < 3; ++tap) {
          const bool valid = (msk[u] >> Taps<1>::bit(tap)) & 1u;
          const float s = valid ? (float)sp[pix[u] + Taps<1>::dc(tap)] : 0.0f;
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) {
            const f32x4 k = *reinterpret_cast<const f32x4*>(kh + tap * C + 16 * mt + 4 * g);
            acc[mt] += k * s;
          }
        }
      };
      // -- epilogue of the vertical kxk: vertical output + y = act(1x1(
      //    DownShift(act(v')))) written one row down
      auto epi_v = [&](int u, const f32x4 (&acc)[MT]) {
        f32x4 av[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) av[mt] = act4<ACT>(acc[mt]);
        if (inp[u]) {
          if constexpr (Prec<PREC>::H16) {
#pragma unroll
            for (int j = 0; j < MT / 2; ++j) {
              const uint32_t so = toff.get(u, 10) ^ (64u * j);
              if (kind == BK_FIRST) L::store_pair_at(V0, so, av[2 * j], av[2 * j + 1]);
              else if (kind == BK_A) L::store_pair_at(V1, so, av[2 * j], av[2 * j + 1]);
              else if (kind == BK_B) {
                f32x4 rl, rh;
                L::load_pair_at(V0, so, rl, rh);
                L::store_pair_at(V0, so, act4<ACT>(rl + acc[2 * j]), act4<ACT>(rh + acc[2 * j + 1]));
              }
            }
          } else {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
              const uint32_t so = L::self_quad(toff.get(u, 10), mt);
              if (kind == BK_FIRST) L::store_quad_at(V0, so, av[mt]);
              else if (kind == BK_A) L::store_quad_at(V1, so, av[mt]);
              else if (kind == BK_B) L::store_quad_at(V0, so, act4<ACT>(L::load_quad_at(V0, so) + acc[mt]));
            }
          }
        }
        frag bf[KS];
        d_to_b<PREC, MT, KS>(av, bf);
        const bool down = pi[u] + 1 < H;
        const int tgt = down ? pix[u] + W : pix[u] - (H - 1) * W;
#pragma unroll
        for (int mh = 0; mh < MTH; ++mh) {
          const f32x4 b0 = bias4(by + 16 * mh + 4 * g);
          f32x4 y = b0;
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) y = P::mfma(wy[mh][ks], bf[ks], y);
          y = act4<ACT>(down ? y : b0);
          if (inp[u] && 16 * mh + 4 * g < C / 2) L::store_quad(CAT, tgt, C / 8 + 4 * mh + g, y);
        }
      };
      // -- epilogue of the horizontal 1xk: x = act(1x1(act(.))) into CAT (the
      //    final block's RightShift folded into the write target)
      auto epi_h = [&](int u, const f32x4 (&acc)[MT]) {
        f32x4 ax[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) ax[mt] = act4<ACT>(acc[mt]);
        frag bf[KS];
        d_to_b<PREC, MT, KS>(ax, bf);
        bool keep = true;
        int tgt = pix[u];
        if (kind == BK_FINAL) {
          keep = pj[u] + 1 < W;
          tgt = keep ? pix[u] + 1 : pix[u] - (W - 1);
        }
#pragma unroll
        for (int mh = 0; mh < MTH; ++mh) {
          const f32x4 b0 = bias4(bx + 16 * mh + 4 * g);
          f32x4 x = b0;
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) x = P::mfma(wx[mh][ks], bf[ks], x);
          x = act4<ACT>(keep ? x : b0);
          if (inp[u] && 16 * mh + 4 * g < C / 2) L::store_quad(CAT, tgt, 4 * mh + g, x);
        }
      };

      // ================= phase VH: vertical stack + horizontal 1xk =========
      if constexpr (TILE_OUTER) {
        {
          ConvW<PREC, C, 0, MT> wv;
          if (kind != BK_FIRST) wv.load(wall + lay.off_v(blk) * 64, lane);
#pragma unroll
          for (int u = 0; u < MAXT; ++u) {
            if (u < nt) {
              f32x4 acc[MT];
#pragma unroll
              for (int mt = 0; mt < MT; ++mt) acc[mt] = bias4(bias + 16 * mt + 4 * g);
              if (kind == BK_FIRST) valu_v(u, acc);
              else conv_tile<PREC, C, 0, MT>(acc, Vin, wv, toff, u);
              epi_v(u, acc);
            }
          }
        }
        {
          ConvW<PREC, C, 1, MT> wh;
          if (kind != BK_FIRST) wh.load(wall + lay.off_h(blk) * 64, lane);
#pragma unroll
          for (int u = 0; u < MAXT; ++u) {
            if (u < nt) {
              f32x4 acch[MT];
#pragma unroll
              for (int mt = 0; mt < MT; ++mt) acch[mt] = bias4(bias + C + 16 * mt + 4 * g);
              if (kind == BK_FIRST) valu_h(u, acch);
              else conv_tile<PREC, C, 1, MT>(acch, Hin, wh, toff, u);
              epi_h(u, acch);
            }
          }
        }
      } else {
        {
          f32x4 acc[MAXT][MT];
#pragma unroll
          for (int u = 0; u < MAXT; ++u)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[u][mt] = bias4(bias + 16 * mt + 4 * g);
          if (kind == BK_FIRST) {
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
              if (u < nt) valu_v(u, acc[u]);
          } else {
            conv_pass<PREC, C, MAXT, 0, MT>(acc, Vin, wall + lay.off_v(blk) * 64, toff, lane);
          }
#pragma unroll
          for (int u = 0; u < MAXT; ++u)
            if (u < nt) epi_v(u, acc[u]);
        }
        {
          f32x4 acc[MAXT][MT];
#pragma unroll
          for (int u = 0; u < MAXT; ++u)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[u][mt] = bias4(bias + C + 16 * mt + 4 * g);
          if (kind == BK_FIRST) {
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
              if (u < nt) valu_h(u, acc[u]);
          } else {
            conv_pass<PREC, C, MAXT, 1, MT>(acc, Hin, wall + lay.off_h(blk) * 64, toff, lane);
          }
#pragma unroll
          for (int u = 0; u < MAXT; ++u)
            if (u < nt) epi_h(u, acc[u]);
        }
      }
      __syncthreads();

      // ================= phase H3: kxk on the concatenation ================
      frag whd[KS];
      f32x4 hb = f32x4{0.f, 0.f, 0.f, 0.f};
      if (kind == BK_FINAL) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) whd[ks] = wall[(lay.head_base(nb) + ks) * 64 + lane];
        hb = load_bias4(a.aux + lay.aux_head_bias(nb), 4 * g, 4);
      }
      auto epi_c = [&](int u, const f32x4 (&acc)[MT]) {
        if (kind != BK_FINAL) {
          if (inp[u]) {
            if constexpr (Prec<PREC>::H16) {
#pragma unroll
              for (int j = 0; j < MT / 2; ++j) {
                const uint32_t so = toff.get(u, 10) ^ (64u * j);
                if (kind == BK_FIRST) L::store_pair_at(H0, so, act4<ACT>(acc[2 * j]), act4<ACT>(acc[2 * j + 1]));
                else if (kind == BK_A) L::store_pair_at(H1, so, act4<ACT>(acc[2 * j]), act4<ACT>(acc[2 * j + 1]));
                else {
                  f32x4 rl, rh;
                  L::load_pair_at(H0, so, rl, rh);
                  L::store_pair_at(H0, so, act4<ACT>(rl + acc[2 * j]), act4<ACT>(rh + acc[2 * j + 1]));
                }
              }
            } else {
#pragma unroll
              for (int mt = 0; mt < MT; ++mt) {
                const uint32_t so = L::self_quad(toff.get(u, 10), mt);
                if (kind == BK_FIRST) L::store_quad_at(H0, so, act4<ACT>(acc[mt]));
                else if (kind == BK_A) L::store_quad_at(H1, so, act4<ACT>(acc[mt]));
                else L::store_quad_at(H0, so, act4<ACT>(L::load_quad_at(H0, so) + acc[mt]));
              }
            }
          }
          return;
        }
        // head: 1x1 C->4 on act(h'), complex logits, 2-way log-normalise,
        // gather at the actual spin (abstract_machine.py:34-37,54-57)
        f32x4 ah[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) ah[mt] = act4<ACT>(acc[mt]);
        frag bf[KS];
        d_to_b<PREC, MT, KS>(ah, bf);
        f32x4 z = hb;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) z = P::mfma(whd[ks], bf[ks], z);
        if (g == 0 && inp[u]) {
          // z = (Re z0, Re z1, Im z0, Im z1) for spin values (+1, -1)
          const float mx = fmaxf(z.x, z.y);
          const float lse = mx + 0.5f * log1pf(expf(-2.0f * fabsf(z.x - z.y)));
          const bool up = sp[pix[u]] > 0;
          part[pix[u]] = make_float2((up ? z.x : z.y) - lse, up ? z.z : z.w);
          if (a.cond) {
            const int t = pix[u] / HW;
            const int64_t m = grp * T + t;
            if (m < M) {
              float* o = a.cond + (m * HW + (pix[u] - t * HW)) * 2;
              o[0] = 2.0f * (z.x - lse);
              o[1] = 2.0f * (z.y - lse);
            }
          }
        }
      };
      if constexpr (TILE_OUTER) {
        ConvW<PREC, C, 2, MT> wc;
        wc.load(wall + lay.off_c(blk) * 64, lane);
#pragma unroll
        for (int u = 0; u < MAXT; ++u) {
          if (u < nt) {
            f32x4 acc[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[mt] = bias4(bias + 3 * C + 16 * mt + 4 * g);
            conv_tile<PREC, C, 2, MT>(acc, CAT, wc, toff, u);
            epi_c(u, acc);
          }
        }
      } else {
        f32x4 acc[MAXT][MT];
#pragma unroll
        for (int u = 0; u < MAXT; ++u)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) acc[u][mt] = bias4(bias + 3 * C + 16 * mt + 4 * g);
        conv_pass<PREC, C, MAXT, 2, MT>(acc, CAT, wall + lay.off_c(blk) * 64, toff, lane);
#pragma unroll
        for (int u = 0; u < MAXT; ++u)
          if (u < nt) epi_c(u, acc[u]);
      }
      __syncthreads();
    }

    // ---- per-configuration site sums (deterministic tree per wave)
    for (int t = wave; t < T; t += NW) {
      float sr = 0.0f, si = 0.0f;
      for (int q = lane; q < HW; q += 64) {
        const float2 v = part[t * HW + q];
        sr += v.x;
        si += v.y;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
      }
      const int64_t m = grp * T + t;
      if (lane == 0 && m < M) {
        a.logpsi[2 * m] = sr;
        a.logpsi[2 * m + 1] = si;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
__global__ void pack_gather_kernel(const float* __restrict__ src, const int32_t* __restrict__ map, int64_t n,
                                   int prec, void* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t k = map[i];
  const float v = k >= 0 ? src[k] : 0.0f;
  if (prec == FK_PREC_BF16) reinterpret_cast<__bf16*>(dst)[i] = (__bf16)v;
  else if (prec == FK_PREC_FP16) reinterpret_cast<_Float16*>(dst)[i] = (_Float16)v;
  else reinterpret_cast<float*>(dst)[i] = v;
}

// ---------------------------------------------------------------------------
struct FwdPlan {
  int T, npix, maxt, nwave;
  size_t lds;
};

// five activation buffers of P + ZERO_PIX pixels and the int8 spins (the
// per-site log-probabilities live in V1)
static size_t fwd_lds_bytes(int P, int pixb) { return (size_t)(P + ZERO_PIX) * pixb * 5 + (size_t)P; }

// One launch shape of the forward: loop order and waves per workgroup.
struct FwdSchedule {
  int sched, nwave;
};

// Schedules compiled for each (precision, channels); index 0 is the default.
static int fwd_schedules(int C, int prec, const FwdSchedule** out) {
  static const FwdSchedule tile_capable[] = {{0, 8}, {1, 4}, {1, 8}, {0, 4}};
  static const FwdSchedule taps_only[] = {{0, 8}, {0, 4}};
  if ((prec == FK_PREC_BF16 && C == 32) || (prec == FK_PREC_FP32 && C == 16)) {
    *out = tile_capable;
    return 4;
  }
  *out = taps_only;
  return 2;
}

static std::mutex g_sched_mu;
static std::map<uint64_t, int> g_sched;

static uint64_t sched_key(int H, int W, int C, int prec) {
  return ((uint64_t)(uint32_t)H << 40) | ((uint64_t)(uint32_t)W << 20) | ((uint64_t)(uint32_t)C << 4) |
         (uint64_t)(prec & 15);
}

static int current_schedule(int H, int W, int C, int prec) {
  std::lock_guard<std::mutex> g(g_sched_mu);
  auto it = g_sched.find(sched_key(H, W, C, prec));
  return it == g_sched.end() ? 0 : it->second;
}

// Largest tile of configurations whose five activation lattices fit in the
// 160 KB LDS with at most 4 (8 waves) or 8 (4 waves) 16-pixel tiles per wave.
static int plan_forward(int HW, int C, int prec, int NW, int64_t m, FwdPlan* plan) {
  const int pixb = C * (prec == FK_PREC_FP32 ? 4 : 2);
  const size_t cap = 160 * 1024;
  const int maxt_cap = (NW == 4) ? 8 : 4;
  int best = 0;
  for (int T = 1; T <= 256; ++T) {
    const int npix = ((T * HW + 15) / 16) * 16;
    if (fwd_lds_bytes(T * HW, pixb) > cap) break;
    const int per_wave = (npix / 16 + NW - 1) / NW;
    if (per_wave > maxt_cap) break;
    best = T;
    if ((int64_t)T >= m) break;
  }
  if (best == 0) return FK_ERR_UNSUPPORTED;
  plan->T = best;
  plan->npix = ((best * HW + 15) / 16) * 16;
  plan->lds = fwd_lds_bytes(best * HW, pixb);
  plan->nwave = NW;
  const int per_wave = (plan->npix / 16 + NW - 1) / NW;
  plan->maxt = per_wave <= 2 ? 2 : (per_wave <= 4 ? 4 : 8);
  return FK_OK;
}

template <int PREC, int C, int MAXT, int ACT, int NWAVE, int SCHED>
static int launch_fwd(const FwdArgs& a, const FwdPlan& pl, hipStream_t s) {
  auto kern = convar_fwd_kernel<PREC, C, MAXT, ACT, NWAVE, SCHED>;
  // raise the dynamic LDS limit once per instantiation (not a stream
  // operation, so launches may later be captured into a graph)
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  FK_CHECK_HIP(attr);
  const int64_t ngroups = (a.m_count + pl.T - 1) / pl.T;
  const int per_cu = (int)((160 * 1024) / pl.lds) > 0 ? (int)((160 * 1024) / pl.lds) : 1;
  int64_t grid = (int64_t)device_cu_count() * per_cu;
  if (grid > ngroups) grid = ngroups;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * NWAVE), pl.lds, s, a);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

template <int PREC, int C, int NW, int SCHED>
static int launch_variant(const FwdArgs& a, const FwdPlan& pl, int act, hipStream_t s) {
  constexpr int LO = NW == 4 ? 4 : 2, HI = NW == 4 ? 8 : 4;
  if (pl.maxt <= LO)
    return act == FK_ACT_ELU ? launch_fwd<PREC, C, LO, FK_ACT_ELU, NW, SCHED>(a, pl, s)
                             : launch_fwd<PREC, C, LO, FK_ACT_RELU, NW, SCHED>(a, pl, s);
  return act == FK_ACT_ELU ? launch_fwd<PREC, C, HI, FK_ACT_ELU, NW, SCHED>(a, pl, s)
                           : launch_fwd<PREC, C, HI, FK_ACT_RELU, NW, SCHED>(a, pl, s);
}

template <int PREC, int C>
static int dispatch_fwd_c(const FwdArgs& a, const FwdPlan& pl, const FwdSchedule& sc, int act, hipStream_t s) {
  if constexpr ((PREC == FK_PREC_BF16 && C == 32) || (PREC == FK_PREC_FP32 && C == 16)) {
    if (sc.sched == 1)
      return sc.nwave == 4 ? launch_variant<PREC, C, 4, 1>(a, pl, act, s) : launch_variant<PREC, C, 8, 1>(a, pl, act, s);
  }
  return sc.nwave == 4 ? launch_variant<PREC, C, 4, 0>(a, pl, act, s) : launch_variant<PREC, C, 8, 0>(a, pl, act, s);
}

bool convar_supported(int channels, int precision) { return kernel_channels(channels, precision) > 0; }

}  // namespace fk

using namespace fk;

extern "C" {

int fk_convar_packed_sizes(int32_t depth, int32_t channels, int32_t precision, int64_t* n_weights, int64_t* n_aux) {
  FK_REQUIRE(n_weights && n_aux, "fk_convar_packed_sizes: null pointer");
  FK_REQUIRE(depth >= 1, "fk_convar_packed_sizes: depth must be >= 1");
  if (!convar_supported(channels, precision)) {
    set_error("ConvNetAutoregressive2D kernels: channels=%d precision=%d not implemented", channels, precision);
    return FK_ERR_UNSUPPORTED;
  }
  const ConvarLayout lay(kernel_channels(channels, precision), precision);
  const int nb = num_causal_blocks(depth);
  *n_weights = lay.total_frags(nb) * 64 * lay.E;
  *n_aux = lay.aux_total(nb);
  return FK_OK;
}

int fk_convar_pack_maps(int32_t depth, int32_t channels, int32_t precision, int32_t* wmap, int32_t* amap) {
  int64_t nw = 0, na = 0;
  int rc = fk_convar_packed_sizes(depth, channels, precision, &nw, &na);
  if (rc != FK_OK) return rc;
  FK_REQUIRE(wmap && amap, "fk_convar_pack_maps: null pointer");
  const int C = channels;
  const int CP = kernel_channels(channels, precision);
  const ConvarLayout lay(CP, precision);
  const int nb = num_causal_blocks(depth);
  for (int64_t i = 0; i < nw; ++i) wmap[i] = -1;
  for (int64_t i = 0; i < na; ++i) amap[i] = -1;
  // canonical offsets
  int64_t src = 0;
  struct Conv {
    int64_t koff, boff;
    int kh, kw, cin, cout;
  };
  auto take = [&](int kh, int kw, int cin, int cout) {
    Conv c{src, src + (int64_t)kh * kw * cin * cout, kh, kw, cin, cout};
    src = c.boff + cout;
    return c;
  };
  // kernel input channel -> channel of the C-channel network (-1: padding);
  // `cat`: the input is the concatenation [x | y] whose y half starts at CP/2
  auto real_cin = [&](int pc, const Conv& cv, bool cat) {
    if (!cat) return pc < cv.cin ? pc : -1;
    if (pc < CP / 2) return pc < C / 2 ? pc : -1;
    return pc - CP / 2 < C / 2 ? C / 2 + (pc - CP / 2) : -1;
  };
  auto fill = [&](int64_t frag_base, const Conv& cv, int mto, bool cat) {
    const int ntap = cv.kh * cv.kw;
    for (int tap = 0; tap < ntap; ++tap)
      for (int mt = 0; mt < mto; ++mt)
        for (int ks = 0; ks < lay.KS; ++ks) {
          const int64_t frag = frag_base + ((int64_t)tap * mto + mt) * lay.KS + ks;
          for (int lane = 0; lane < 64; ++lane) {
            const int m = lane % 16, g = lane / 16;
            for (int e = 0; e < lay.E; ++e) {
              const int pc = precision != FK_PREC_FP32 ? slot_channel_bf16(32 * ks + 8 * g + e)
                                                       : slot_channel_f32(ks, g);
              const int cin = real_cin(pc, cv, cat);
              const int cout = 16 * mt + m;
              const int64_t idx = (frag * 64 + lane) * lay.E + e;
              if (cout < cv.cout && cin >= 0)
                wmap[idx] = (int32_t)(cv.koff + ((int64_t)tap * cv.cin + cin) * cv.cout + cout);
            }
          }
        }
  };
  auto bias_copy = [&](int64_t dst, const Conv& cv) {
    for (int c = 0; c < cv.cout; ++c) amap[dst + c] = (int32_t)(cv.boff + c);
  };
  for (int b = 0; b < nb; ++b) {
    const int cin = b == 0 ? 1 : C;
    Conv v = take(3, 3, cin, C), h = take(1, 3, cin, C), x = take(1, 1, C, C / 2), y = take(1, 1, C, C / 2),
         cc = take(3, 3, C, C);
    const int64_t ab = lay.aux_bias(b);
    bias_copy(ab, v);
    bias_copy(ab + CP, h);
    bias_copy(ab + 2 * CP, x);
    bias_copy(ab + 2 * CP + CP / 2, y);
    bias_copy(ab + 3 * CP, cc);
    if (b == 0) {
      // (3,3,1,C) and (1,3,1,C) kernels, [tap][channel] with stride CP
      for (int t = 0; t < 9; ++t)
        for (int c = 0; c < C; ++c) amap[lay.aux_kv0(nb) + t * CP + c] = (int32_t)(v.koff + t * C + c);
      for (int t = 0; t < 3; ++t)
        for (int c = 0; c < C; ++c) amap[lay.aux_kh0(nb) + t * CP + c] = (int32_t)(h.koff + t * C + c);
    } else {
      fill(lay.off_v(b), v, lay.MT, false);
      fill(lay.off_h(b), h, lay.MT, false);
    }
    fill(lay.off_x(b), x, lay.MTH, false);
    fill(lay.off_y(b), y, lay.MTH, false);
    fill(lay.off_c(b), cc, lay.MT, true);
  }
  Conv head = take(1, 1, C, 4);
  fill(lay.head_base(nb), head, 1, false);
  for (int c = 0; c < 4; ++c) amap[lay.aux_head_bias(nb) + c] = (int32_t)(head.boff + c);
  return FK_OK;
}

int fk_pack_gather(const float* params, const int32_t* map, int64_t n, int32_t precision, void* dst, void* stream) {
  FK_REQUIRE(params && map && dst, "fk_pack_gather: null pointer");
  FK_REQUIRE(precision == FK_PREC_FP32 || precision == FK_PREC_BF16 || precision == FK_PREC_FP16,
             "fk_pack_gather: bad precision");
  if (n <= 0) return FK_OK;
  hipLaunchKernelGGL(pack_gather_kernel, dim3(div_up(n, 256)), dim3(256), 0, (hipStream_t)stream, params, map, n,
                     precision, dst);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

int fk_convar_forward(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W, int32_t depth,
                      int32_t channels, int32_t activation, const int8_t* spins, const int32_t* pairs,
                      const int64_t* m_dev, int64_t m_count, int32_t conn_kind, const int32_t* conn_a,
                      const int32_t* conn_b, float* logpsi, float* cond_logp, void* stream) {
  FK_REQUIRE(wpacked && aux && spins && logpsi, "fk_convar_forward: null pointer");
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1, "fk_convar_forward: bad shape");
  FK_REQUIRE(activation == FK_ACT_RELU || activation == FK_ACT_ELU, "fk_convar_forward: bad activation");
  FK_REQUIRE(!pairs || (conn_a && conn_b), "fk_convar_forward: pairs need a connection table");
  if (!convar_supported(channels, precision)) {
    set_error("ConvNetAutoregressive2D kernels: channels=%d precision=%d not implemented", channels, precision);
    return FK_ERR_UNSUPPORTED;
  }
  if (m_count <= 0) return FK_OK;
  const int CP = kernel_channels(channels, precision);
  const FwdSchedule* table = nullptr;
  const int nsched = fwd_schedules(CP, precision, &table);
  int si = current_schedule(H, W, CP, precision);
  if (si < 0 || si >= nsched) si = 0;
  const FwdSchedule sc = table[si];
  FwdPlan pl;
  if (plan_forward(H * W, CP, precision, sc.nwave, m_count, &pl) != FK_OK) {
    set_error("ConvNetAutoregressive2D kernels: lattice %dx%d with %d channels does not fit in LDS", H, W, channels);
    return FK_ERR_UNSUPPORTED;
  }
  FwdArgs a;
  a.w = wpacked;
  a.aux = aux;
  a.H = H;
  a.W = W;
  a.HW = H * W;
  a.nblocks = num_causal_blocks(depth);
  a.T = pl.T;
  a.npix = pl.npix;
  a.spins = spins;
  a.pairs = pairs;
  a.m_dev = m_dev;
  a.m_count = m_count;
  a.conn_kind = conn_kind;
  a.conn_a = conn_a;
  a.conn_b = conn_b;
  a.logpsi = logpsi;
  a.cond = cond_logp;
  hipStream_t s = (hipStream_t)stream;
  if (precision == FK_PREC_BF16) {
    if (CP == 32) return dispatch_fwd_c<FK_PREC_BF16, 32>(a, pl, sc, activation, s);
    return dispatch_fwd_c<FK_PREC_BF16, 64>(a, pl, sc, activation, s);
  }
  if (precision == FK_PREC_FP16) {
    if (CP == 32) return dispatch_fwd_c<FK_PREC_FP16, 32>(a, pl, sc, activation, s);
    return dispatch_fwd_c<FK_PREC_FP16, 64>(a, pl, sc, activation, s);
  }
  if (CP == 16) return dispatch_fwd_c<FK_PREC_FP32, 16>(a, pl, sc, activation, s);
  if (CP == 32) return dispatch_fwd_c<FK_PREC_FP32, 32>(a, pl, sc, activation, s);
  return dispatch_fwd_c<FK_PREC_FP32, 64>(a, pl, sc, activation, s);
}

int32_t fk_convar_num_schedules(int32_t channels, int32_t precision) {
  if (!convar_supported(channels, precision)) return 0;
  const FwdSchedule* t = nullptr;
  return fwd_schedules(kernel_channels(channels, precision), precision, &t);
}

int fk_convar_schedule_info(int32_t channels, int32_t precision, int32_t schedule, int32_t* tile_outer,
                            int32_t* waves) {
  FK_REQUIRE(tile_outer && waves, "fk_convar_schedule_info: null pointer");
  const int n = fk_convar_num_schedules(channels, precision);
  FK_REQUIRE(schedule >= 0 && schedule < n, "fk_convar_schedule_info: schedule out of range");
  const FwdSchedule* t = nullptr;
  fwd_schedules(kernel_channels(channels, precision), precision, &t);
  *tile_outer = t[schedule].sched;
  *waves = t[schedule].nwave;
  return FK_OK;
}

int fk_convar_set_schedule(int32_t H, int32_t W, int32_t channels, int32_t precision, int32_t schedule) {
  const int n = fk_convar_num_schedules(channels, precision);
  FK_REQUIRE(n > 0, "fk_convar_set_schedule: unsupported channels/precision");
  FK_REQUIRE(schedule >= -1 && schedule < n, "fk_convar_set_schedule: schedule out of range");
  const int CP = kernel_channels(channels, precision);
  std::lock_guard<std::mutex> g(g_sched_mu);
  if (schedule < 0) g_sched.erase(sched_key(H, W, CP, precision));
  else g_sched[sched_key(H, W, CP, precision)] = schedule;
  return FK_OK;
}

int32_t fk_convar_get_schedule(int32_t H, int32_t W, int32_t channels, int32_t precision) {
  return current_schedule(H, W, kernel_channels(channels, precision), precision);
}

}  // extern "C"
