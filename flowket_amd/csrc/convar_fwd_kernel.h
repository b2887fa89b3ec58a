// Fused ConvNetAutoregressive2D forward kernel template (machines/
// conv_net_autoregressive_2D.py:33-74 + machines/abstract_machine.py:31-71),
// instantiated by convar_fwd_<precision>_lds.hip (lattices resident in LDS)
// and convar_fwd_<precision>_hg.hip (the horizontal-stack lattices in a
// per-workgroup global scratch, for lattices whose five buffers exceed the
// LDS); convar_fwd.hip holds the host side.
//
// One workgroup evaluates a tile of T configurations through the whole network
// (buffers V0, V1, CAT, H0, H1 of T*H*W pixels x C channels).  Each
// convolution is an implicit GEMM on MFMA: A = weights (16 output channels x k
// input slots), B = activations (k slots x 16 pixels) read at the tap offset,
// D = 16 channels x 16 pixels.  The 1x1 convolutions (x, y, head) take their B
// operand straight from the registers of the preceding convolution's D
// fragment: input slots are numbered so that a D fragment *is* a B fragment
// (see convar_layout.h).  Zero padding, DownShift and RightShift are index
// arithmetic (tap validity masks and shifted write targets), never
// materialised.
#pragma once
#include "fk_common.h"
#include "convar_layout.h"
#include "convar_dev.h"

namespace fk {

// ---------------------------------------------------------------------------
// all-zero pixels after each activation buffer (one per bank placement of
// the 8-pixel swizzle period)
constexpr int ZERO_PIX = 8;
// LDS of the forward (all of it dynamic)
constexpr int FWD_STATIC_LDS = 0;
constexpr int FWD_DYN_LDS = 160 * 1024;

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
  // 16-bit precisions: an already packed quad / fragment
  template <typename Q>
  __device__ static __forceinline__ void store_quad_raw(uint8_t* buf, int p, int q, Q v) {
    *reinterpret_cast<Q*>(buf + quad(p, q)) = v;
  }
  template <typename F>
  __device__ static __forceinline__ void store_frag_at(uint8_t* buf, uint32_t off, F v) {
    *reinterpret_cast<F*>(buf + off) = v;
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
  // chunk at off ^ 64j, so they are read as one access.
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
  uint8_t* hscr;  // HG kernels: 2 lattices per workgroup of global scratch
  // Row reuse.  A connected configuration equal to its sample on every row
  // before R0 has the sample's activations on those rows (every activation
  // of row r depends on input rows <= r only).  Such a launch keeps rows
  // R0-2 .. H-1 of each configuration (HR rows): the first ROW0 = 2 are halo
  // rows whose block inputs (vertical stack, concatenation) are copied from
  // the sample's cache, the HC = H - R0 rows after them are computed, and
  // the sample's prefix sum of site log-probabilities over rows < R0 is added.
  // R0 = 0: the whole lattice (HR = HC = H, ROW0 = 0).
  int R0, HR, HC, ROW0;
  uint8_t* rcache;           // per sample: for every block its V input and CAT lattices
  int64_t rcache_stride;     // bytes per sample
  float* prefix;             // per sample: (H + 1) x (re, im) prefix sums over rows
  int cache_write;           // sample pass: write rcache and prefix
  const int64_t* out_index;  // output row of configuration m (null: m)
  const int64_t* m_range;    // device (first, end) of the configurations within pairs (null: 0, count)
};

// Byte offset, within one sample's cache, of lattice `which` (0 = the block's
// vertical input, 1 = its concatenation) of block blk: pixels in global
// order, PIXB bytes each, 16-byte chunks in their unswizzled order.
__device__ __forceinline__ int64_t rcache_off(int blk, int which, int HW, int PIXB) {
  return ((int64_t)blk * 2 + which) * HW * PIXB;
}

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

// Prefetch depth of a convolution's weights (taps ahead), bounded by the
// registers one tap's fragments take.
template <int PREC, int C, int MTO>
struct WeightDepth {
  static constexpr int KS = C / Prec<PREC>::KW;
  static constexpr int FR = (int)(sizeof(typename Prec<PREC>::frag) / 4);
  static constexpr int WREGS = MTO * KS * FR;
  static constexpr int WD = WREGS <= 16 ? 2 : (WREGS <= 32 ? 1 : 0);
};

// The head of a convolution pass -- its bias and the weights of its first WD
// taps -- requested one phase early (during the previous phase's epilogue,
// across the barrier), so that a pass starts without waiting on L2.
template <int PREC, int C, int MTO>
struct PassHead {
  using frag = typename Prec<PREC>::frag;
  static constexpr int KS = C / Prec<PREC>::KW;
  // taps whose weights travel in the head: the 64-channel variants have no
  // registers to spare across an epilogue, so their passes fetch their own
  static constexpr int WD = C <= 32 ? WeightDepth<PREC, C, MTO>::WD : 0;
  frag w[WD > 0 ? WD : 1][MTO][KS];
  f32x4 bias[MTO];
  __device__ __forceinline__ void load_bias(const float* b, int g) {
#pragma unroll
    for (int mt = 0; mt < MTO; ++mt) bias[mt] = *reinterpret_cast<const f32x4*>(b + 16 * mt + 4 * g);
  }
  __device__ __forceinline__ void load(const frag* __restrict__ wg, const float* b, int lane, int g) {
#pragma unroll
    for (int t = 0; t < WD; ++t)
#pragma unroll
      for (int mt = 0; mt < MTO; ++mt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) w[t][mt][ks] = wg[((t * MTO + mt) * KS + ks) * 64 + lane];
    load_bias(b, g);
  }
};

// acc[u][mt] += sum over taps of W[tap][mt] x B(pixel + tap offset).
// toff.get(u, k) is this lane's byte offset of k-step 0 at tap k of tile u,
// pointing at the all-zero pixel for padding taps -- tiles past
// the wave's count hold only such offsets, so every tile is computed without
// a branch (their epilogues are skipped).
// Software pipeline: the weights (read from L2) are requested WD taps ahead
// -- the first taps already by the caller, in hd, where the head carries
// them -- and the B fragments of all tiles BD taps ahead (depths bounded by
// the registers they take); scheduling fences keep the compiler from sinking
// those requests next to their first use.
template <int PREC, int C, int MAXT, int KIND, int MTO>
__device__ __forceinline__ void conv_pass(f32x4 (&acc)[MAXT][MTO], const uint8_t* buf,
                                          const typename Prec<PREC>::frag* __restrict__ wg,
                                          const TapOff<MAXT>& toff, int lane, const PassHead<PREC, C, MTO>& hd) {
  using P = Prec<PREC>;
  using frag = typename P::frag;
  using L = Lds<PREC, C>;
  constexpr int KS = C / P::KW;
  constexpr int NTAP = Taps<KIND>::N;
  constexpr int FR = (int)(sizeof(frag) / 4);
  constexpr int BREGS = MAXT * KS * FR;
  constexpr int WD = WeightDepth<PREC, C, MTO>::WD;  // weight prefetch distance (taps)
  constexpr int BD = BREGS <= 32 ? 1 : 0;             // B prefetch distance (taps)
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
  for (int t = 0; t < WD && t < NTAP; ++t) {
    if (t < PassHead<PREC, C, MTO>::WD) {
#pragma unroll
      for (int mt = 0; mt < MTO; ++mt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) w[t][mt][ks] = hd.w[t][mt][ks];
    } else {
      load_w(t, w[t]);
    }
  }
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

// Row-reuse cache copies.  In LDS a configuration's pixels are swizzled
// 16-byte chunks (Lds::chunk); in a sample's cache the pixels are in global
// order with their chunks unswizzled (the layout of pixel 0).
// rcache_out: every local pixel of every configuration, LDS -> cache (the
// sample pass, whose local pixels are the global ones).
template <int PREC, int C>
__device__ __forceinline__ void rcache_out(const uint8_t* buf, const FwdArgs& a, int blk, int which, int T, int HWL,
                                           const int* s_samp) {
  using L = Lds<PREC, C>;
  constexpr int CPP = L::CPP;
  const int per = HWL * CPP;
  for (int idx = threadIdx.x; idx < T * per; idx += blockDim.x) {
    const int t = idx / per, r = idx - t * per;
    const int b = s_samp[t];
    if (b < 0) continue;
    const int ql = r / CPP, c = r % CPP;
    uint8_t* cp = a.rcache + (int64_t)b * a.rcache_stride + rcache_off(blk, which, a.HW, L::PIXB) +
                  (int64_t)ql * L::PIXB + 16 * c;
    *reinterpret_cast<uint4*>(cp) = *reinterpret_cast<const uint4*>(buf + L::chunk(t * HWL + ql, c));
  }
}

// rcache_in (restricted pass): the ROW0 halo rows of every configuration,
// cache -> LDS, and with yrow also the y half (quads C/8 .. C/4-1) of the
// first computed row.  Each thread issues up to RC_BATCH reads before it
// writes any, so one memory latency covers the copy.
constexpr int RC_BATCH = 4;
template <int PREC, int C>
__device__ __forceinline__ void rcache_in(uint8_t* buf, const FwdArgs& a, int blk, int which, int T, int HWL,
                                          bool yrow, const int* s_samp) {
  using L = Lds<PREC, C>;
  constexpr int CPP = L::CPP;
  constexpr bool Q8 = Prec<PREC>::H16;  // 16-bit quads are 8 bytes, fp32 quads 16
  constexpr int NQ = C / 8;             // quads of the y half of a pixel
  constexpr int HALF8 = 1 << 30;        // marks an 8-byte piece
  const int W = a.W;
  const int nfull = a.ROW0 * W * CPP;
  const int per = nfull + (yrow ? W * NQ : 0);
  const int total = T * per;
  const int gofs = (a.R0 - a.ROW0) * W;
  for (int base = threadIdx.x; base < total; base += RC_BATCH * (int)blockDim.x) {
    uint4 v[RC_BATCH];
    int lo[RC_BATCH];
#pragma unroll
    for (int r = 0; r < RC_BATCH; ++r) {
      const int idx = base + r * (int)blockDim.x;
      lo[r] = -1;
      v[r] = make_uint4(0u, 0u, 0u, 0u);
      if (idx < total) {
        const int t = idx / per, k = idx - t * per;
        const int b = s_samp[t];
        if (b >= 0) {
          const uint8_t* src = a.rcache + (int64_t)b * a.rcache_stride + rcache_off(blk, which, a.HW, L::PIXB);
          if (k < nfull) {
            const int ql = k / CPP, c = k % CPP;
            v[r] = *reinterpret_cast<const uint4*>(src + (int64_t)(ql + gofs) * L::PIXB + 16 * c);
            lo[r] = L::chunk(t * HWL + ql, c);
          } else {
            const int kk = k - nfull;
            const int ql = a.ROW0 * W + kk / NQ, qd = NQ + kk % NQ;
            const uint8_t* g = src + (int64_t)(ql + gofs) * L::PIXB + L::quad(0, qd);
            if constexpr (Q8) {
              const uint2 x = *reinterpret_cast<const uint2*>(g);
              v[r] = make_uint4(x.x, x.y, 0u, 0u);
              lo[r] = L::quad(t * HWL + ql, qd) | HALF8;
            } else {
              v[r] = *reinterpret_cast<const uint4*>(g);
              lo[r] = L::quad(t * HWL + ql, qd);
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RC_BATCH; ++r) {
      if (lo[r] < 0) continue;
      if (Q8 && (lo[r] & HALF8)) *reinterpret_cast<uint2*>(buf + (lo[r] & ~HALF8)) = make_uint2(v[r].x, v[r].y);
      else *reinterpret_cast<uint4*>(buf + lo[r]) = v[r];
    }
  }
}

// SCHED 0: taps outer -- each tap's weights are loaded once and serve all the
// wave's tiles (accumulators of every tile live in registers).
// SCHED 1: tiles outer -- a 3x3 convolution's weights stay in registers for
// the whole phase and the wave walks its tiles one by one (fits bf16 C=32 and
// fp32 C=16).
// HG: the horizontal-stack lattices H0, H1 live in global scratch (a.hscr,
// two lattices per workgroup, L2-resident) instead of LDS.
// ROWS: the row-reuse sample pass (cache and prefix writes) and restricted
// passes (halo rows, prefix add); whole-lattice launches compile without
// them (device ranges and output scatter, m_range / out_index, work in both).
template <int PREC, int C, int MAXT, int ACT, int NWAVE, int SCHED, bool HG, bool ROWS>
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
  // per configuration: HWL pixels in LDS (halo rows included), HWC computed
  const int HWL = ROWS ? a.HR * a.W : a.HW, HWC = ROWS ? a.HC * a.W : a.HW;
  const int NPX = a.T * HWL;
  const int BUFB = (NPX + ZERO_PIX) * PIXB;
  uint8_t* V0 = lds;
  uint8_t* V1 = V0 + BUFB;
  uint8_t* CAT = V1 + BUFB;
  uint8_t* H0;
  uint8_t* H1;
  int8_t* sp;
  if constexpr (HG) {
    H0 = a.hscr + (int64_t)blockIdx.x * 2 * BUFB;
    H1 = H0 + BUFB;
    sp = reinterpret_cast<int8_t*>(CAT + BUFB);
  } else {
    H0 = CAT + BUFB;
    H1 = H0 + BUFB;
    sp = reinterpret_cast<int8_t*>(H1 + BUFB);
  }
  // after the spins: the sample of each configuration of the group (-1: none)
  int* s_samp = reinterpret_cast<int*>(sp + ((NPX + 3) & ~3));
  // per-site log-probabilities (and, when asked for, the conditionals) of
  // the final block: V1 is neither read nor written by the final block, so
  // they reuse its space
  float2* part = reinterpret_cast<float2*>(V1);
  float2* cpart = part + NPX;

  const int lane = threadIdx.x & 63;
  // wave index and tile count are wave-uniform: keep them in SGPRs so tile
  // predicates are scalar branches
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int NW = blockDim.x >> 6;
  const int n16 = lane & 15, g = lane >> 4;
  const int H = a.H, W = a.W, HW = a.HW, T = a.T;
  const int HR = ROWS ? a.HR : H, ROW0 = ROWS ? a.ROW0 : 0;
  const int gofs = ROWS ? (a.R0 - ROW0) * W : 0;  // global index of a configuration's LDS pixel 0
  const int ntiles = npix >> 4;
  const int nt = (ntiles > wave) ? (ntiles - wave + NW - 1) / NW : 0;

  int pix[MAXT], pi[MAXT], pj[MAXT];
  uint32_t msk[MAXT];
  TapOff<MAXT> toff;
  bool inp[MAXT];
#pragma unroll
  for (int u = 0; u < MAXT; ++u) {
    // tiles run over the computed pixels; pix is the LDS pixel
    const int pc = (wave + u * NW) * 16 + n16;
    const bool in = (u < nt) && (pc < T * HWC);
    const int t = in ? pc / HWC : 0;
    const int rem = in ? pc - t * HWC : 0;
    const int i = ROW0 + rem / W, j = rem - (rem / W) * W;
    const int p = in ? t * HWL + i * W + j : NPX + (n16 & (ZERO_PIX - 1));
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
  for (int z = threadIdx.x; z < 3 * ZERO_PIX * (PIXB / 4); z += blockDim.x) {
    const int buf = z / (ZERO_PIX * PIXB / 4), w4 = z - buf * (ZERO_PIX * PIXB / 4);
    reinterpret_cast<uint32_t*>(lds + buf * BUFB + NPX * PIXB)[w4] = 0u;
  }
  for (int z = threadIdx.x; z < ZERO_PIX * (PIXB / 4); z += blockDim.x) {
    reinterpret_cast<uint32_t*>(H0 + NPX * PIXB)[z] = 0u;
    reinterpret_cast<uint32_t*>(H1 + NPX * PIXB)[z] = 0u;
  }

  int64_t M = a.m_dev ? *a.m_dev : a.m_count;
  int64_t m0 = 0;
  if (a.m_range) {
    m0 = a.m_range[0];
    M = a.m_range[1] - m0;
  }
  if (M > a.m_count) M = a.m_count;
  const int64_t ngroups = (M + T - 1) / T;
  const int nb = a.nblocks;
  const frag* wall = reinterpret_cast<const frag*>(a.w);

  // head of the next vertical pass (block 0 convolves the 1-channel input on
  // VALU: bias only)
  PassHead<PREC, C, MT> hv;
  hv.load_bias(a.aux + lay.aux_bias(0), g);

  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    // ---- configurations -> LDS (int8 spins; connection applied on the fly):
    //      one wave per configuration, so the (sample, connection) lookup is
    //      wave-uniform and the lanes read the spin row
    for (int t = wave; t < T; t += NW) {
      const int64_t m = grp * T + t;
      const bool live = m < M;
      int64_t b = m;
      int c = 0;
      if (live && a.pairs) {
        b = a.pairs[2 * (m0 + m)];
        c = a.pairs[2 * (m0 + m) + 1];
      }
      int ca = -1, cb = -1;
      if (live && c > 0) {
        ca = a.conn_a[c];
        if (a.conn_kind != FK_CONN_FLIP && ca >= 0) cb = a.conn_b[c];
      }
      const int8_t* row = a.spins + (live ? b : 0) * HW;
      for (int ql = lane; ql < HWL; ql += 64) {
        const int q = ql + gofs;
        int8_t v = 0;
        if (live) {
          v = row[q];
          if (a.conn_kind == FK_CONN_FLIP) {
            if (q == ca) v = (int8_t)-v;
          } else if (ca >= 0) {
            if (q == ca) v = row[cb];
            else if (q == cb) v = row[ca];
          }
        }
        sp[t * HWL + ql] = v;
      }
      if (ROWS && lane == 0) s_samp[t] = live ? b : -1;
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
      // biases of the two 1x1 convolutions, read once per block (not per
      // tile, where each read would stall the epilogue on its load), and --
      // for 16-bit precisions -- their activations packed, which is what
      // the DownShift row and the RightShift column receive
      f32x4 by0[MTH], bx0[MTH];
      typename Prec<PREC == FK_PREC_FP32 ? FK_PREC_BF16 : PREC>::quad aby[MTH], abx[MTH];
#pragma unroll
      for (int mh = 0; mh < MTH; ++mh) {
        by0[mh] = bias4(by + 16 * mh + 4 * g);
        bx0[mh] = bias4(bx + 16 * mh + 4 * g);
        if constexpr (P::H16) {
          aby[mh] = act_quad<PREC, ACT>(by0[mh]);
          abx[mh] = act_quad<PREC, ACT>(bx0[mh]);
        }
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
        frag bf[KS];
        if constexpr (P::H16) {
          // act(v') packed once: the B operand of the y 1x1 and, in the
          // first and A blocks, the stored vertical output
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) bf[ks] = act_frag<PREC, ACT>(acc[2 * ks], acc[2 * ks + 1]);
          if (inp[u]) {
#pragma unroll
            for (int j = 0; j < KS; ++j) {
              const uint32_t so = toff.get(u, 10) ^ (64u * j);
              if (kind == BK_FIRST) L::store_frag_at(V0, so, bf[j]);
              else if (kind == BK_A) L::store_frag_at(V1, so, bf[j]);
              else if (kind == BK_B) {
                f32x4 rl, rh;
                L::load_pair_at(V0, so, rl, rh);
                L::store_frag_at(V0, so, act_frag<PREC, ACT>(rl + acc[2 * j], rh + acc[2 * j + 1]));
              }
            }
          }
        } else {
          f32x4 av[MT];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) av[mt] = act4<ACT>(acc[mt]);
          if (inp[u]) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
              const uint32_t so = L::self_quad(toff.get(u, 10), mt);
              if (kind == BK_FIRST) L::store_quad_at(V0, so, av[mt]);
              else if (kind == BK_A) L::store_quad_at(V1, so, av[mt]);
              else if (kind == BK_B) L::store_quad_at(V0, so, act4<ACT>(L::load_quad_at(V0, so) + acc[mt]));
            }
          }
          d_to_b<PREC, MT, KS>(av, bf);
        }
        // the last row's output wraps to row 0 as the DownShift's zero row
        // (bias only); in a restricted pass row 0 is a halo row, which holds
        // the sample's values instead
        const bool down = pi[u] + 1 < HR;
        const int tgt = down ? pix[u] + W : pix[u] - (HR - 1) * W;
        const bool ystore = inp[u] && (down || ROW0 == 0);
#pragma unroll
        for (int mh = 0; mh < MTH; ++mh) {
          f32x4 y = by0[mh];
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) y = P::mfma(wy[mh][ks], bf[ks], y);
          if (ystore && 16 * mh + 4 * g < C / 2) {
            if constexpr (P::H16) L::store_quad_raw(CAT, tgt, C / 8 + 4 * mh + g, down ? act_quad<PREC, ACT>(y) : aby[mh]);
            else L::store_quad(CAT, tgt, C / 8 + 4 * mh + g, act4<ACT>(down ? y : by0[mh]));
          }
        }
      };
      // -- epilogue of the horizontal 1xk: x = act(1x1(act(.))) into CAT (the
      //    final block's RightShift folded into the write target)
      auto epi_h = [&](int u, const f32x4 (&acc)[MT]) {
        frag bf[KS];
        if constexpr (P::H16) {
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) bf[ks] = act_frag<PREC, ACT>(acc[2 * ks], acc[2 * ks + 1]);
        } else {
          f32x4 ax[MT];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) ax[mt] = act4<ACT>(acc[mt]);
          d_to_b<PREC, MT, KS>(ax, bf);
        }
        bool keep = true;
        int tgt = pix[u];
        if (kind == BK_FINAL) {
          keep = pj[u] + 1 < W;
          tgt = keep ? pix[u] + 1 : pix[u] - (W - 1);
        }
#pragma unroll
        for (int mh = 0; mh < MTH; ++mh) {
          f32x4 x = bx0[mh];
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) x = P::mfma(wx[mh][ks], bf[ks], x);
          if (inp[u] && 16 * mh + 4 * g < C / 2) {
            if constexpr (P::H16) L::store_quad_raw(CAT, tgt, 4 * mh + g, keep ? act_quad<PREC, ACT>(x) : abx[mh]);
            else L::store_quad(CAT, tgt, 4 * mh + g, act4<ACT>(keep ? x : bx0[mh]));
          }
        }
      };

      // ================= phase VH: vertical stack + horizontal 1xk =========
      PassHead<PREC, C, MT> hc;  // head of this block's h3 pass
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
        PassHead<PREC, C, MT> hh;
        {
          f32x4 acc[MAXT][MT];
#pragma unroll
          for (int u = 0; u < MAXT; ++u)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) acc[u][mt] = hv.bias[mt];
          if (kind == BK_FIRST) {
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
              if (u < nt) valu_v(u, acc[u]);
            hh.load_bias(bias + C, g);
          } else {
            conv_pass<PREC, C, MAXT, 0, MT>(acc, Vin, wall + lay.off_v(blk) * 64, toff, lane, hv);
            hh.load(wall + lay.off_h(blk) * 64, bias + C, lane, g);
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
            for (int mt = 0; mt < MT; ++mt) acc[u][mt] = hh.bias[mt];
          if (kind == BK_FIRST) {
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
              if (u < nt) valu_h(u, acc[u]);
          } else {
            conv_pass<PREC, C, MAXT, 1, MT>(acc, Hin, wall + lay.off_h(blk) * 64, toff, lane, hh);
          }
          hc.load(wall + lay.off_c(blk) * 64, bias + 3 * C, lane, g);
#pragma unroll
          for (int u = 0; u < MAXT; ++u)
            if (u < nt) epi_h(u, acc[u]);
        }
      }
      // row reuse, restricted pass: the halo rows of the concatenation (and
      // the y half of its first computed row, which the DownShift takes from
      // the halo) from the sample's cache.  This phase neither reads nor
      // writes those rows, so the copy shares the phase's closing barrier.
      if (ROWS && a.rcache && !a.cache_write && ROW0 > 0)
        rcache_in<PREC, C>(CAT, a, blk, 1, T, HWL, true, s_samp);
      __syncthreads();
      // row reuse, sample pass: this block's concatenation and the next
      // block's vertical input (this phase's vertical output, complete now)
      // to the cache; phase H3 changes neither
      if (ROWS && a.rcache && a.cache_write) {
        rcache_out<PREC, C>(CAT, a, blk, 1, T, HWL, s_samp);
        if (blk + 1 < nb) rcache_out<PREC, C>(block_kind(blk + 1, nb) == BK_B ? V1 : V0, a, blk + 1, 0, T, HWL, s_samp);
      }

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
                if (kind == BK_FIRST) L::store_frag_at(H0, so, act_frag<PREC, ACT>(acc[2 * j], acc[2 * j + 1]));
                else if (kind == BK_A) L::store_frag_at(H1, so, act_frag<PREC, ACT>(acc[2 * j], acc[2 * j + 1]));
                else {
                  f32x4 rl, rh;
                  L::load_pair_at(H0, so, rl, rh);
                  L::store_frag_at(H0, so, act_frag<PREC, ACT>(rl + acc[2 * j], rh + acc[2 * j + 1]));
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
        frag bf[KS];
        if constexpr (P::H16) {
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) bf[ks] = act_frag<PREC, ACT>(acc[2 * ks], acc[2 * ks + 1]);
        } else {
          f32x4 ah[MT];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) ah[mt] = act4<ACT>(acc[mt]);
          d_to_b<PREC, MT, KS>(ah, bf);
        }
        f32x4 z = hb;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) z = P::mfma(whd[ks], bf[ks], z);
        if (g == 0 && inp[u]) {
          // z = (Re z0, Re z1, Im z0, Im z1) for spin values (+1, -1)
          const float mx = fmaxf(z.x, z.y);
          const float lse = mx + 0.5f * log1pf(expf(-2.0f * fabsf(z.x - z.y)));
          const bool up = sp[pix[u]] > 0;
          part[pix[u]] = make_float2((up ? z.x : z.y) - lse, up ? z.z : z.w);
          if (a.cond) cpart[pix[u]] = make_float2(2.0f * (z.x - lse), 2.0f * (z.y - lse));
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
          for (int mt = 0; mt < MT; ++mt) acc[u][mt] = hc.bias[mt];
        conv_pass<PREC, C, MAXT, 2, MT>(acc, CAT, wall + lay.off_c(blk) * 64, toff, lane, hc);
        // the next block's vertical pass (block 0 of the next group: bias only)
        if (blk + 1 < nb) hv.load(wall + lay.off_v(blk + 1) * 64, a.aux + lay.aux_bias(blk + 1), lane, g);
        else hv.load_bias(a.aux + lay.aux_bias(0), g);
#pragma unroll
        for (int u = 0; u < MAXT; ++u)
          if (u < nt) epi_c(u, acc[u]);
      }
      // row reuse, restricted pass: the halo rows of the next block's
      // vertical input from the sample's cache.  That lattice is the other
      // V buffer from this block's input, and this block writes it only on
      // computed rows, so the copy shares the block's closing barrier.
      if (ROWS && a.rcache && !a.cache_write && ROW0 > 0 && blk + 1 < nb)
        rcache_in<PREC, C>(block_kind(blk + 1, nb) == BK_B ? V1 : V0, a, blk + 1, 0, T, HWL, false, s_samp);
      __syncthreads();
    }

    // ---- per-configuration sums of the site log-probabilities: each lane
    //      sums one row in column order, then the rows are added in order
    //      (starting from the sample's prefix over the rows before R0); the
    //      sample pass records the prefix sums over rows
    for (int t = wave; t < T; t += NW) {
      const int64_t m = grp * T + t;
      float rr = 0.0f, ri = 0.0f;
      const int nrows = HR - ROW0;
      if (lane < nrows) {
        const float2* rowp = part + t * HWL + (ROW0 + lane) * W;
        for (int j = 0; j < W; ++j) {
          rr += rowp[j].x;
          ri += rowp[j].y;
        }
      }
      float sr = 0.0f, si = 0.0f;
      const int b = ROWS ? s_samp[t] : -1;
      if (ROWS && ROW0 > 0 && b >= 0) {
        sr = a.prefix[((int64_t)b * (H + 1) + a.R0) * 2];
        si = a.prefix[((int64_t)b * (H + 1) + a.R0) * 2 + 1];
      }
      float* pref = (ROWS && a.cache_write && b >= 0) ? a.prefix + (int64_t)b * (H + 1) * 2 : nullptr;
      if (pref && lane == 0) {
        pref[0] = 0.0f;
        pref[1] = 0.0f;
      }
      for (int r = 0; r < nrows; ++r) {
        sr += __shfl(rr, r, 64);
        si += __shfl(ri, r, 64);
        if (pref && lane == 0) {
          pref[2 * (r + 1)] = sr;
          pref[2 * (r + 1) + 1] = si;
        }
      }
      if (lane == 0 && m < M) {
        const int64_t mo = a.out_index ? a.out_index[m0 + m] : m;
        a.logpsi[2 * mo] = sr;
        a.logpsi[2 * mo + 1] = si;
      }
      if (a.cond && m < M) {
        for (int q = lane; q < HWL; q += 64) {
          const float2 c = cpart[t * HWL + q];
          float* o = a.cond + (m * HWL + q) * 2;
          o[0] = c.x;
          o[1] = c.y;
        }
      }
    }
    __syncthreads();
  }
}

struct FwdPlan {
  int T, npix, maxt, nwave;
  size_t lds;
};

// workgroups that fit a CU's 160 KB of LDS (dynamic + static)
inline int fwd_per_cu(const FwdPlan& pl) {
  const int n = (int)((160 * 1024) / (pl.lds + FWD_STATIC_LDS));
  return n > 0 ? n : 1;
}

// persistent grid: CU count x workgroups per CU, at most one per group
inline int64_t fwd_grid(const FwdArgs& a, const FwdPlan& pl) {
  const int64_t ngroups = (a.m_count + pl.T - 1) / pl.T;
  const int per_cu = fwd_per_cu(pl);
  int64_t grid = (int64_t)device_cu_count() * per_cu;
  if (grid > ngroups) grid = ngroups;
  if (grid < 1) grid = 1;
  return grid;
}

template <int PREC, int C, int MAXT, int ACT, int NWAVE, int SCHED, bool HG, bool ROWS>
int launch_fwd(const FwdArgs& a, const FwdPlan& pl, hipStream_t s) {
  auto kern = convar_fwd_kernel<PREC, C, MAXT, ACT, NWAVE, SCHED, HG, ROWS>;
  // raise the dynamic LDS limit once per instantiation (not a stream
  // operation, so launches may later be captured into a graph)
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, FWD_DYN_LDS);
  FK_CHECK_HIP(attr);
  hipLaunchKernelGGL(kern, dim3((unsigned)fwd_grid(a, pl)), dim3(64 * NWAVE), pl.lds, s, a);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

template <int PREC, int C, int NW, int SCHED, bool HG, bool ROWS>
int launch_variant(const FwdArgs& a, const FwdPlan& pl, int act, hipStream_t s) {
  constexpr int LO = NW == 4 ? 4 : 2, HI = NW == 4 ? 8 : 4;
  if (pl.maxt <= LO)
    return act == FK_ACT_ELU ? launch_fwd<PREC, C, LO, FK_ACT_ELU, NW, SCHED, HG, ROWS>(a, pl, s)
                             : launch_fwd<PREC, C, LO, FK_ACT_RELU, NW, SCHED, HG, ROWS>(a, pl, s);
  return act == FK_ACT_ELU ? launch_fwd<PREC, C, HI, FK_ACT_ELU, NW, SCHED, HG, ROWS>(a, pl, s)
                           : launch_fwd<PREC, C, HI, FK_ACT_RELU, NW, SCHED, HG, ROWS>(a, pl, s);
}

// Launch of one forward: the precision's kernel variant for channel width
// cp, loop order sched (0 taps outer, 1 tiles outer), nwave waves, the
// horizontal stack in LDS or global scratch (hg), whole-lattice or row-reuse
// (rows; those always run taps outer).  Each precision's kernels are compiled
// in a file of their own (convar_fwd_<precision>_{lds,hg}.hip).
template <int PREC, bool HG>
int dispatch_fwd(const FwdArgs& a, const FwdPlan& pl, int cp, int sched, int nwave, int act, bool rows,
                 hipStream_t s);
#define FK_DECLARE_DISPATCH_FWD(PREC, HG)                                                                      \
  template <>                                                                                                  \
  int dispatch_fwd<PREC, HG>(const FwdArgs& a, const FwdPlan& pl, int cp, int sched, int nwave, int act,      \
                             bool rows, hipStream_t s);
FK_DECLARE_DISPATCH_FWD(FK_PREC_BF16, false)
FK_DECLARE_DISPATCH_FWD(FK_PREC_BF16, true)
FK_DECLARE_DISPATCH_FWD(FK_PREC_FP16, false)
FK_DECLARE_DISPATCH_FWD(FK_PREC_FP16, true)
FK_DECLARE_DISPATCH_FWD(FK_PREC_FP32, false)
FK_DECLARE_DISPATCH_FWD(FK_PREC_FP32, true)

// The body those files instantiate: channel widths 32/64 (16-bit) or
// 16/32/64 (fp32); tiles-outer only where it is compiled (bf16 C=32, fp32
// C=16, LDS-resident, whole lattice).
template <int PREC, int C, bool HG>
int dispatch_fwd_c(const FwdArgs& a, const FwdPlan& pl, int sched, int nwave, int act, bool rows, hipStream_t s) {
  if (rows)
    return nwave == 4 ? launch_variant<PREC, C, 4, 0, HG, true>(a, pl, act, s)
                      : launch_variant<PREC, C, 8, 0, HG, true>(a, pl, act, s);
  if constexpr (!HG && ((PREC == FK_PREC_BF16 && C == 32) || (PREC == FK_PREC_FP32 && C == 16))) {
    if (sched == 1)
      return nwave == 4 ? launch_variant<PREC, C, 4, 1, false, false>(a, pl, act, s)
                        : launch_variant<PREC, C, 8, 1, false, false>(a, pl, act, s);
  }
  return nwave == 4 ? launch_variant<PREC, C, 4, 0, HG, false>(a, pl, act, s)
                    : launch_variant<PREC, C, 8, 0, HG, false>(a, pl, act, s);
}

#define FK_DEFINE_DISPATCH_FWD(PREC, HG)                                                                       \
  template <>                                                                                                  \
  int dispatch_fwd<PREC, HG>(const FwdArgs& a, const FwdPlan& pl, int cp, int sched, int nwave, int act,      \
                             bool rows, hipStream_t s) {                                                       \
    if constexpr (PREC == FK_PREC_FP32) {                                                                      \
      if (cp == 16) return dispatch_fwd_c<PREC, 16, HG>(a, pl, sched, nwave, act, rows, s);                   \
    }                                                                                                          \
    if (cp == 32) return dispatch_fwd_c<PREC, 32, HG>(a, pl, sched, nwave, act, rows, s);                     \
    return dispatch_fwd_c<PREC, 64, HG>(a, pl, sched, nwave, act, rows, s);                                   \
  }

}  // namespace fk
