// Fast autoregressive sampling for ConvNetAutoregressive2D
// (deepar/samplers/fast_autoregressive.py:13-76 with the per-site topologies
// of deepar/graph_analysis/{convolutional,padding,masking,sampling}_topology.py).
//
// One wavefront owns 16 samples (the MFMA N dimension) and walks the lattice
// in raster order.  Every (layer, site) activation is computed exactly once:
//   * site (i, j): the masked final block gives the conditional of (i, j) from
//     sites before it; the spin is drawn from the injected uniform with the
//     AutoregressiveSampler rule (exp(log p(+1)) > u); then the horizontal
//     stack of every other block is advanced by one site, the value of the
//     current site flowing between blocks in registers;
//   * after row i: the vertical stack of every block for row i, which also
//     produces the down-shifted y input of row i+1.
// Activations of earlier sites/rows live in a per-tile global-memory cache in
// MFMA B-fragment order (one 16-byte load per lane per pixel), so the state of
// 16 samples is (3*nb-2) lattices of 16 x C values.  The arithmetic, the
// accumulation order and the roundings are those of the fused forward kernel,
// so the draws equal the plain sampler's draws given the same uniforms.
#include "fk_common.h"
#include "convar_layout.h"
#include "convar_dev.h"

#include <type_traits>

namespace fk {

// Per-lane data of one pixel of a C-channel tensor for 16 samples: MT quads
// of 4 channels (channels 16*mt + 4*g + r for lane group g) in M-tile order.
template <int PREC, int C>
struct LaneData;

// 16-bit precisions (bf16, fp16): 8 bytes per quad
template <int PREC, int C>
struct LaneData {
  static constexpr int MT = C / 16;
  static constexpr int LB = MT * 8;  // bytes per lane
  uint2 q[MT];
  __device__ __forceinline__ void load(const uint8_t* p) {
#pragma unroll
    for (int k = 0; k < MT / 2; ++k) {
      const uint4 v = *reinterpret_cast<const uint4*>(p + 16 * k);
      q[2 * k] = make_uint2(v.x, v.y);
      q[2 * k + 1] = make_uint2(v.z, v.w);
    }
  }
  __device__ __forceinline__ void load_quads(const uint8_t* p, int m0, int m1) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
      if (m >= m0 && m < m1) q[m] = *reinterpret_cast<const uint2*>(p + 8 * m);
  }
  // the same loads through a buffer resource: per-lane offset voff, uniform
  // offset soff (bytes)
  __device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
#pragma unroll
    for (int k = 0; k < MT / 2; ++k) {
      const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16 * k, soff, 0);
      q[2 * k] = make_uint2(v.x, v.y);
      q[2 * k + 1] = make_uint2(v.z, v.w);
    }
  }
  __device__ __forceinline__ void bload_quads(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int m0, int m1) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
      if (m >= m0 && m < m1) {
        const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff + 8 * m, soff, 0);
        q[m] = make_uint2(v.x, v.y);
      }
  }
  __device__ __forceinline__ void store_quads(uint8_t* p, int m0, int m1) const {
#pragma unroll
    for (int m = 0; m < MT; ++m)
      if (m >= m0 && m < m1) *reinterpret_cast<uint2*>(p + 8 * m) = q[m];
  }
  __device__ __forceinline__ void store(uint8_t* p) const {
#pragma unroll
    for (int k = 0; k < MT / 2; ++k)
      *reinterpret_cast<uint4*>(p + 16 * k) = make_uint4(q[2 * k].x, q[2 * k].y, q[2 * k + 1].x, q[2 * k + 1].y);
  }
  __device__ __forceinline__ void bstore_quads(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int m0,
                                               int m1) const {
#pragma unroll
    for (int m = 0; m < MT; ++m)
      if (m >= m0 && m < m1) __builtin_amdgcn_raw_buffer_store_b64(u32x2{q[m].x, q[m].y}, r, voff + 8 * m, soff, 0);
  }
  __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) const {
#pragma unroll
    for (int k = 0; k < MT / 2; ++k)
      __builtin_amdgcn_raw_buffer_store_b128(u32x4{q[2 * k].x, q[2 * k].y, q[2 * k + 1].x, q[2 * k + 1].y}, r,
                                             voff + 16 * k, soff, 0);
  }
  __device__ __forceinline__ void set(int m, f32x4 v) {
    typename Prec<PREC>::quad h = to_quad<PREC>(v);
    q[m] = *reinterpret_cast<uint2*>(&h);
  }
  __device__ __forceinline__ f32x4 get(int m) const {
    return from_quad<PREC>(*reinterpret_cast<const typename Prec<PREC>::quad*>(&q[m]));
  }
  template <int KS>
  __device__ __forceinline__ void frags(typename Prec<PREC>::frag (&b)[KS]) const {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      uint4 v = make_uint4(q[2 * ks].x, q[2 * ks].y, q[2 * ks + 1].x, q[2 * ks + 1].y);
      b[ks] = *reinterpret_cast<typename Prec<PREC>::frag*>(&v);
    }
  }
};

template <int C>
struct LaneData<FK_PREC_FP32, C> {
  static constexpr int MT = C / 16;
  static constexpr int LB = MT * 16;
  f32x4 q[MT];
  __device__ __forceinline__ void load(const uint8_t* p) {
#pragma unroll
    for (int m = 0; m < MT; ++m) q[m] = *reinterpret_cast<const f32x4*>(p + 16 * m);
  }
  __device__ __forceinline__ void load_quads(const uint8_t* p, int m0, int m1) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
      if (m >= m0 && m < m1) q[m] = *reinterpret_cast<const f32x4*>(p + 16 * m);
  }
  __device__ __forceinline__ void store_quads(uint8_t* p, int m0, int m1) const {
#pragma unroll
    for (int m = 0; m < MT; ++m)
      if (m >= m0 && m < m1) *reinterpret_cast<f32x4*>(p + 16 * m) = q[m];
  }
  __device__ __forceinline__ void store(uint8_t* p) const { store_quads(p, 0, MT); }
  __device__ __forceinline__ void set(int m, f32x4 v) { q[m] = v; }
  __device__ __forceinline__ f32x4 get(int m) const { return q[m]; }
  template <int KS>
  __device__ __forceinline__ void frags(float (&b)[KS]) const {
#pragma unroll
    for (int t = 0; t < KS; ++t) b[t] = q[t >> 2][t & 3];
  }
};

// Per-lane activation loads of one horizontal-stack block step of the fast
// sampler (from the per-tile cache in HBM): the two earlier horizontal-stack
// pixels, the eight earlier concatenation pixels, and the y half of the
// current one (the batched form; otherwise only the current pixel).  With
// row partials (RP) the six concatenation pixels of the two rows above are
// replaced by the partial sum of their taps (FP32, bias included), computed
// for the whole row after the row above it.
template <int PREC, int C, bool BATCH, bool RP = false>
struct HLoads {
  LaneData<PREC, C> hd[BATCH ? 2 : 1];
  LaneData<PREC, C> cd[RP ? 2 : (BATCH ? 8 : 1)];
  LaneData<PREC, C> cur2;
  f32x4 r[RP ? C / 16 : 1];
};

// Per-lane activation loads of one pixel of the vertical-stack pass: the
// nine vertical inputs and the residual (batched form only).
template <int PREC, int C, bool BATCH>
struct VLoads {
  LaneData<PREC, C> vd[BATCH ? 9 : 1];
  LaneData<PREC, C> vr;
};

struct SampArgs {
  const void* w;
  const float* aux;
  int H, W, HW, nb;
  int64_t batch;
  const double* uniforms;
  int8_t* out;
  float* cond;
  uint8_t* cache;
  int64_t tile_bytes;
};

// WR: the whole-row vertical pass (16-bit 32-channel nets, rows of at most
// VROW_MAX pixels), with its rows in dynamic LDS (fast_vrow_lds bytes).
constexpr int VROW_MAX = 10;
template <int PREC, int C>
constexpr int fast_lb() {
  return PREC == FK_PREC_FP32 ? (C / 16) * 16 : (C / 16) * 8;
}
inline int fast_vrow_lds(int W, int lb) { return (6 * W + 1) * 64 * lb + 3 * W * 16 * 4; }

// Whether a block step's operands are all requested ahead of its MFMAs (the
// tap fragments are small: bf16/fp16 C=32), and the tap-loop unroll that goes
// with it.
template <int PREC, int C>
struct FastBatch {
  static constexpr bool value =
      (C / 16) * (C / Prec<PREC>::KW) * (int)(sizeof(typename Prec<PREC>::frag) / 4) <= 8;
  static constexpr int unr = value ? 9 : 1;
};

// Four consecutive biases of one lane group (16-byte aligned: every bias
// offset of the aux buffer is a multiple of 4 floats, and channels 16 m + 4 g
// .. + 3 stay inside the C or C/2 biases of the M-tiles that read them).
__device__ __forceinline__ f32x4 ldq4(const float* p) { return *reinterpret_cast<const f32x4*>(p); }

template <int PREC, int C, int ACT, bool WR>
__global__ __launch_bounds__(64) void fast_sample_kernel(SampArgs a) {
  using P = Prec<PREC>;
  using frag = typename P::frag;
  using LD = LaneData<PREC, C>;
  constexpr int MT = C / 16;
  constexpr int KS = C / P::KW;
  constexpr int MTH = MT / 2;  // quads of a C/2-channel 1x1 output
  constexpr int LB = LD::LB;
  constexpr int PB = 64 * LB;  // bytes per pixel per tile
  const ConvarLayout lay(C, PREC);

  const int lane = threadIdx.x;
  const int n = lane & 15, g = lane >> 4;
  const int H = a.H, W = a.W, HW = a.HW, nb = a.nb, F = nb - 1;
  const int64_t sample = (int64_t)blockIdx.x * 16 + n;
  const bool live = sample < a.batch;
  const frag* wall = reinterpret_cast<const frag*>(a.w);

  uint8_t* base = a.cache + (int64_t)blockIdx.x * a.tile_bytes;
  // every buffer holds HW pixels, at index HW a zero pixel that padding taps
  // read (their MFMAs add exact zeros, as the fused forward's do) and at HW+1
  // a pixel that takes the stores a step has nothing for, so the batched
  // steps have no branches around their memory operations
  const int64_t BUF = (int64_t)(HW + 2) * PB;
  auto CATp = [&](int l, int q) { return base + (int64_t)l * BUF + (int64_t)q * PB + lane * LB; };
  auto HAp = [&](int l, int q) { return base + (int64_t)(nb + l - 1) * BUF + (int64_t)q * PB + lane * LB; };
  auto VAp = [&](int l, int q) { return base + (int64_t)(2 * nb - 1 + l - 1) * BUF + (int64_t)q * PB + lane * LB; };
  float* spins = reinterpret_cast<float*>(base + (int64_t)(3 * nb - 2) * BUF);  // [HW][16]

  // Buffer resources over the weights, the biases and this tile's cache: a
  // load takes a per-lane 32-bit offset (mostly loop-invariant) and a uniform
  // offset in a scalar register, so no 64-bit address arithmetic per load.
  constexpr int BUF_FLAGS = 0x00020000;  // raw buffer, 32-bit elements (gfx9)
  const __amdgpu_buffer_rsrc_t rs_w = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.w), (short)0, 0x7fffffff, BUF_FLAGS);
  const __amdgpu_buffer_rsrc_t rs_a =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(a.aux), (short)0, 0x7fffffff, BUF_FLAGS);
  // (16-bit nets; the FP32 kernels, whose fragments are single floats, keep
  // plain global loads, which measured faster for them)
  constexpr bool BUFLD = Prec<PREC>::H16;
  auto wfrag = [&](int64_t fragidx) {
    if constexpr (BUFLD) {
      static_assert(sizeof(frag) == 16, "16-byte fragments");
      const int soff = (int)(fragidx * 64 * (int64_t)sizeof(frag));
      return __builtin_bit_cast(frag, __builtin_amdgcn_raw_buffer_load_b128(rs_w, lane * 16u, soff, 0));
    } else {
      return wall[fragidx * 64 + lane];
    }
  };
  const __amdgpu_buffer_rsrc_t rs_t =
      __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)(a.tile_bytes < 0x7fffffff ? a.tile_bytes : 0x7fffffff), BUF_FLAGS);
  // four biases of this lane's group: aux[fidx + 4 g .. + 3] (fidx uniform)
  auto ldaux4 = [&](int fidx) {
    if constexpr (BUFLD) return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs_a, 16u * g, fidx * 4, 0));
    else return ldq4(a.aux + fidx + 4 * g);
  };
  // batch every load of a block step ahead of its MFMAs when a tap's
  // fragments are small (bf16/fp16 C=32); wider nets load tap by tap
  constexpr bool BATCH = FastBatch<PREC, C>::value;
  constexpr int BT = BATCH ? 9 : 1;
  static_assert(!WR || BATCH, "the whole-row vertical pass needs the batched loads");
  constexpr int VROW = VROW_MAX;
  extern __shared__ __align__(16) uint8_t s_dyn[];

  for (int k = 0; k < 3 * nb - 2; ++k) LD{}.store(base + (int64_t)k * BUF + (int64_t)HW * PB + lane * LB);

  // row 0 of every block's concatenation: y = act(1x1(0) + b) (DownShift zeros)
  for (int l = 0; l <= F; ++l) {
    const float* by = a.aux + lay.aux_bias(l) + 2 * C + C / 2;
    LD d;
#pragma unroll
    for (int m = 0; m < MTH; ++m) d.set(MTH + m, act4<ACT>(ldq4(by + 16 * m + 4 * g)));
    for (int j = 0; j < W; ++j) d.store_quads(CATp(l, j), MTH, MT);
  }

  // Whole-row kernel (WR) staging: LDS rows of W pixels (the vertical pass
  // and the row partials below), then one zero pixel and the spins of three
  // rows.  Direct global-to-LDS loads move 16 bytes per lane,
  // lane-consecutive, which is the pixels' lane layout; every lane reads back
  // only its own bytes.
  uint8_t* const s_vrow = s_dyn;
  uint8_t* const vzero = s_dyn + 6 * W * PB + lane * LB;  // the zero pixel
  auto vrow = [&](int which, int c) { return s_vrow + ((which * W + c) * 64 + lane) * LB; };
  auto vdma = [&](const uint8_t* src, int which, int c) {
    static_assert(!WR || LB == 16, "direct LDS loads move 16 bytes per lane");
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)(src),
                                     (void __attribute__((address_space(3)))*)(s_vrow + (which * W + c) * PB), 16, 0,
                                     0);
  };

  // Row partials (WR).  R(l, j) of row i is the bias plus taps 0..5 of block
  // l's 3x3 concatenation conv at site (i, j) -- the taps that read rows i-2
  // and i-1 -- in FP32, for every block l = 0..F and column j.  The step of
  // site (i, j) continues the same accumulator with taps 6..8, so every sum
  // is the one the fused forward forms, in the same order.  Row i's partials
  // of block l are made in the vertical pass of row i-1, block l: the block's
  // two rows (VROW_MAX pixels each, in registers; columns from W on and rows
  // above the lattice read the cache's zero pixel) and weights are requested
  // with the pass's own operands, so they share its one wait per block.
  // Columns from W on fill spare slots.
  uint8_t* const rbuf = reinterpret_cast<uint8_t*>(spins + HW * 16);
  constexpr int RB = MT * 16;  // bytes per lane of one partial
  struct RPre {
    frag w[6][MT][KS];
    f32x4 b[MT];
    LD x[2][VROW_MAX];
  };
  // block l's operands for the partials of row i
  auto rfetch = [&](RPre& R, int l, int i) {
    const float* bl = a.aux + lay.aux_bias(l) + 3 * C;
#pragma unroll
    for (int m = 0; m < MT; ++m) R.b[m] = ldq4(bl + 16 * m + 4 * g);
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) R.w[t][m][ks] = wfrag(lay.off_c(l) + (t * MT + m) * KS + ks);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int row = i - 2 + r;
#pragma unroll
      for (int c = 0; c < VROW_MAX; ++c) {
        const int q = row >= 0 && c < W ? row * W + c : HW;
        if constexpr (BUFLD) R.x[r][c].bload(rs_t, lane * LB, (int)(l * BUF) + q * PB);
        else R.x[r][c].load(CATp(l, q));
      }
    }
  };
  // Pixels in groups of RG: the MFMAs of a group's pixels alternate, so the
  // accumulator chains overlap in the MFMA pipeline.
  constexpr int RG = 2;
  static_assert(VROW_MAX % RG == 0, "whole pixel groups");
  auto rpixels = [&](const RPre& R, int l) {
    uint8_t* const rl = rbuf + (int64_t)l * VROW_MAX * 64 * RB + lane * RB;
#pragma unroll
    for (int j0 = 0; j0 < VROW_MAX; j0 += RG) {
      f32x4 acc[RG][MT];
#pragma unroll
      for (int u = 0; u < RG; ++u)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[u][m] = R.b[m];
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        frag b[RG][KS];
#pragma unroll
        for (int u = 0; u < RG; ++u) {
          const int c = j0 + u + t % 3 - 2;
          if (c >= 0) {
            R.x[t / 3][c >= 0 ? c : 0].template frags<KS>(b[u]);
          } else {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) b[u][ks] = frag{};
          }
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int u = 0; u < RG; ++u) acc[u][m] = P::mfma(R.w[t][m][ks], b[u][ks], acc[u][m]);
      }
#pragma unroll
      for (int u = 0; u < RG; ++u)
#pragma unroll
        for (int m = 0; m < MT; ++m) *reinterpret_cast<f32x4*>(rl + (j0 + u) * 64 * RB + 16 * m) = acc[u][m];
    }
  };
  if constexpr (WR) {
    // the zero pixel of the LDS rows, and row 0's partials (bias and taps
    // that read the zero rows above the lattice)
    LD{}.store(vzero);
    for (int l = 0; l <= F; ++l) {
      RPre R;
      rfetch(R, l, 0);
      rpixels(R, l);
    }
  }

  for (int i = 0; i < H; ++i) {
    for (int j = 0; j < W; ++j) {
      const int q = i * W + j;
      // ============ conditional of site (i, j): masked final block =========
      {
        f32x4 z;
        if constexpr (BATCH) {
          // Every load of the conditional (weights, biases, the earlier
          // pixels it reads) is requested first and waited for once.
          const float* bias = a.aux + lay.aux_bias(F);
          // (whole-row kernel: the concatenation conv starts from the row
          // partial, so only its taps 6..8 and two earlier pixels remain)
          constexpr int CT0 = WR ? 6 : 0;
          f32x4 bh[MT], bc[MT], bxv[MTH];
#pragma unroll
          for (int m = 0; m < MT; ++m) bh[m] = ldq4(bias + C + 16 * m + 4 * g);
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            if constexpr (WR)
              bc[m] = *reinterpret_cast<const f32x4*>(rbuf + ((int64_t)F * VROW_MAX + j) * 64 * RB + lane * RB + 16 * m);
            else bc[m] = ldq4(bias + 3 * C + 16 * m + 4 * g);
          }
#pragma unroll
          for (int m = 0; m < MTH; ++m) bxv[m] = ldq4(bias + 2 * C + 16 * m + 4 * g);
          const f32x4 hb4 = ldq4(a.aux + lay.aux_head_bias(nb));  // the head's 4 biases (lane group 0)
          const f32x4 zb = g == 0 ? hb4 : f32x4{0.f, 0.f, 0.f, 0.f};
          frag wh[3][MT][KS], wx[MTH][KS], wc[9 - CT0][MT][KS], wz[KS];
#pragma unroll
          for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) wh[t][m][ks] = wfrag(lay.off_h(F) + (t * MT + m) * KS + ks);
#pragma unroll
          for (int m = 0; m < MTH; ++m)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) wx[m][ks] = wfrag(lay.off_x(F) + m * KS + ks);
#pragma unroll
          for (int t = CT0; t < 9; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) wc[t - CT0][m][ks] = wfrag(lay.off_c(F) + (t * MT + m) * KS + ks);
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) wz[ks] = wfrag(lay.head_base(nb) + ks);
          LD hd[3], cd[8 - CT0], cur;
#pragma unroll
          for (int t = 0; t < 3; ++t) hd[t].load(HAp(F, j - 3 + t >= 0 ? i * W + j - 3 + t : HW));
#pragma unroll
          for (int t = CT0; t < 8; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 2;
            cd[t - CT0].load(CATp(F, r >= 0 && c >= 0 ? r * W + c : HW));
          }
          cur.load_quads(CATp(F, q), MTH, MT);
          f32x4 x2[MTH];
          if (j >= 1) {
            f32x4 acc[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m] = bh[m];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
              frag b[KS];
              hd[t].template frags<KS>(b);
#pragma unroll
              for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) acc[m] = P::mfma(wh[t][m][ks], b[ks], acc[m]);
            }
            f32x4 ax[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) ax[m] = act4<ACT>(acc[m]);
            frag b[KS];
            d_to_b<PREC, MT, KS>(ax, b);
#pragma unroll
            for (int m = 0; m < MTH; ++m) {
              f32x4 x = bxv[m];
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) x = P::mfma(wx[m][ks], b[ks], x);
              x2[m] = act4<ACT>(x);
            }
          } else {
#pragma unroll
            for (int m = 0; m < MTH; ++m) x2[m] = act4<ACT>(bxv[m]);
          }
#pragma unroll
          for (int m = 0; m < MTH; ++m) cur.set(m, x2[m]);
          f32x4 acc[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[m] = bc[m];
#pragma unroll
          for (int t = CT0; t < 9; ++t) {
            frag b[KS];
            if (t == 8) cur.template frags<KS>(b);
            else cd[t < 8 ? t - CT0 : 0].template frags<KS>(b);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) acc[m] = P::mfma(wc[t - CT0][m][ks], b[ks], acc[m]);
          }
          f32x4 ah[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) ah[m] = act4<ACT>(acc[m]);
          frag b[KS];
          d_to_b<PREC, MT, KS>(ah, b);
          z = zb;
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) z = P::mfma(wz[ks], b[ks], z);
          cur.store_quads(CATp(F, q), 0, MTH);
        } else {
        const float* bias = a.aux + lay.aux_bias(F);
        f32x4 x2[MTH];
        const float* bx = bias + 2 * C;
        if (j >= 1) {
          f32x4 acc[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[m] = ldq4(bias + C + 16 * m + 4 * g);
          for (int t = 0; t < 3; ++t) {
            const int c = j - 1 + t - 2;
            if (c < 0) continue;
            LD hd;
            hd.load(HAp(F, i * W + c));
            frag b[KS];
            hd.template frags<KS>(b);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks)
                acc[m] = P::mfma(wfrag(lay.off_h(F) + (t * MT + m) * KS + ks), b[ks], acc[m]);
          }
          f32x4 ax[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) ax[m] = act4<ACT>(acc[m]);
          frag b[KS];
          d_to_b<PREC, MT, KS>(ax, b);
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            f32x4 x = ldq4(bx + 16 * m + 4 * g);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) x = P::mfma(wfrag(lay.off_x(F) + m * KS + ks), b[ks], x);
            x2[m] = act4<ACT>(x);
          }
        } else {
#pragma unroll
          for (int m = 0; m < MTH; ++m) x2[m] = act4<ACT>(ldq4(bx + 16 * m + 4 * g));
        }
        LD cur;
        cur.load_quads(CATp(F, q), MTH, MT);
#pragma unroll
        for (int m = 0; m < MTH; ++m) cur.set(m, x2[m]);
        cur.store_quads(CATp(F, q), 0, MTH);
        f32x4 acc[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m] = ldq4(bias + 3 * C + 16 * m + 4 * g);
        for (int t = 0; t < 9; ++t) {
          const int r = i + t / 3 - 2, c = j + t % 3 - 2;
          if (r < 0 || c < 0) continue;
          LD cd;
          if (t == 8) cd = cur;
          else cd.load(CATp(F, r * W + c));
          frag b[KS];
          cd.template frags<KS>(b);
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
              acc[m] = P::mfma(wfrag(lay.off_c(F) + (t * MT + m) * KS + ks), b[ks], acc[m]);
        }
        f32x4 ah[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) ah[m] = act4<ACT>(acc[m]);
        frag b[KS];
        d_to_b<PREC, MT, KS>(ah, b);
        z = load_bias4(a.aux + lay.aux_head_bias(nb), 4 * g, 4);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) z = P::mfma(wfrag(lay.head_base(nb) + ks), b[ks], z);
        }
        float s = 0.0f;
        if (g == 0) {
          const float mx = fmaxf(z.x, z.y);
          const float lse = mx + 0.5f * log1pf(expf(-2.0f * fabsf(z.x - z.y)));
          const float lp0 = 2.0f * (z.x - lse);
          if (live) {
            const double u = a.uniforms[sample * HW + q];
            s = ((double)expf(lp0) > u) ? 1.0f : -1.0f;
            a.out[sample * HW + q] = (int8_t)s;
            if (a.cond) {
              a.cond[(sample * HW + q) * 2] = lp0;
              a.cond[(sample * HW + q) * 2 + 1] = 2.0f * (z.y - lse);
            }
          } else {
            s = 1.0f;
          }
          spins[q * 16 + n] = s;
        }
        // later sites read this spin from other lanes of the wave
        __threadfence_block();
        s = __shfl(s, n, 64);

        // ============ advance the horizontal stack of blocks 0..F-1 ========
        f32x4 hcur[MT], hres[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) hcur[m] = hres[m] = f32x4{0.f, 0.f, 0.f, 0.f};
        // The activation loads of block step l (earlier sites and rows, from
        // HBM) do not depend on this site's chain.  With batched loads the
        // steps go in pairs over two register sets: step l+1's activations
        // are requested before step l computes, so their latency hides behind
        // step l's dependent MFMA chain; weights (L2) are requested at the
        // start of their own step.
        // this site's pixel offsets inside a buffer (32-bit, per lane): the
        // loads then take a uniform buffer base plus one offset register
        uint32_t o_hd[2], o_cd[8];
#pragma unroll
        for (int t = 0; t < 2; ++t) o_hd[t] = (uint32_t)((j + t - 2 >= 0 ? i * W + j + t - 2 : HW) * PB + lane * LB);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int r = i + t / 3 - 2, c = j + t % 3 - 2;
          o_cd[t] = (uint32_t)((r >= 0 && c >= 0 ? r * W + c : HW) * PB + lane * LB);
        }
        const uint32_t o_cur = (uint32_t)(q * PB + lane * LB);
        // whole-row kernel: the concatenation conv of a step starts from the
        // row partial and reads two earlier pixels (taps 6, 7) and the current
        constexpr int CT0 = WR ? 6 : 0;
        const uint32_t o_r = (uint32_t)(j * 64 * RB + lane * RB);
        // (tile-relative byte offsets: buffers at l * BUF, partials at rbuf)
        const int rbuf_off = (int)(rbuf - base);
        auto issue = [&](HLoads<PREC, C, BATCH, WR>& L, int l) {
          const int cs = (int)(l * BUF);
          if constexpr (BATCH) {
            if (l > 0) {
              const int hs = (int)((nb + l - 1) * BUF);
#pragma unroll
              for (int t = 0; t < 2; ++t) L.hd[t].bload(rs_t, o_hd[t], hs);
            }
#pragma unroll
            for (int t = CT0; t < 8; ++t) L.cd[t - CT0].bload(rs_t, o_cd[t], cs);
            if constexpr (WR) {
              const int rs = rbuf_off + l * VROW_MAX * 64 * RB;
#pragma unroll
              for (int m = 0; m < MT; ++m)
                L.r[m] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs_t, o_r + 16 * m, rs, 0));
            }
          }
          if constexpr (BUFLD) L.cur2.bload_quads(rs_t, o_cur, cs, MTH, MT);
          else L.cur2.load_quads(base + cs + o_cur, MTH, MT);
        };
        // One block step: its own loads (biases, weights, the first block's
        // spins) are requested first, then the next step's activations
        // (Ln, ln; ln < 0: none), then the step computes.  Waiting for a
        // load waits for every load issued before it, so this order lets
        // the step start on its own operands while the prefetch is in flight.
        // A step's two stores (its concatenation x half and its output) are
        // issued by the next step after that step's loads: vector memory
        // operations complete in issue order, so a store issued before a load
        // would make the wait for the load wait for the store as well.
        LD pcx, phn;
        uint8_t* const trash = base + (int64_t)(HW + 1) * PB + lane * LB;
        uint8_t *ppc = trash, *pph = trash;
        // (16-bit: the pending stores as uniform offsets from this pixel's
        // offset o_cur; the trash pixel of buffer 0 is HW + 1 - q pixels on)
        const int strash = (HW + 1 - q) * PB;
        int spc = strash, sph = strash;
        auto flush = [&]() {
          if constexpr (BUFLD) {
            pcx.bstore_quads(rs_t, o_cur, spc, 0, MTH);
            phn.bstore(rs_t, o_cur, sph);
            spc = sph = strash;
          } else {
            pcx.store_quads(ppc, 0, MTH);
            phn.store(pph);
            ppc = pph = trash;
          }
        };
        // FIRST: block 0, whose horizontal input is the spins
        auto step = [&](auto first, HLoads<PREC, C, BATCH, WR>& L, int l, HLoads<PREC, C, BATCH, WR>& Ln, int ln) {
          constexpr bool FIRST = decltype(first)::value;
          const int kind = block_kind(l, nb);
          const float* bl = a.aux + lay.aux_bias(l);
          f32x4 hacc[MT], cacc[MT], xb[MTH];
          const int bo = lay.aux_bias(l);
#pragma unroll
          for (int m = 0; m < MT; ++m) hacc[m] = ldaux4(bo + C + 16 * m);
          if constexpr (!WR) {
#pragma unroll
            for (int m = 0; m < MT; ++m) cacc[m] = ldaux4(bo + 3 * C + 16 * m);
          }
#pragma unroll
          for (int m = 0; m < MTH; ++m) xb[m] = ldaux4(bo + 2 * C + 16 * m);
          frag wh[BATCH ? 3 : 1][MT][KS], wc[BATCH ? 9 - CT0 : 1][MT][KS], wx[MTH][KS];
#pragma unroll
          for (int m = 0; m < MTH; ++m)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) wx[m][ks] = wfrag(lay.off_x(l) + m * KS + ks);
          float sv[3] = {0.0f, 0.0f, 0.0f};
          f32x4 kh0[FIRST ? 3 : 1][MT];
          if constexpr (FIRST) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
              const int c = j + t - 2;
              const float v = spins[(i * W + (c >= 0 ? c : 0)) * 16 + n];
              sv[t] = c >= 0 ? v : 0.0f;
            }
            sv[2] = s;
            const float* kh = a.aux + lay.aux_kh0(nb);
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
              for (int m = 0; m < MT; ++m) kh0[t][m] = ldq4(kh + t * C + 16 * m + 4 * g);
          } else if constexpr (BATCH) {
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
              for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) wh[t][m][ks] = wfrag(lay.off_h(l) + (t * MT + m) * KS + ks);
          }
          if constexpr (BATCH) {
#pragma unroll
            for (int t = CT0; t < 9; ++t)
#pragma unroll
              for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) wc[t - CT0][m][ks] = wfrag(lay.off_c(l) + (t * MT + m) * KS + ks);
          }
          // scheduling fences: the compiler would otherwise sink loads next
          // to their first use, and a load issued after the prefetch makes
          // its wait cover the prefetch too
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (BATCH) issue(Ln, ln);
          flush();
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (FIRST) {
#pragma unroll
            for (int t = 0; t < 3; ++t) {
              if (!BATCH && j + t - 2 < 0) continue;  // (batched: sv = 0 adds zeros)
#pragma unroll
              for (int m = 0; m < MT; ++m) hacc[m] += kh0[t][m] * sv[t];
            }
          } else {
#pragma unroll FastBatch<PREC, C>::unr
            for (int t = 0; t < 3; ++t) {
              const int c = j + t - 2;
              if (!BATCH && c < 0) continue;  // (batched: the zero pixel)
              frag b[KS];
              if (t == 2) {
                d_to_b<PREC, MT, KS>(hcur, b);
              } else if constexpr (BATCH) {
                L.hd[t].template frags<KS>(b);
              } else {
                LD hd;
                hd.load(HAp(l, i * W + c));
                hd.template frags<KS>(b);
              }
#pragma unroll
              for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                  frag wk;
                  if constexpr (BATCH) wk = wh[t][m][ks];
                  else wk = wfrag(lay.off_h(l) + (t * MT + m) * KS + ks);
                  hacc[m] = P::mfma(wk, b[ks], hacc[m]);
                }
            }
          }
          f32x4 ax[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) ax[m] = act4<ACT>(hacc[m]);
          frag bxf[KS];
          d_to_b<PREC, MT, KS>(ax, bxf);
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            f32x4 x = xb[m];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) x = P::mfma(wx[m][ks], bxf[ks], x);
            L.cur2.set(m, act4<ACT>(x));
          }
          pcx = L.cur2;
          if constexpr (BUFLD) spc = (int)(l * BUF);
          else ppc = CATp(l, q);
          if constexpr (WR) {
#pragma unroll
            for (int m = 0; m < MT; ++m) cacc[m] = L.r[m];
          }
#pragma unroll FastBatch<PREC, C>::unr
          for (int t = CT0; t < 9; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 2;
            if (!BATCH && (r < 0 || c < 0)) continue;  // (batched: the zero pixel)
            frag b[KS];
            if (t == 8) {
              L.cur2.template frags<KS>(b);
            } else if constexpr (BATCH) {
              L.cd[t - CT0].template frags<KS>(b);
            } else {
              LD cd;
              cd.load(CATp(l, r * W + c));
              cd.template frags<KS>(b);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) {
                frag wk;
                if constexpr (BATCH) wk = wc[t - CT0][m][ks];
                else wk = wfrag(lay.off_c(l) + (t * MT + m) * KS + ks);
                cacc[m] = P::mfma(wk, b[ks], cacc[m]);
              }
          }
          LD hn;
          if (kind == BK_A) {
#pragma unroll
            for (int m = 0; m < MT; ++m) hres[m] = hcur[m];
          }
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            const f32x4 v = (kind == BK_B) ? act4<ACT>(hres[m] + cacc[m]) : act4<ACT>(cacc[m]);
            hn.set(m, v);
          }
          phn = hn;
          if constexpr (BUFLD) sph = (int)((nb + l) * BUF);
          else pph = HAp(l + 1, q);
          // the value the next block sees is the stored (rounded) one
#pragma unroll
          for (int m = 0; m < MT; ++m) hcur[m] = hn.get(m);
        };
        HLoads<PREC, C, BATCH, WR> LA, LB;
        if constexpr (BATCH) {
          // every step requests a next step's activations (the last one its
          // own again), so the memory operations per step are fixed
          if (F > 0) {
            issue(LA, 0);
            step(std::true_type{}, LA, 0, LB, 1 < F ? 1 : 0);
          }
          int l = 1;
          for (; l + 1 < F; l += 2) {
            step(std::false_type{}, LB, l, LA, l + 1);
            step(std::false_type{}, LA, l + 1, LB, l + 2 < F ? l + 2 : l + 1);
          }
          if (l < F) step(std::false_type{}, LB, l, LA, l);
        } else {
          for (int l = 0; l < F; ++l) {
            issue(LA, l);
            if (l == 0) step(std::true_type{}, LA, l, LB, -1);
            else step(std::false_type{}, LA, l, LB, -1);
          }
        }
        flush();
      }
    }

    // ============ vertical stack of row i (all blocks) ======================
    if constexpr (WR) {
      {
        // Whole-row form.  Per block, rows i-2 and i-1 of its input (written
        // by earlier rows' passes) are requested together with its weights,
        // so the block waits on memory once, and are staged in LDS; row i of
        // its input is the previous block's output, kept in LDS (the stored,
        // rounded values), and so is the residual (the input of the block
        // before).  LDS rows: 0, 1 (4, 5 for odd blocks) = input rows i-2,
        // i-1; 2 + cur = row i of this block's input, 2 + (cur ^ 1) = the
        // previous block's, which pixel j replaces with its output once it
        // has read its residual; then one zero pixel and the spins of rows
        // i-2..i.
        // Every lane reads back only the bytes it wrote, except the spins
        // (a barrier after they are staged).
        int cur = 0;
        float* s_vspin = reinterpret_cast<float*>(s_dyn + 6 * W * PB + PB);
        for (int idx = lane; idx < 3 * W * 16; idx += 64) {
          const int r = idx / (W * 16), c = (idx / 16) % W, nn = idx % 16;
          const int rr = i + r - 2;
          s_vspin[idx] = rr >= 0 ? spins[(rr * W + c) * 16 + nn] : 0.0f;
        }
        __syncthreads();
        // Block l's operands -- weights, biases and rows i-2, i-1 of its
        // input -- are requested before block l-1's pixel loop (which issues
        // no loads), in one of two register sets, so their latency hides
        // behind that loop.
        struct VPre {
          frag wv[9][MT][KS], wy[MTH][KS];
          f32x4 bv0[MT], by0[MTH];
        };
        // LDS rows of block l's input rows i-2, i-1: 0, 1 for even l and
        // 4, 5 for odd l (filled by direct global-to-LDS loads, whose lane
        // layout -- 16 bytes per lane, lane-consecutive -- is the rows')
        auto vbase = [&](int l) { return (l & 1) ? 4 : 0; };
        auto vrows = [&](int l) {
          for (int c = 0; c < W; ++c) {
            if (i >= 2) vdma(VAp(l, (i - 2) * W + c), vbase(l), c);
            if (i >= 1) vdma(VAp(l, (i - 1) * W + c), vbase(l) + 1, c);
          }
        };
        auto vfetch = [&](VPre& V, int l) {
          const float* bl = a.aux + lay.aux_bias(l);
#pragma unroll
          for (int m = 0; m < MT; ++m) V.bv0[m] = ldq4(bl + 16 * m + 4 * g);
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            V.by0[m] = ldq4(bl + 2 * C + C / 2 + 16 * m + 4 * g);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) V.wy[m][ks] = wfrag(lay.off_y(l) + m * KS + ks);
          }
#pragma unroll
          for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) V.wv[t][m][ks] = wfrag(lay.off_v(l) + (t * MT + m) * KS + ks);
        };
        // the pixel loop of block l (FIRST: block 0, whose input is the
        // spins); V holds its operands, all landed
        auto vpixels = [&](auto first, const VPre& V, const f32x4 (&kv0)[9][MT], int l) {
          constexpr bool FIRST = decltype(first)::value;
          const int kind = block_kind(l, nb);
          // pixels in pairs (the second of an odd row's last pair reads only
          // the zero pixel and stores nothing): the two pixels' MFMAs
          // alternate, so their accumulator chains overlap
          for (int j0 = 0; j0 < W; j0 += 2) {
            f32x4 acc[2][MT];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
              for (int m = 0; m < MT; ++m) acc[u][m] = V.bv0[m];
            if constexpr (FIRST) {
#pragma unroll
              for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                  const int c = j0 + u + t % 3 - 1;
                  const float sv = (c >= 0 && c < W) ? s_vspin[((t / 3) * W + c) * 16 + n] : 0.0f;
#pragma unroll
                  for (int m = 0; m < MT; ++m) acc[u][m] += kv0[t][m] * sv;
                }
            } else {
              // every tap's fragment first (padding taps read the zero pixel:
              // their MFMAs add exact zeros, as the fused forward's do), then
              // the MFMAs
              LD x[2][9];
#pragma unroll
              for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                  const int r = i + t / 3 - 2, c = j0 + u + t % 3 - 1;
                  const bool pad = r < 0 || c < 0 || c >= W;
                  x[u][t].load(pad ? vzero : vrow(t / 3 == 2 ? 2 + cur : vbase(l) + t / 3, c));
                }
#pragma unroll
              for (int t = 0; t < 9; ++t) {
                frag b[2][KS];
#pragma unroll
                for (int u = 0; u < 2; ++u) x[u][t].template frags<KS>(b[u]);
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                    for (int u = 0; u < 2; ++u) acc[u][m] = P::mfma(V.wv[t][m][ks], b[u][ks], acc[u][m]);
              }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              const int j = j0 + u;
              if (j >= W) break;
              const int q = i * W + j;
              f32x4 av[MT];
#pragma unroll
              for (int m = 0; m < MT; ++m) av[m] = act4<ACT>(acc[u][m]);
              if (l < F) {
                LD vn;
                if (kind == BK_B) {
                  LD vr;
                  vr.load(vrow(2 + (cur ^ 1), j));
#pragma unroll
                  for (int m = 0; m < MT; ++m) vn.set(m, act4<ACT>(vr.get(m) + acc[u][m]));
                } else {
#pragma unroll
                  for (int m = 0; m < MT; ++m) vn.set(m, av[m]);
                }
                vn.store(VAp(l + 1, q));
                vn.store(vrow(2 + (cur ^ 1), j));
              }
              if (i + 1 < H) {
                frag b[KS];
                d_to_b<PREC, MT, KS>(av, b);
                LD yd;
#pragma unroll
                for (int m = 0; m < MTH; ++m) {
                  f32x4 y = V.by0[m];
#pragma unroll
                  for (int ks = 0; ks < KS; ++ks) y = P::mfma(V.wy[m][ks], b[ks], y);
                  yd.set(MTH + m, act4<ACT>(y));
                }
                yd.store_quads(CATp(l, q + W), MTH, MT);
              }
            }
          }
          cur ^= 1;
        };
        // block l >= 1 with its operands in V and its rows in LDS (requested
        // earlier): wait for them (and for every earlier store), request
        // block ln's operands into Vn, then run the pixel loop
        f32x4 kv0[9][MT];
        VPre V;
        // (with row i+1's partials of the same block)
        const bool rnext = i + 1 < H;
        auto vblock = [&](int l) {
          vfetch(V, l);
          RPre RR;
          if (rnext) rfetch(RR, l, i + 1);
          __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
          if (l + 1 <= F) vrows(l + 1);
          if (rnext) rpixels(RR, l);
          vpixels(std::false_type{}, V, kv0, l);
        };
        VPre& VB = V;
        {
          // block 0: its own operands, then block 1's into VA
          const float* kv = a.aux + lay.aux_kv0(nb);
#pragma unroll
          for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m) kv0[t][m] = *reinterpret_cast<const f32x4*>(kv + t * C + 16 * m + 4 * g);
          const float* bl = a.aux + lay.aux_bias(0);
#pragma unroll
          for (int m = 0; m < MT; ++m) VB.bv0[m] = ldq4(bl + 16 * m + 4 * g);
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            VB.by0[m] = ldq4(bl + 2 * C + C / 2 + 16 * m + 4 * g);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) VB.wy[m][ks] = wfrag(lay.off_y(0) + m * KS + ks);
          }
          RPre RR;
          if (rnext) rfetch(RR, 0, i + 1);
          __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
          if (F >= 1) vrows(1);
          if (rnext) rpixels(RR, 0);
          vpixels(std::true_type{}, VB, kv0, 0);
        }
        for (int l = 1; l <= F; ++l) vblock(l);
        __syncthreads();  // the spins staging is rewritten by the next row
      }
    } else {
    for (int l = 0; l <= F; ++l) {
      const int kind = block_kind(l, nb);
      const float* bl = a.aux + lay.aux_bias(l);
      // this block's weights and biases serve the whole row
      frag wv[BT][MT][KS], wy[MTH][KS];
      f32x4 bv0[MT], by0[MTH];
#pragma unroll
      for (int m = 0; m < MT; ++m) bv0[m] = ldq4(bl + 16 * m + 4 * g);
#pragma unroll
      for (int m = 0; m < MTH; ++m) {
        by0[m] = ldq4(bl + 2 * C + C / 2 + 16 * m + 4 * g);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) wy[m][ks] = wfrag(lay.off_y(l) + m * KS + ks);
      }
      if constexpr (BATCH) {
        if (l > 0) {
#pragma unroll
          for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) wv[t][m][ks] = wfrag(lay.off_v(l) + (t * MT + m) * KS + ks);
        }
      }
      // the activations pixel j reads (the vertical inputs of rows i-2..i
      // and the residual) were written by earlier blocks' passes, so with
      // batched loads pixel j+1's are requested before pixel j computes
      auto vissue = [&](VLoads<PREC, C, BATCH>& L, int j) {
        if constexpr (BATCH) {
          if (l > 0) {
#pragma unroll
            for (int t = 0; t < 9; ++t) {
              const int r = i + t / 3 - 2, c = j + t % 3 - 1;
              if (r >= 0 && c >= 0 && c < W) L.vd[t].load(VAp(l, r * W + c));
            }
          }
          if (l < F && kind == BK_B) L.vr.load(VAp(l - 1, i * W + j));
        }
      };
      auto vstep = [&](VLoads<PREC, C, BATCH>& L, int j, VLoads<PREC, C, BATCH>& Ln, int jn) {
        const int q = i * W + j;
        f32x4 acc[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m] = bv0[m];
        float sv[9];
        if (l == 0) {
          for (int t = 0; t < 9; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 1;
            sv[t] = (r >= 0 && c >= 0 && c < W) ? spins[(r * W + c) * 16 + n] : 0.0f;
          }
        }
        if (jn >= 0) vissue(Ln, jn);
        if (l == 0) {
          const float* kv = a.aux + lay.aux_kv0(nb);
          for (int t = 0; t < 9; ++t) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
              acc[m] += *reinterpret_cast<const f32x4*>(kv + t * C + 16 * m + 4 * g) * sv[t];
          }
        } else {
#pragma unroll FastBatch<PREC, C>::unr
          for (int t = 0; t < 9; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 1;
            if (r < 0 || c < 0 || c >= W) continue;
            frag b[KS];
            if constexpr (BATCH) {
              L.vd[t].template frags<KS>(b);
            } else {
              LD vd;
              vd.load(VAp(l, r * W + c));
              vd.template frags<KS>(b);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) {
                frag wk;
                if constexpr (BATCH) wk = wv[t][m][ks];
                else wk = wfrag(lay.off_v(l) + (t * MT + m) * KS + ks);
                acc[m] = P::mfma(wk, b[ks], acc[m]);
              }
          }
        }
        f32x4 av[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = act4<ACT>(acc[m]);
        if (l < F) {
          LD vn;
          if (kind == BK_B) {
            LD vr;
            if constexpr (BATCH) vr = L.vr;
            else vr.load(VAp(l - 1, q));
#pragma unroll
            for (int m = 0; m < MT; ++m) vn.set(m, act4<ACT>(vr.get(m) + acc[m]));
          } else {
#pragma unroll
            for (int m = 0; m < MT; ++m) vn.set(m, av[m]);
          }
          vn.store(VAp(l + 1, q));
        }
        if (i + 1 < H) {
          frag b[KS];
          d_to_b<PREC, MT, KS>(av, b);
          LD yd;
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            f32x4 y = by0[m];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) y = P::mfma(wy[m][ks], b[ks], y);
            yd.set(MTH + m, act4<ACT>(y));
          }
          yd.store_quads(CATp(l, q + W), MTH, MT);
        }
      };
      VLoads<PREC, C, BATCH> VA, VB;
      if constexpr (BATCH) {
        vissue(VA, 0);
        for (int j = 0; j < W; j += 2) {
          vstep(VA, j, VB, j + 1 < W ? j + 1 : -1);
          if (j + 1 < W) vstep(VB, j + 1, VA, j + 2 < W ? j + 2 : -1);
        }
      } else {
        for (int j = 0; j < W; ++j) vstep(VA, j, VB, -1);
      }
    }
    }
  }
}

// the fast sampler runs the 32- and 64-channel kernel widths (networks of
// other even widths up to 64 are zero-padded to them, see kernel_channels;
// the fp32 16-wide packing has no fast-sampler kernel)
static int fast_channels(int channels, int precision) {
  const int cp = kernel_channels(channels, precision);
  return cp == 16 ? 0 : cp;
}
bool fast_sampler_supported(int channels, int precision) { return fast_channels(channels, precision) > 0; }

// whether a net takes the whole-row kernel (16-bit 32-channel nets, rows of
// at most VROW_MAX pixels; see launch_fast)
static bool fast_whole_row(int W, int channels, int precision) {
  return precision != FK_PREC_FP32 && channels == 32 && W <= VROW_MAX;
}

// Per-tile cache: (3 nb - 2) activation buffers of HW + 2 pixels, the spins
// and, for the whole-row kernel, one row of concatenation partial sums per
// block (FP32, MT quads per lane, VROW_MAX columns).
int64_t fast_tile_bytes(int H, int W, int nb, int channels, int precision) {
  const int esz = precision == FK_PREC_FP32 ? 4 : 2;
  const int64_t pb = 16LL * channels * esz;
  const int64_t HW = (int64_t)H * W;
  int64_t b = (int64_t)(3 * nb - 2) * (HW + 2) * pb + HW * 16 * 4;
  if (fast_whole_row(W, channels, precision)) b += (int64_t)nb * VROW_MAX * 64 * (channels / 16) * 16;
  return b;
}

template <int PREC, int C, bool WR>
static int launch_fast_wr(const SampArgs& a, int act, int64_t tiles, hipStream_t s) {
  const int lds = WR ? fast_vrow_lds(a.W, fast_lb<PREC, C>()) : 0;
  if (WR && lds > 64 * 1024) {
    auto k = act == FK_ACT_ELU ? fast_sample_kernel<PREC, C, FK_ACT_ELU, WR> : fast_sample_kernel<PREC, C, FK_ACT_RELU, WR>;
    FK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  }
  if (act == FK_ACT_ELU)
    hipLaunchKernelGGL((fast_sample_kernel<PREC, C, FK_ACT_ELU, WR>), dim3((unsigned)tiles), dim3(64), lds, s, a);
  else
    hipLaunchKernelGGL((fast_sample_kernel<PREC, C, FK_ACT_RELU, WR>), dim3((unsigned)tiles), dim3(64), lds, s, a);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

// 16-bit 32-channel nets with rows of at most VROW_MAX pixels take the
// whole-row vertical pass
template <int PREC, int C>
static int launch_fast(const SampArgs& a, int act, int64_t tiles, hipStream_t s) {
  if constexpr (PREC != FK_PREC_FP32 && C == 32) {
    if (fast_whole_row(a.W, C, PREC)) return launch_fast_wr<PREC, C, true>(a, act, tiles, s);
  }
  return launch_fast_wr<PREC, C, false>(a, act, tiles, s);
}

}  // namespace fk

using namespace fk;

extern "C" {

int64_t fk_fast_sampler_workspace_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t precision,
                                        int64_t batch) {
  if (H < 1 || W < 1 || depth < 1 || batch < 0 || !fast_sampler_supported(channels, precision)) return -1;
  const int64_t tiles = (batch + 15) / 16;
  return tiles * fast_tile_bytes(H, W, num_causal_blocks(depth), fast_channels(channels, precision), precision);
}

int fk_fast_sample(const void* wpacked, const float* aux, int32_t precision, int32_t H, int32_t W, int32_t depth,
                   int32_t channels, int32_t activation, int64_t batch, const double* uniforms, void* workspace,
                   int8_t* out, float* cond_logp, void* stream) {
  FK_REQUIRE(H >= 1 && W >= 1 && depth >= 1 && batch >= 0, "fk_fast_sample: bad shape");
  FK_REQUIRE(activation == FK_ACT_RELU || activation == FK_ACT_ELU, "fk_fast_sample: bad activation");
  if (!fast_sampler_supported(channels, precision)) {
    set_error("fast sampler: channels=%d precision=%d not implemented (17..64 channels, bf16/fp16 2..64)", channels,
              precision);
    return FK_ERR_UNSUPPORTED;
  }
  if (batch == 0) return FK_OK;
  FK_REQUIRE(wpacked && aux && uniforms && workspace && out, "fk_fast_sample: null pointer");
  SampArgs a;
  a.w = wpacked;
  a.aux = aux;
  a.H = H;
  a.W = W;
  a.HW = H * W;
  a.nb = num_causal_blocks(depth);
  a.batch = batch;
  a.uniforms = uniforms;
  a.out = out;
  a.cond = cond_logp;
  a.cache = reinterpret_cast<uint8_t*>(workspace);
  const int CP = fast_channels(channels, precision);
  a.tile_bytes = fast_tile_bytes(H, W, a.nb, CP, precision);
  const int64_t tiles = (batch + 15) / 16;
  FK_REQUIRE(tiles <= 0x7fffffff, "fk_fast_sample: batch too large");
  hipStream_t s = (hipStream_t)stream;
  if (precision == FK_PREC_BF16) {
    if (CP == 32) return launch_fast<FK_PREC_BF16, 32>(a, activation, tiles, s);
    return launch_fast<FK_PREC_BF16, 64>(a, activation, tiles, s);
  }
  if (precision == FK_PREC_FP16) {
    if (CP == 32) return launch_fast<FK_PREC_FP16, 32>(a, activation, tiles, s);
    return launch_fast<FK_PREC_FP16, 64>(a, activation, tiles, s);
  }
  if (CP == 32) return launch_fast<FK_PREC_FP32, 32>(a, activation, tiles, s);
  return launch_fast<FK_PREC_FP32, 64>(a, activation, tiles, s);
}

}  // extern "C"
