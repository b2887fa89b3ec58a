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

template <int PREC, int C, int ACT>
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
  const int64_t BUF = (int64_t)HW * PB;
  auto CATp = [&](int l, int q) { return base + (int64_t)l * BUF + (int64_t)q * PB + lane * LB; };
  auto HAp = [&](int l, int q) { return base + (int64_t)(nb + l - 1) * BUF + (int64_t)q * PB + lane * LB; };
  auto VAp = [&](int l, int q) { return base + (int64_t)(2 * nb - 1 + l - 1) * BUF + (int64_t)q * PB + lane * LB; };
  float* spins = reinterpret_cast<float*>(base + (int64_t)(3 * nb - 2) * BUF);  // [HW][16]

  auto wfrag = [&](int64_t fragidx) { return wall[fragidx * 64 + lane]; };
  // batch every load of a block step ahead of its MFMAs when a tap's
  // fragments are small (bf16/fp16 C=32); wider nets load tap by tap
  constexpr bool BATCH = MT * KS * (int)(sizeof(frag) / 4) <= 8;
  constexpr int BT = BATCH ? 9 : 1;
  constexpr int UNR = BATCH ? 9 : 1;  // tap loops unrolled only with batched loads

  // row 0 of every block's concatenation: y = act(1x1(0) + b) (DownShift zeros)
  for (int l = 0; l <= F; ++l) {
    const float* by = a.aux + lay.aux_bias(l) + 2 * C + C / 2;
    LD d;
#pragma unroll
    for (int m = 0; m < MTH; ++m) d.set(MTH + m, act4<ACT>(load_bias4(by, 16 * m + 4 * g, C / 2)));
    for (int j = 0; j < W; ++j) d.store_quads(CATp(l, j), MTH, MT);
  }

  for (int i = 0; i < H; ++i) {
    for (int j = 0; j < W; ++j) {
      const int q = i * W + j;
      // ============ conditional of site (i, j): masked final block =========
      {
        const float* bias = a.aux + lay.aux_bias(F);
        f32x4 x2[MTH];
        const float* bx = bias + 2 * C;
        if (j >= 1) {
          f32x4 acc[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[m] = load_bias4(bias + C, 16 * m + 4 * g, C);
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
            f32x4 x = load_bias4(bx, 16 * m + 4 * g, C / 2);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) x = P::mfma(wfrag(lay.off_x(F) + m * KS + ks), b[ks], x);
            x2[m] = act4<ACT>(x);
          }
        } else {
#pragma unroll
          for (int m = 0; m < MTH; ++m) x2[m] = act4<ACT>(load_bias4(bx, 16 * m + 4 * g, C / 2));
        }
        LD cur;
        cur.load_quads(CATp(F, q), MTH, MT);
#pragma unroll
        for (int m = 0; m < MTH; ++m) cur.set(m, x2[m]);
        cur.store_quads(CATp(F, q), 0, MTH);
        f32x4 acc[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m] = load_bias4(bias + 3 * C, 16 * m + 4 * g, C);
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
        f32x4 z = load_bias4(a.aux + lay.aux_head_bias(nb), 4 * g, 4);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) z = P::mfma(wfrag(lay.head_base(nb) + ks), b[ks], z);
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
        for (int l = 0; l < F; ++l) {
          const int kind = block_kind(l, nb);
          const float* bl = a.aux + lay.aux_bias(l);
          f32x4 hacc[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) hacc[m] = load_bias4(bl + C, 16 * m + 4 * g, C);
          if (l == 0) {
            const float* kh = a.aux + lay.aux_kh0(nb);
            for (int t = 0; t < 3; ++t) {
              const int c = j + t - 2;
              const float sv = (c < 0) ? 0.0f : (c == j ? s : spins[(i * W + c) * 16 + n]);
#pragma unroll
              for (int m = 0; m < MT; ++m)
                hacc[m] += *reinterpret_cast<const f32x4*>(kh + t * C + 16 * m + 4 * g) * sv;
            }
          } else {
            // every load of this step is issued before the first MFMA: the
            // step is a chain of dependent MFMAs, so serialised L2 round trips
            // would dominate it
            frag wh[BATCH ? 3 : 1][MT][KS];
            LD hd[BATCH ? 2 : 1];
            if constexpr (BATCH) {
#pragma unroll
              for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                  for (int ks = 0; ks < KS; ++ks) wh[t][m][ks] = wfrag(lay.off_h(l) + (t * MT + m) * KS + ks);
#pragma unroll
              for (int t = 0; t < 2; ++t)
                if (j + t - 2 >= 0) hd[t].load(HAp(l, i * W + j + t - 2));
            }
#pragma unroll UNR
            for (int t = 0; t < 3; ++t) {
              const int c = j + t - 2;
              if (c < 0) continue;
              frag b[KS];
              if (t == 2) {
                d_to_b<PREC, MT, KS>(hcur, b);
              } else if constexpr (BATCH) {
                hd[t].template frags<KS>(b);
              } else {
                hd[0].load(HAp(l, i * W + c));
                hd[0].template frags<KS>(b);
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
          LD cur2;
          cur2.load_quads(CATp(l, q), MTH, MT);
          frag wc[BT][MT][KS];
          LD cd[BATCH ? 8 : 1];
          if constexpr (BATCH) {
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
              for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) wc[t][m][ks] = wfrag(lay.off_c(l) + (t * MT + m) * KS + ks);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
              const int r = i + t / 3 - 2, c = j + t % 3 - 2;
              if (r >= 0 && c >= 0) cd[t].load(CATp(l, r * W + c));
            }
          }
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            f32x4 x = load_bias4(bl + 2 * C, 16 * m + 4 * g, C / 2);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) x = P::mfma(wfrag(lay.off_x(l) + m * KS + ks), bxf[ks], x);
            cur2.set(m, act4<ACT>(x));
          }
          cur2.store_quads(CATp(l, q), 0, MTH);
          f32x4 cacc[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) cacc[m] = load_bias4(bl + 3 * C, 16 * m + 4 * g, C);
#pragma unroll UNR
          for (int t = 0; t < 9; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 2;
            if (r < 0 || c < 0) continue;
            frag b[KS];
            if (t == 8) {
              cur2.template frags<KS>(b);
            } else if constexpr (BATCH) {
              cd[t].template frags<KS>(b);
            } else {
              cd[0].load(CATp(l, r * W + c));
              cd[0].template frags<KS>(b);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
              for (int ks = 0; ks < KS; ++ks) {
                frag wk;
                if constexpr (BATCH) wk = wc[t][m][ks];
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
          hn.store(HAp(l + 1, q));
          // the value the next block sees is the stored (rounded) one
#pragma unroll
          for (int m = 0; m < MT; ++m) hcur[m] = hn.get(m);
        }
      }
    }

    // ============ vertical stack of row i (all blocks) ======================
    for (int l = 0; l <= F; ++l) {
      const int kind = block_kind(l, nb);
      const float* bl = a.aux + lay.aux_bias(l);
      // this block's vertical and y weights serve the whole row
      frag wv[BT][MT][KS];
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
      for (int j = 0; j < W; ++j) {
        const int q = i * W + j;
        f32x4 acc[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[m] = load_bias4(bl, 16 * m + 4 * g, C);
        if (l == 0) {
          const float* kv = a.aux + lay.aux_kv0(nb);
          for (int t = 0; t < 9; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 1;
            const bool valid = r >= 0 && c >= 0 && c < W;
            const float sv = valid ? spins[(r * W + c) * 16 + n] : 0.0f;
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m] += *reinterpret_cast<const f32x4*>(kv + t * C + 16 * m + 4 * g) * sv;
          }
        } else {
          LD vd[BT];
          if constexpr (BATCH) {
#pragma unroll
            for (int t = 0; t < 9; ++t) {
              const int r = i + t / 3 - 2, c = j + t % 3 - 1;
              if (r >= 0 && c >= 0 && c < W) vd[t].load(VAp(l, r * W + c));
            }
          }
#pragma unroll UNR
          for (int t = 0; t < 9; ++t) {
            const int r = i + t / 3 - 2, c = j + t % 3 - 1;
            if (r < 0 || c < 0 || c >= W) continue;
            frag b[KS];
            if constexpr (BATCH) {
              vd[t].template frags<KS>(b);
            } else {
              vd[0].load(VAp(l, r * W + c));
              vd[0].template frags<KS>(b);
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
            vr.load(VAp(l - 1, q));
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
          const float* by = bl + 2 * C + C / 2;
          LD yd;
#pragma unroll
          for (int m = 0; m < MTH; ++m) {
            f32x4 y = load_bias4(by, 16 * m + 4 * g, C / 2);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) y = P::mfma(wfrag(lay.off_y(l) + m * KS + ks), b[ks], y);
            yd.set(MTH + m, act4<ACT>(y));
          }
          yd.store_quads(CATp(l, q + W), MTH, MT);
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

int64_t fast_tile_bytes(int HW, int nb, int channels, int precision) {
  const int esz = precision == FK_PREC_FP32 ? 4 : 2;
  const int64_t pb = 16LL * channels * esz;
  return (int64_t)(3 * nb - 2) * HW * pb + (int64_t)HW * 16 * 4;
}

template <int PREC, int C>
static int launch_fast(const SampArgs& a, int act, int64_t tiles, hipStream_t s) {
  if (act == FK_ACT_ELU)
    hipLaunchKernelGGL((fast_sample_kernel<PREC, C, FK_ACT_ELU>), dim3((unsigned)tiles), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((fast_sample_kernel<PREC, C, FK_ACT_RELU>), dim3((unsigned)tiles), dim3(64), 0, s, a);
  FK_CHECK_LAUNCH();
  return FK_OK;
}

}  // namespace fk

using namespace fk;

extern "C" {

int64_t fk_fast_sampler_workspace_bytes(int32_t H, int32_t W, int32_t depth, int32_t channels, int32_t precision,
                                        int64_t batch) {
  if (H < 1 || W < 1 || depth < 1 || batch < 0 || !fast_sampler_supported(channels, precision)) return -1;
  const int64_t tiles = (batch + 15) / 16;
  return tiles * fast_tile_bytes(H * W, num_causal_blocks(depth), fast_channels(channels, precision), precision);
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
  a.tile_bytes = fast_tile_bytes(H * W, a.nb, CP, precision);
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
