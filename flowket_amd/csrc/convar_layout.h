// Packed-parameter layout of ConvNetAutoregressive2D shared by the host map
// builder and the device kernels.
//
// Flat parameter vector (canonical, Keras order of
// machines/conv_net_autoregressive_2D.py:33-51,73): per causal block b
//   vconv K (3,3,cin,C) b(C) | hconv K (1,3,cin,C) b(C) | x1x1 K (1,1,C,C/2) b(C/2)
//   | y1x1 K (1,1,C,C/2) b(C/2) | h3conv K (3,3,C,C) b(C)
// with cin = 1 for b = 0 and C otherwise; then head K (1,1,C,4) b(4).
//
// Packed weight buffer: MFMA A-operand fragments (64 lanes x E elements each,
// E = 8 bf16 or 1 fp32), the output channel (m = lane % 16 of M-tile mt) as
// the MFMA row and the input-channel *slot* as the contraction index.
//   block 0      : [x1x1][y1x1][h3conv]
//   block b >= 1 : [vconv][hconv][x1x1][y1x1][h3conv]
//   head         : [head]
// Within a convolution fragments are ordered (tap, mt, kstep).
// Aux fp32 buffer: per block b at 4*C*b: bv(C) bh(C) bx(C/2) by(C/2) bc(C);
// head bias (4) at 4*C*nb; block-0 vertical kernel (9*C, [tap][c]) and
// horizontal kernel (3*C) after it.
//
// Channel padding: a network of C channels runs on the kernels compiled for
// CP = kernel_channels(C) >= C (bf16/fp16: 32, 64; fp32: 16, 32, 64).  The layout
// above is then taken with CP; channels >= C of every convolution and the
// x/y halves past C/2 get zero weights and biases, so their activations stay
// exactly zero and the result is the C-channel network's.  In the padded
// concatenation [x | y] the y half starts at channel CP/2.
#pragma once
#include <cstdint>

#ifndef FK_HD
#define FK_HD __host__ __device__
#endif

namespace fk {

struct ConvarLayout {
  int C, prec, E, KW, MT, KS, MTH;
  int64_t nV, nH, nX, nCc, nHead, blk0, blkN;

  FK_HD ConvarLayout(int channels, int precision) {
    C = channels;
    prec = precision;
    E = (prec != 0) ? 8 : 1;       // elements per lane per fragment
    KW = (prec != 0) ? 32 : 4;     // input-channel slots per MFMA k-step
    MT = C / 16;                   // 16-row output tiles of a C-output conv
    KS = C / KW;                   // k-steps per tap
    MTH = (C / 2 + 15) / 16;       // output tiles of a C/2-output 1x1 conv
    nV = 9LL * MT * KS;
    nH = 3LL * MT * KS;
    nX = (int64_t)MTH * KS;
    nCc = 9LL * MT * KS;
    nHead = KS;
    blk0 = 2 * nX + nCc;
    blkN = nV + nH + 2 * nX + nCc;
  }
  // fragment index of the start of block b
  FK_HD int64_t block_base(int b) const { return b == 0 ? 0 : blk0 + (int64_t)(b - 1) * blkN; }
  FK_HD int64_t head_base(int nb) const { return block_base(nb); }
  FK_HD int64_t total_frags(int nb) const { return head_base(nb) + nHead; }
  // offsets of the parts inside a block (fragments)
  FK_HD int64_t off_v(int b) const { return block_base(b); }
  FK_HD int64_t off_h(int b) const { return block_base(b) + nV; }
  FK_HD int64_t off_x(int b) const { return b == 0 ? 0 : block_base(b) + nV + nH; }
  FK_HD int64_t off_y(int b) const { return off_x(b) + nX; }
  FK_HD int64_t off_c(int b) const { return off_y(b) + nX; }
  // aux buffer
  FK_HD int64_t aux_bias(int b) const { return 4LL * C * b; }
  FK_HD int64_t aux_head_bias(int nb) const { return 4LL * C * nb; }
  FK_HD int64_t aux_kv0(int nb) const { return 4LL * C * nb + 4; }
  FK_HD int64_t aux_kh0(int nb) const { return aux_kv0(nb) + 9LL * C; }
  FK_HD int64_t aux_total(int nb) const { return aux_kh0(nb) + 3LL * C; }
};

// Channel stored in input slot s (the MFMA contraction index).  bf16: within
// each 32-slot k-step, slot 8*g+e holds channel 4*g+e (e<4) or 16+4*g+(e-4)
// (e>=4): slot bit 2 moves to channel bit 4, slot bits 3-4 to channel bits
// 2-3.  fp32: k-step t, lane group g holds channel 16*(t/4) + 4*g + t%4.
FK_HD inline int slot_channel_bf16(int s) {
  return (s & ~31) | (((s >> 2) & 1) << 4) | (((s >> 3) & 3) << 2) | (s & 3);
}
FK_HD inline int slot_channel_f32(int t, int g) { return 16 * (t >> 2) + 4 * g + (t & 3); }

// Kernel channel width for a C-channel network, 0 if none fits.
inline int kernel_channels(int channels, int precision) {
  if (channels < 2 || (channels & 1)) return 0;
  if (precision < 0 || precision > 2) return 0;
  const int widths[3] = {16, 32, 64};
  for (int i = (precision != 0) ? 1 : 0; i < 3; ++i)
    if (channels <= widths[i]) return widths[i];
  return 0;
}

inline int num_causal_blocks(int depth) { return 2 + 2 * (depth > 2 ? depth - 2 : 0); }

}  // namespace fk
