// ConvNetAutoregressive2D fused forward, fp16 kernels for lattices whose five
// activation lattices exceed the LDS (e.g. 20x20 with 64 16-bit channels:
// 256 KB per configuration): the vertical stacks and the concatenation stay
// in LDS, the horizontal-stack lattices H0/H1 (read by the 1x3 convolution
// only) live in a per-workgroup global scratch that stays L2-resident.  Same
// kernel template and arithmetic as the LDS-resident variants (taps outer, the
// plan's wave count).
#include "convar_fwd_kernel.h"

namespace fk {
FK_DEFINE_DISPATCH_FWD(FK_PREC_FP16, true)
}  // namespace fk
