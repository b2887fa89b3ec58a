// ConvNetAutoregressive2D forward for lattices whose five activation
// lattices exceed the LDS (e.g. 20x20 with 64 bf16 channels: 256 KB per
// configuration): the vertical stacks and the concatenation stay in LDS, the
// horizontal-stack lattices H0/H1 (read by the 1x3 convolution only, 3 taps)
// live in a per-workgroup global scratch that stays L2-resident.  Same kernel
// template and arithmetic as convar_fwd.hip (tap-outer, 8 or 4 waves: the
// plan's wave count).
#include "convar_fwd_kernel.h"

namespace fk {

template <int PREC, int C>
static int launch_hg(const FwdArgs& a, const FwdPlan& pl, int act, hipStream_t s) {
  if (pl.nwave == 4) return launch_variant<PREC, C, 4, 0, true>(a, pl, act, s);
  return launch_variant<PREC, C, 8, 0, true>(a, pl, act, s);
}

int dispatch_fwd_hg(const FwdArgs& a, const FwdPlan& pl, int prec, int cp, int act, hipStream_t s) {
  if (prec == FK_PREC_BF16) return cp == 32 ? launch_hg<FK_PREC_BF16, 32>(a, pl, act, s)
                                            : launch_hg<FK_PREC_BF16, 64>(a, pl, act, s);
  if (prec == FK_PREC_FP16) return cp == 32 ? launch_hg<FK_PREC_FP16, 32>(a, pl, act, s)
                                            : launch_hg<FK_PREC_FP16, 64>(a, pl, act, s);
  if (cp == 16) return launch_hg<FK_PREC_FP32, 16>(a, pl, act, s);
  if (cp == 32) return launch_hg<FK_PREC_FP32, 32>(a, pl, act, s);
  return launch_hg<FK_PREC_FP32, 64>(a, pl, act, s);
}

}  // namespace fk
