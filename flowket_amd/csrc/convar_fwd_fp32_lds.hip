// ConvNetAutoregressive2D fused forward, fp32 kernels with the five activation
// lattices resident in LDS (kernel template: convar_fwd_kernel.h; host side:
// convar_fwd.hip).  One file per precision and lattice placement, so the
// kernel variants compile in parallel.
#include "convar_fwd_kernel.h"

namespace fk {
FK_DEFINE_DISPATCH_FWD(FK_PREC_FP32, false)
}  // namespace fk
