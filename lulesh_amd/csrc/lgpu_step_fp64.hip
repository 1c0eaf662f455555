// FP64 instantiation of the step kernels.  Built with -ffp-contract=off so
// that every operation rounds as the reference's (FMA-free) x86-64 build.
#include "lgpu_step.inc"

namespace lgpu {
template struct Launch<PolicyFp64>;
}
