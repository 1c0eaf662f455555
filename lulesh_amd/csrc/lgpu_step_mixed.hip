// Mixed-precision instantiation of the step kernels (FP32 force integration,
// Q gradients/limiter, EOS pressure and sound speed).  FMA contraction is
// allowed here.
#include "lgpu_step.inc"

namespace lgpu {
template struct Launch<PolicyMixed>;
}
