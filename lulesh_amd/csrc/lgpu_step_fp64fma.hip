// FP64 instantiation of the step kernels built with -ffp-contract=fast: the
// reference's association order, with each a*b+c fused into one FMA.
#include "lgpu_step.inc"

namespace lgpu {
template struct Launch<PolicyFp64Fma>;
}
