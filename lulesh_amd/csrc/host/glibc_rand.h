// glibc_rand.h -- the pseudo-random sequence of glibc's srand()/rand().
//
// The reference draws its region layout from srand(0)/rand()
// (lulesh-init.cc:406-480).  This generator reproduces glibc's default
// (TYPE_3) generator from its published description -- an additive lagged
// Fibonacci generator over a 31-word state, r[i] = r[i-3] + r[i-31]
// (mod 2^32), seeded by a Lehmer LCG (multiplier 16807, modulus 2^31 - 1),
// with the first 310 outputs discarded and each output shifted right by one
// -- so the layout does not depend on the process-wide rand() state.
#pragma once

#include <stdint.h>

namespace lhost {

class GlibcRand {
 public:
   explicit GlibcRand(uint32_t seed);
   int32_t next();

 private:
   uint32_t state_[31];
   int front_ = 3; // r[i-3] side, updated in place
   int rear_ = 0;  // r[i-31] side
};

} // namespace lhost
