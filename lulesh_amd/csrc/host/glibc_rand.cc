#include "host/glibc_rand.h"

namespace lhost {

GlibcRand::GlibcRand(uint32_t seed)
{
   // glibc replaces a zero seed by one.
   int32_t word = (int32_t)(seed == 0 ? 1 : seed);
   state_[0] = (uint32_t)word;
   for (int i = 1; i < 31; ++i) {
      // 16807 * word mod (2^31 - 1) without overflow (Schrage's method).
      const int32_t hi = word / 127773, lo = word % 127773;
      word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      state_[i] = (uint32_t)word;
   }
   front_ = 3;
   rear_ = 0;
   for (int i = 0; i < 310; ++i) next();
}

int32_t GlibcRand::next()
{
   state_[front_] += state_[rear_];
   const int32_t out = (int32_t)(state_[front_] >> 1);
   front_ = (front_ + 1 == 31) ? 0 : front_ + 1;
   rear_ = (rear_ + 1 == 31) ? 0 : rear_ + 1;
   return out;
}

} // namespace lhost
