// lgpu_time.h -- device-side time state and TimeIncrement (lulesh.cc:166-222)
// of the device loop (lgpu_run).
#pragma once

#include <hip/hip_runtime.h>

#include "lgpu_kernels.h"

namespace lgpu {

// TimeIncrement operands plus the loop flag the kernels test.
struct DevTime {
   double time, deltatime, dtcourant, dthydro, stoptime, dtfixed, dtmax, multlb, multub;
   int32_t cycle, maxCycle, active, pad;
};

// TimeIncrement (lulesh.cc:166-222) at the top of a cycle, given the dt
// constraints of the cycle before (used when cycle > 0); also decides whether
// the cycle runs at all (the loop test of lulesh.cc:2745, and an error flag
// of the cycle before ends the loop).  One thread.
__device__ inline void time_increment(DevTime *t, double dtc, double dth, const int32_t *err)
{
   if (!t->active) return;
   bool failed = false;
   for (int i = 0; i < kNumErrSites; ++i) failed = failed || (err[i] != 0);
   if (t->cycle > 0) {
      t->dtcourant = dtc;
      t->dthydro = dth;
   }
   if (failed || !(t->time < t->stoptime && t->cycle < t->maxCycle)) {
      t->active = 0;
      return;
   }
   double targetdt = t->stoptime - t->time;
   if (t->dtfixed <= 0.0 && t->cycle != 0) {
      const double olddt = t->deltatime;
      double newdt = 1.0e+20;
      if (t->dtcourant < newdt) newdt = t->dtcourant / 2.0;
      if (t->dthydro < newdt) newdt = t->dthydro * 2.0 / 3.0;
      const double ratio = newdt / olddt;
      if (ratio >= 1.0) {
         if (ratio < t->multlb) newdt = olddt;
         else if (ratio > t->multub) newdt = olddt * t->multub;
      }
      if (newdt > t->dtmax) newdt = t->dtmax;
      t->deltatime = newdt;
   }
   if (targetdt > t->deltatime && targetdt < 4.0 * t->deltatime / 3.0)
      targetdt = 2.0 * t->deltatime / 3.0;
   if (targetdt < t->deltatime) t->deltatime = targetdt;
   t->time += t->deltatime;
   t->cycle += 1;
}

} // namespace lgpu
