// lgpu_launch.h -- host-side launchers of the step kernels, one explicit
// instantiation per precision policy (lgpu_step_fp64.hip, lgpu_step_mixed.hip,
// lgpu_step_fp64fma.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "lgpu_kernels.h"

namespace lgpu {

constexpr int kEosBlock = 256;
// Slot pairs of the one-problem dt minima (k_eos atomics, k_dt_slots); a
// power of two
constexpr int kDtSlots = 256;

inline unsigned eos_blocks(int numElem)
{
   return (unsigned)((numElem + kEosBlock - 1) / kEosBlock);
}

// k_eos dispatch slots for the perm blocks in `order`: each block's repeat
// counts from the work-list groups (regStart: nGroups+1 perm positions,
// possibly empty groups; regRep: nGroups counts) -- its first position's
// group and, where the block reaches into it, the next non-empty group.  A
// block over three or more groups gets rep0 = -1 (k_eos searches regStart).
inline std::vector<EosSlot> eos_slots(const std::vector<int32_t> &order,
                                      const std::vector<int32_t> &regStart,
                                      const std::vector<int32_t> &regRep, int64_t numPositions)
{
   const int ng = (int)regRep.size();
   auto group_of = [&](int64_t p) {
      const int g = (int)(std::upper_bound(regStart.begin(), regStart.begin() + ng, (int32_t)p) -
                          regStart.begin()) - 1;
      // the last group starting at or before p: of several groups that start
      // at p, the last, which is the one holding p where any does (empty
      // groups only precede it, or cover padding after a member's elements)
      return std::max(0, g);
   };
   std::vector<EosSlot> slots(order.size());
   for (size_t i = 0; i < order.size(); ++i) {
      const int64_t p0 = (int64_t)order[i] * kEosBlock;
      const int64_t p1 = std::max(p0, std::min<int64_t>(p0 + kEosBlock, numPositions) - 1);
      EosSlot &sl = slots[i];
      sl.block = order[i];
      const int g0 = group_of(p0), g1 = group_of(p1);
      if (g0 == g1) {
         sl.rep0 = sl.rep1 = regRep[(size_t)g0];
         sl.split = (int32_t)(p0 + kEosBlock);
      } else if (group_of(regStart[(size_t)g0 + 1]) == g1) {
         sl.rep0 = regRep[(size_t)g0];
         sl.rep1 = regRep[(size_t)g1];
         sl.split = regStart[(size_t)g0 + 1];
      } else {
         sl.rep0 = -1;
         sl.split = sl.rep1 = 0;
      }
   }
   return slots;
}

template <class P>
struct Launch {
   static void force(const DevState &s, const Consts &k, int64_t planeStride, int block,
                     hipStream_t st);
   static void node(const DevState &s, const Consts &k, int64_t planeStride, double dt, int block,
                    hipStream_t st);
   static void nodal_brick(const DevState &s, const Consts &k, double dt, hipStream_t st);
   // resident k_nodal_brick workgroups per CU (0 if the query fails)
   static int nodal_brick_blocks_per_cu();
   static void kin(const DevState &s, const Consts &k, double dt, int block, hipStream_t st);
   static void kin_brick(const DevState &s, const Consts &k, double dt, hipStream_t st);
   // resident k_kin_brick workgroups per CU (0 if the query fails)
   static int kin_brick_blocks_per_cu();
   // ensemble on the brick path: one launch over every member's workgroups
   // (16x16-thread tiles), the grid sizes from the member tables
   static void nodal_brick_members(const DevState &s, const Consts &k, unsigned grid,
                                   hipStream_t st);
   static void kin_brick_members(const DevState &s, const Consts &k, unsigned grid, hipStream_t st);
   static void eos(const DevState &s, const Consts &k, hipStream_t st);
   static void dt_reduce(const DevState &s, hipStream_t st);
};

extern template struct Launch<PolicyFp64>;
extern template struct Launch<PolicyMixed>;
extern template struct Launch<PolicyFp64Fma>;

} // namespace lgpu
