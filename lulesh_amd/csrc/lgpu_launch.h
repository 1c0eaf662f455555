// lgpu_launch.h -- host-side launchers of the step kernels, one explicit
// instantiation per precision policy (lgpu_step_fp64.hip, lgpu_step_mixed.hip,
// lgpu_step_fp64fma.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
