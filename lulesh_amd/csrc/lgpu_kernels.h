// lgpu_kernels.h -- device kernels of the Lagrange leapfrog step.
//
// Four passes per cycle, each a single launch over its whole index space:
//
//   k_force  (elements)  stress integration + Flanagan-Belytschko hourglass,
//                        per-corner forces to corner planes
//   k_node   (nodes)     deterministic corner gather in nodeElemCornerList
//                        order, acceleration, symmetry BCs, velocity, position
//   k_kin    (elements)  volume, vnew, characteristic length, strain rate,
//                        monotonic-Q gradients
//   k_eos    (region-ordered elements)  qstop check, monotonic-Q limiter,
//                        delv = vnew - v, EOS with its region repeat count,
//                        sound speed, volume update, Courant/hydro block minima
//   k_dt_slots / k_dt_reduce (one block per problem) Courant/hydro minima
//                        over the k_eos slots (one problem: atomic minima
//                        into 256 slot pairs) or block partials (ensembles)
//
// The Policy template parameter selects the arithmetic of each part:
// PolicyFp64 keeps everything in FP64 with the reference's association order,
// PolicyMixed runs force integration (corner forces and their node sums),
// Q gradients/limiter and EOS pressure and sound speed in FP32 on
// element-local coordinates; node accelerations, velocities and positions,
// element volume, delv, strain rate, characteristic length, the energy
// update and the dt candidates stay FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>

#include "lgpu_geom.h"

namespace lgpu {

// Error sites, in the order the reference checks them within one cycle.
enum ErrSite {
   kErrForceDeterm = 0, // lulesh.cc:1082-1091  determ <= 0       -> VolumeError
   kErrHourglassV = 1,  // lulesh.cc:1033-1040  v <= 0            -> VolumeError
   kErrKinVnew = 2,     // lulesh.cc:1597-1605  vnew <= 0         -> VolumeError
   kErrQStop = 3,       // lulesh.cc:1993-2008  q > qstop         -> QStopError
   kErrMaterialV = 4,   // lulesh.cc:2366-2384  clamped v <= 0    -> VolumeError
   kNumErrSites = 8
};

// One k_eos workgroup's perm block and its repeat counts (a block spans at
// most two work-list groups unless a group holds fewer than a block).
struct EosSlot {
   int32_t block, rep0, split, rep1;
};

struct Consts {
   double e_cut, p_cut, q_cut, v_cut, u_cut;
   double hgcoef, ss4o3, qstop, monoq_max_slope, monoq_limiter_mult;
   double qlc_monoq, qqc_monoq, qqc, eosvmax, eosvmin, pmin, emin, dvovmax, refdens;
   int32_t fastEos; // fast_eos_consts(*this), set by the host
};

// The constants k_eos's fast FP64 EOS repeats assume: e_cut and p_cut
// positive (the values then clamped below by emin and pmin are never -0),
// emin and pmin not -0 and of magnitude <= 2^100, refdens in [2^-64, 2^64].
inline bool fast_eos_consts(const Consts &k)
{
   const double big = 0x1p100;
   return k.e_cut > 0. && k.p_cut > 0. && std::fabs(k.emin) <= big && std::fabs(k.pmin) <= big &&
          !(k.emin == 0. && std::signbit(k.emin)) && !(k.pmin == 0. && std::signbit(k.pmin)) &&
          k.refdens >= 0x1p-64 && k.refdens <= 0x1p64;
}

struct PolicyFp64 {
   using force_t = double;  // force integration arithmetic
   using corner_t = double; // stored per-corner force
   using grad_t = double;   // Q gradients + limiter arithmetic and storage
   using eos_t = double;    // EOS pressure / sound-speed arithmetic
   using nodesum_t = double; // per-node sums of the corner forces
   using pqs_t = double;     // stored p, q, ss
   static constexpr bool kLocalCoords = false;
   // geometry in the compact regrouped forms (different roundings) instead
   // of the reference's association order
   static constexpr bool kCompact = false;
   // waves per SIMD (min, max) for the nodal brick kernel's register
   // allocation and scheduling; (1, 8) leaves it to the compiler
   static constexpr int kNodalWavesMin = 1, kNodalWavesMax = 8;
   // brick kernel: load the node update's operands before the element work
   static constexpr bool kPrefetchNode = true;
   // brick kernel: 32-bit byte offsets from scalar bases (at32) instead of
   // 64-bit per-lane addresses; FP64 runs 1.4% slower with them (more VGPRs)
   static constexpr bool kOffset32 = false;
   // brick kernel: issue the first hourglass velocity gather before the
   // hourglass modes (true) or after them; the later components are always
   // loaded one component ahead.  FP64 s=192: after, 598 us against 748 (256
   // VGPRs before); FMA and mixed: before, 1.4% and 0.5% faster
   static constexpr bool kHgVelEarly = false;
   // brick kernel: a scheduling barrier after the element-field loads (p, q,
   // v, volo, ss, elemMass), so that none is sunk to its use.  Mixed s=192:
   // 416 -> 411 us; FP64 582 -> 600 us, FMA 505 -> 534 us (the stress work
   // then waits for every load at once)
   static constexpr bool kPinElemFields = false;
   // brick kernel: load the next element plane's coordinates and fields in
   // the node phase of the current one (between its barriers).  FP64_FMA
   // s=192: 498 -> 486 us; FP64 no gain (255 VGPRs)
   static constexpr bool kPrefetchElem = false;
   // k_eos waves per SIMD: FP64 at 7 (at most 72 VGPRs); 8 (64 VGPRs, four
   // spilled) ran 326 -> 332 us at s=192
   static constexpr int kEosWaves = 7;
   // brick kernel: the node phase at raised wave priority (s_setprio 2).
   // s=192, every kernel timed: FP64 576.4 -> 568.9 us, FP64_FMA 489 ->
   // 476.5 us (device loop: FP64 unchanged, FP64_FMA 484 -> 472 us); mixed
   // 417.5 -> 420.4 us
   static constexpr bool kNodePhasePrio = true;
};

// The FP64 policy compiled with multiply-add contraction (its own
// translation unit); the types are those of PolicyFp64.
struct PolicyFp64Fma : PolicyFp64 {
   static constexpr bool kHgVelEarly = true;
   static constexpr bool kPrefetchElem = true;
};

struct PolicyMixed {
   using force_t = float;
   using corner_t = float;
   using grad_t = float;
   using eos_t = float;
   using nodesum_t = float;
   using pqs_t = float;
   static constexpr bool kLocalCoords = true;
   static constexpr bool kCompact = true;
   // exactly 3: aiming for 4 (128 VGPRs) makes the scheduler issue the
   // hourglass velocity loads one at a time, each waited on at once
   static constexpr int kNodalWavesMin = 3, kNodalWavesMax = 3;
   static constexpr bool kPrefetchNode = true;
   static constexpr bool kOffset32 = true;
   static constexpr bool kHgVelEarly = true;
   static constexpr bool kPinElemFields = true;
   static constexpr bool kPrefetchElem = false; // 56 spills at 3 waves/SIMD
   // k_eos at 8 waves per SIMD (at most 64 VGPRs): 214 -> 211 us at s=192
   static constexpr int kEosWaves = 8;
   static constexpr bool kNodePhasePrio = false;
};

// Element i of a device array through a 32-bit byte offset: the access
// compiles to the scalar-base + 32-bit vector-offset form of the global
// load/store instead of per-lane 64-bit address arithmetic.  The caller
// guarantees i * sizeof(T) < 2^32 (the brick kernels: lgpu_create admits at
// most 2^28 elements, so a brick has fewer than 2^28.1 nodes).
template <typename T>
__device__ __forceinline__ T &at32(T *base, uint32_t i)
{
   using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
   return *reinterpret_cast<T *>(reinterpret_cast<B *>(base) + i * (uint32_t)sizeof(T));
}

// Division by a run-time constant d (< 2^32) as a 64-bit multiply-high:
// q = hi64(n * m) with m = floor((2^64 - 1) / d) + 1 is exact for n < 2^32.
struct FastDiv {
   uint64_t m;
   uint32_t d;
};

// Structured-brick description: s elements per edge, lattice node/element
// numbering of lulesh-init.cc:217-267.  Set when the mesh is the reference's
// brick, which lets the nodal pass run as one fused kernel.
struct Brick {
   int32_t s, en;          // elements and nodes per edge
   int32_t tilesX, tilesY; // k_nodal_brick column tiles of (tileW-1) x (rows-1) nodes
   int32_t chunkZ, chunks; // node planes per z-chunk, number of chunks
   int32_t rows;           // thread rows per k_nodal_brick workgroup: 16 (256 threads) or 8
                           // for one problem; up to 256 / tileW in ensembles
   int32_t tileW;          // threads per k_nodal_brick row: 16 for one problem; chosen per
                           // ensemble member (4..64)
   FastDiv divS, divS2;    // element index -> (a, b, c)
   // k_kin_brick: 16x16-element column tiles per edge, element planes per z-chunk
   int32_t kinTiles, kinChunk, kinChunks;
   // ensemble members: kinW x kinH element tiles, kinTilesX x kinTilesY per plane
   int32_t kinW, kinH, kinTilesX, kinTilesY;
};

// One ensemble member on the brick path: its brick, where its nodes and
// elements start in the concatenated arrays, and its first workgroup in the
// k_nodal_brick and k_kin_brick grids.
struct MemberBrick {
   Brick g;
   int32_t nodeOff, elemOff, nodalBlock0, kinBlock0;
};

// Everything a cycle touches on the device.
struct DevState {
   int32_t numElem, numNode, numReg;
   int32_t xcdRemap; // 1: map consecutive blocks onto one XCD's L2
   // node fields (current buffer; the brick path writes the other one)
   double *x, *y, *z, *xd, *yd, *zd, *nodalMass;
   double *xo, *yo, *zo, *xdo, *ydo, *zdo; // second node buffer (brick path)
   double *fx, *fy, *fz, *xdd, *ydd, *zdd; // diagnostics (may be null)
   const uint8_t *nodeBC;                  // bit0 symmX, bit1 symmY, bit2 symmZ
   const int32_t *nodeElemStart, *cornerList;
   // element fields
   const int32_t *nodelist;
   const int32_t *lxim, *lxip, *letam, *letap, *lzetam, *lzetap, *elemBC;
   double *e, *v, *volo;
   // P::pqs_t (FP64, or FP32 in the mixed form): p, q, ss
   void *p, *q, *ss;
   double *elemMass;
   double *vnew, *delv;
   double *arealg, *vdov;
   double *ql, *qq; // diagnostics (may be null)
   double *delvOut; // diagnostics: delv = vnew - v, stored by k_eos (may be null)
   void *delx_xi, *delx_eta, *delx_zeta, *delv_xi, *delv_eta, *delv_zeta; // grad_t
   // corner planes: [(component*2 + kind)*8 + corner][elem], kind 0 stress, 1 hourglass
   void *corner;
   // region order
   int32_t numPerm;         // entries of perm (>= numElem; -1 marks padding)
   const int32_t *perm;     // region-major element order
   const int32_t *regStart; // numReg+1 offsets into perm
   const int32_t *regRep;   // EOS repeat count per region
   // control
   int32_t *err;            // kNumErrSites flags
   double *dtPartial;       // one problem: kDtSlots pairs of min_key values; ensemble: 2 per
                            // k_eos block
   double *dtOut;           // [0] dtcourant, [1] dthydro
   const double *dtDevice;  // device-side deltatime (null: use argument)
   const int32_t *active;   // device-side loop flag (null: always active)
   // device-side time state: k_dt_slots also runs the next cycle's
   // TimeIncrement (null: host-driven time stepping)
   struct DevTime *time;
   Brick brick;             // valid when isBrick
   int32_t isBrick;
   int32_t brickFaces; // k_eos computes face neighbours and BC masks from the brick
   // k_eos dispatch table (one problem): workgroup i processes perm block
   // eosSlot[i].block, whose positions below eosSlot[i].split take repeat
   // count rep0 and the others rep1 (rep0 < 0: look each position up in
   // regStart).  Null (ensembles): block i, counts looked up.
   const struct EosSlot *eosSlot;
   // Ensemble of independent problems (nMembers > 0): members are
   // concatenated; each has its own dt, loop flag and error flags.
   int32_t nMembers;
   const int32_t *memberNodeStart; // nMembers+1
   const int32_t *memberElemStart; // nMembers+1
   const double *memberDt;         // per-member deltatime
   const int32_t *memberActive;    // per-member loop flag
   int32_t *memberErr;             // nMembers x kNumErrSites
   const int32_t *blockMember;     // k_eos block -> member (-1: padding)
   const int32_t *memberBlockStart; // nMembers+1: k_eos blocks of each member
   double *memberDtOut;            // 2 per member: dtcourant, dthydro
   // ensemble on the brick path (memberBrick non-null)
   const MemberBrick *memberBrick;
   const int32_t *nodalBlockMember; // k_nodal_brick workgroup -> member
   const int32_t *kinBlockMember;   // k_kin_brick workgroup -> member
};

} // namespace lgpu
