// lgpu_ensemble.hip -- a batch of independent Sedov problems advanced
// together: one launch per kernel covers every member, each member keeps its
// own dt (TimeIncrement, lulesh.cc:166-222, evaluated per member on the
// device), and one cycle is captured once as a hipGraph and replayed.
//
// Members are concatenated in one device arena: node arrays, element arrays
// and node-to-corner lists are offset per member, and the region-ordered EOS
// list is padded per member to whole k_eos blocks so that every block (and
// so every dt partial) belongs to exactly one member.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "lgpu_brick.h"
#include "lgpu_kernels.h"
#include "lgpu_launch.h"
#include "lulesh_gpu.h"

using namespace lgpu;

namespace {

enum EnsKernel {
   kEForce = 0, kENode, kEKin, kEEos, kEDt, kETime, kENodalBrick, kEKinBrick, kENumKernels
};
const char *const kEnsKernelNames[kENumKernels] = {"k_force", "k_node", "k_kin", "k_eos",
                                                   "k_dt_reduce", "k_time_increment",
                                                   "k_nodal_brick", "k_kin_brick"};

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Per-member time state, struct of arrays (the kernels read deltatime and
// active as plain arrays).
struct EnsTime {
   double *time, *deltatime, *dtcourant, *dthydro, *stoptime, *dtfixed, *dtmax, *multlb, *multub;
   int32_t *cycle, *active;
   // the run's cycle limit, one value in device memory: set by each
   // lgpu_ensemble_run, so that the cycle graphs do not depend on it
   int32_t *maxCycle;
};

// TimeIncrement per member (lulesh.cc:166-222) and the loop test of
// lulesh.cc:2745; a member whose step failed stops.
__global__ void k_time_increment_ens(EnsTime t, const double *dtOut, const int32_t *err,
                                     int nMembers)
{
   const int m = blockIdx.x * blockDim.x + threadIdx.x;
   if (m >= nMembers || !t.active[m]) return;
   const int maxCycle = *t.maxCycle;
   bool failed = false;
   for (int i = 0; i < kNumErrSites; ++i) failed = failed || (err[m * kNumErrSites + i] != 0);
   if (t.cycle[m] > 0) {
      t.dtcourant[m] = dtOut[2 * m];
      t.dthydro[m] = dtOut[2 * m + 1];
   }
   if (failed || !(t.time[m] < t.stoptime[m] && t.cycle[m] < maxCycle)) {
      t.active[m] = 0;
      return;
   }
   double targetdt = t.stoptime[m] - t.time[m];
   double dt = t.deltatime[m];
   if (t.dtfixed[m] <= 0.0 && t.cycle[m] != 0) {
      const double olddt = dt;
      double newdt = 1.0e+20;
      if (t.dtcourant[m] < newdt) newdt = t.dtcourant[m] / 2.0;
      if (t.dthydro[m] < newdt) newdt = t.dthydro[m] * 2.0 / 3.0;
      const double ratio = newdt / olddt;
      if (ratio >= 1.0) {
         if (ratio < t.multlb[m]) newdt = olddt;
         else if (ratio > t.multub[m]) newdt = olddt * t.multub[m];
      }
      if (newdt > t.dtmax[m]) newdt = t.dtmax[m];
      dt = newdt;
   }
   if (targetdt > dt && targetdt < 4.0 * dt / 3.0) targetdt = 2.0 * dt / 3.0;
   if (targetdt < dt) dt = targetdt;
   t.deltatime[m] = dt;
   t.time[m] += dt;
   t.cycle[m] += 1;
}

int32_t region_rep(int32_t r, int32_t numReg, int32_t cost)
{
   if (r < numReg / 2) return 1;
   if (r < numReg - (numReg + 15) / 20) return 1 + cost;
   return 10 * (1 + cost);
}

} // namespace

struct lgpu_ensemble {
   hipStream_t stream = nullptr;
   int32_t nMembers = 0, precision = 0;
   Consts k{};
   DevState s{};
   int64_t planeStride = 0;
   char *arena = nullptr;
   size_t arenaBytes = 0;
   size_t stateBytes = 0; // leading part of the arena holding the state (for snapshots)
   std::vector<int32_t> nodeStart, elemStart, nodeCount, elemCount;
   EnsTime t{};
   int block[3] = {256, 256, 256}; // force, node, kin
   bool useGraph = true;
   bool timing = false;
   hipGraph_t graph = nullptr;
   hipGraphExec_t exec = nullptr;
   // Brick path: every member is a reference brick; the fused nodal and
   // kinematics kernels run over all members' workgroups, node state
   // ping-pongs between two buffers (parity = which one is current), and a
   // cycle graph is captured for each parity.
   bool brick = false;
   int parity = 0;
   hipGraph_t graph2 = nullptr;
   hipGraphExec_t exec2 = nullptr;
   std::vector<int32_t> memberSize;
   int nodalChunk = 8, kinChunk = 12;
   unsigned nodalBlocks = 0, kinBlocks = 0;
   char *tables = nullptr;
   size_t tablesBytes = 0;
   int64_t kLaunches[kENumKernels] = {};
   double kMs[kENumKernels] = {};
   std::string lastError;
   // pinned host mirror of the time state
   char *host = nullptr;
};

namespace {

int fail_hip(lgpu_ensemble *e, hipError_t err, const char *what)
{
   if (err == hipSuccess) return LGPU_OK;
   if (e) e->lastError = std::string(what) + ": " + hipGetErrorString(err);
   return err == hipErrorOutOfMemory ? LGPU_ENOMEM : LGPU_EDEVICE;
}

#define ENS_HIP(e, call)                                            \
   do {                                                             \
      const hipError_t err_ = (call);                               \
      if (err_ != hipSuccess) return fail_hip((e), err_, #call);    \
   } while (0)

void swap_nodes(lgpu_ensemble *e)
{
   DevState &s = e->s;
   std::swap(s.x, s.xo);
   std::swap(s.y, s.yo);
   std::swap(s.z, s.zo);
   std::swap(s.xd, s.xdo);
   std::swap(s.yd, s.ydo);
   std::swap(s.zd, s.zdo);
   e->parity ^= 1;
}

template <class P>
void launch_cycle(lgpu_ensemble *e)
{
   hipLaunchKernelGGL(k_time_increment_ens, dim3((e->nMembers + 63) / 64), dim3(64), 0, e->stream,
                      e->t, e->s.memberDtOut, e->s.memberErr, e->nMembers);
   if (e->brick) {
      Launch<P>::nodal_brick_members(e->s, e->k, e->nodalBlocks, e->stream);
      swap_nodes(e);
      Launch<P>::kin_brick_members(e->s, e->k, e->kinBlocks, e->stream);
   } else {
      Launch<P>::force(e->s, e->k, e->planeStride, e->block[0], e->stream);
      Launch<P>::node(e->s, e->k, e->planeStride, 0.0, e->block[1], e->stream);
      Launch<P>::kin(e->s, e->k, 0.0, e->block[2], e->stream);
   }
   Launch<P>::eos(e->s, e->k, e->stream);
   Launch<P>::dt_reduce(e->s, e->stream);
}

void enqueue_cycle(lgpu_ensemble *e)
{
   switch (e->precision) {
   case LGPU_MIXED: launch_cycle<PolicyMixed>(e); break;
   case LGPU_FP64_FMA: launch_cycle<PolicyFp64Fma>(e); break;
   default: launch_cycle<PolicyFp64>(e); break;
   }
}

// Per-kernel timing of one cycle (used by the autotuner and the timing
// option): events around each launch.
template <class P>
int timed_cycle(lgpu_ensemble *e, float ms[kENumKernels])
{
   hipEvent_t ev[kENumKernels + 1];
   for (auto &x : ev) ENS_HIP(e, hipEventCreate(&x));
   ENS_HIP(e, hipEventRecord(ev[0], e->stream));
   hipLaunchKernelGGL(k_time_increment_ens, dim3((e->nMembers + 63) / 64), dim3(64), 0, e->stream,
                      e->t, e->s.memberDtOut, e->s.memberErr, e->nMembers);
   ENS_HIP(e, hipEventRecord(ev[1], e->stream));
   if (e->brick) {
      Launch<P>::nodal_brick_members(e->s, e->k, e->nodalBlocks, e->stream);
      swap_nodes(e);
      ENS_HIP(e, hipEventRecord(ev[2], e->stream));
      ENS_HIP(e, hipEventRecord(ev[3], e->stream));
      Launch<P>::kin_brick_members(e->s, e->k, e->kinBlocks, e->stream);
   } else {
      Launch<P>::force(e->s, e->k, e->planeStride, e->block[0], e->stream);
      ENS_HIP(e, hipEventRecord(ev[2], e->stream));
      Launch<P>::node(e->s, e->k, e->planeStride, 0.0, e->block[1], e->stream);
      ENS_HIP(e, hipEventRecord(ev[3], e->stream));
      Launch<P>::kin(e->s, e->k, 0.0, e->block[2], e->stream);
   }
   ENS_HIP(e, hipEventRecord(ev[4], e->stream));
   Launch<P>::eos(e->s, e->k, e->stream);
   ENS_HIP(e, hipEventRecord(ev[5], e->stream));
   Launch<P>::dt_reduce(e->s, e->stream);
   ENS_HIP(e, hipEventRecord(ev[6], e->stream));
   ENS_HIP(e, hipGetLastError());
   ENS_HIP(e, hipStreamSynchronize(e->stream));
   // ev[i]..ev[i+1] brackets: 0 time, 1 force (brick: nodal), 2 node (brick:
   // empty), 3 kin, 4 eos, 5 dt
   const int order[6] = {kETime, e->brick ? kENodalBrick : kEForce, kENode,
                         e->brick ? kEKinBrick : kEKin, kEEos, kEDt};
   for (int i = 0; i < kENumKernels; ++i) ms[i] = 0.f;
   for (int i = 0; i < 6; ++i) ENS_HIP(e, hipEventElapsedTime(&ms[order[i]], ev[i], ev[i + 1]));
   if (e->brick) ms[kENode] = 0.f;
   for (auto &x : ev) (void)hipEventDestroy(x);
   return LGPU_OK;
}

int timed_cycle_any(lgpu_ensemble *e, float ms[kENumKernels])
{
   switch (e->precision) {
   case LGPU_MIXED: return timed_cycle<PolicyMixed>(e, ms);
   case LGPU_FP64_FMA: return timed_cycle<PolicyFp64Fma>(e, ms);
   default: return timed_cycle<PolicyFp64>(e, ms);
   }
}

void drop_graph(lgpu_ensemble *e)
{
   if (e->exec) (void)hipGraphExecDestroy(e->exec);
   if (e->graph) (void)hipGraphDestroy(e->graph);
   if (e->exec2) (void)hipGraphExecDestroy(e->exec2);
   if (e->graph2) (void)hipGraphDestroy(e->graph2);
   e->exec = e->exec2 = nullptr;
   e->graph = e->graph2 = nullptr;
}

int capture_cycle(lgpu_ensemble *e, hipGraph_t &g, hipGraphExec_t &x)
{
   ENS_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
   enqueue_cycle(e);
   const hipError_t endErr = hipStreamEndCapture(e->stream, &g);
   ENS_HIP(e, endErr);
   ENS_HIP(e, hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
   return LGPU_OK;
}

// One cycle graph; on the brick path a second one for the other node-buffer
// parity (exec: captured at parity 0, exec2: at parity 1).  Capturing
// advances the host's parity; it is put back afterwards.
int build_graph(lgpu_ensemble *e)
{
   drop_graph(e);
   if (e->brick) {
      const int p0 = e->parity;
      if (p0) swap_nodes(e);
      int r = capture_cycle(e, e->graph, e->exec);
      if (!r) r = capture_cycle(e, e->graph2, e->exec2);
      if (e->parity != p0) swap_nodes(e);
      if (r) return r;
   } else {
      const int r = capture_cycle(e, e->graph, e->exec);
      if (r) return r;
   }
   return LGPU_OK;
}

// Replays one cycle: the graph of the current parity (the host mirrors the
// pointer swap the graph's nodal launch implies).
int launch_graph(lgpu_ensemble *e)
{
   if (e->brick) {
      ENS_HIP(e, hipGraphLaunch(e->parity ? e->exec2 : e->exec, e->stream));
      swap_nodes(e);
   } else {
      ENS_HIP(e, hipGraphLaunch(e->exec, e->stream));
   }
   return LGPU_OK;
}

// Member tables of the brick path for the current chunk choices: per-member
// bricks with offsets and first workgroups, and workgroup -> member maps.
int build_brick_tables(lgpu_ensemble *e)
{
   const int M = e->nMembers;
   std::vector<MemberBrick> mb((size_t)M);
   std::vector<int32_t> nodalOf, kinOf;
   int64_t nb = 0, kb = 0;
   for (int m = 0; m < M; ++m) {
      Brick g{};
      init_brick(g, e->memberSize[(size_t)m]);
      pick_member_tile(g);
      pick_member_kin_tile(g);
      set_brick_chunk(g, e->nodalChunk);
      set_kin_chunk(g, e->kinChunk);
      mb[(size_t)m].g = g;
      mb[(size_t)m].nodeOff = e->nodeStart[(size_t)m];
      mb[(size_t)m].elemOff = e->elemStart[(size_t)m];
      mb[(size_t)m].nodalBlock0 = (int32_t)nb;
      mb[(size_t)m].kinBlock0 = (int32_t)kb;
      const int64_t n1 = nodal_blocks(g), k1 = kin_blocks(g);
      nodalOf.insert(nodalOf.end(), (size_t)n1, m);
      kinOf.insert(kinOf.end(), (size_t)k1, m);
      nb += n1;
      kb += k1;
   }
   const size_t b0 = align_up(sizeof(MemberBrick) * (size_t)M);
   const size_t b1 = align_up(4 * (size_t)nb), b2 = align_up(4 * (size_t)kb);
   const size_t need = b0 + b1 + b2;
   ENS_HIP(e, hipStreamSynchronize(e->stream));
   if (need > e->tablesBytes) {
      if (e->tables) ENS_HIP(e, hipFree(e->tables));
      e->tables = nullptr;
      e->tablesBytes = 0;
      ENS_HIP(e, hipMalloc((void **)&e->tables, need));
      e->tablesBytes = need;
   }
   ENS_HIP(e, hipMemcpy(e->tables, mb.data(), sizeof(MemberBrick) * (size_t)M, hipMemcpyHostToDevice));
   ENS_HIP(e, hipMemcpy(e->tables + b0, nodalOf.data(), 4 * (size_t)nb, hipMemcpyHostToDevice));
   ENS_HIP(e, hipMemcpy(e->tables + b0 + b1, kinOf.data(), 4 * (size_t)kb, hipMemcpyHostToDevice));
   e->s.memberBrick = reinterpret_cast<const MemberBrick *>(e->tables);
   e->s.nodalBlockMember = reinterpret_cast<const int32_t *>(e->tables + b0);
   e->s.kinBlockMember = reinterpret_cast<const int32_t *>(e->tables + b0 + b1);
   e->nodalBlocks = (unsigned)nb;
   e->kinBlocks = (unsigned)kb;
   drop_graph(e);
   return LGPU_OK;
}

// Device array of a field.  `narrow`: stored as FP32 in the mixed form (the
// Q gradients, p, q and ss).
double *field_base(lgpu_ensemble *e, int f, bool &isElem, bool &narrow)
{
   isElem = f >= LGPU_E;
   narrow = e->precision == LGPU_MIXED &&
            (f >= LGPU_DELX_XI || f == LGPU_P || f == LGPU_Q || f == LGPU_SS);
   DevState &s = e->s;
   switch (f) {
   case LGPU_X: return s.x;
   case LGPU_Y: return s.y;
   case LGPU_Z: return s.z;
   case LGPU_XD: return s.xd;
   case LGPU_YD: return s.yd;
   case LGPU_ZD: return s.zd;
   case LGPU_NODALMASS: return s.nodalMass;
   case LGPU_E: return s.e;
   case LGPU_P: return static_cast<double *>(s.p);
   case LGPU_Q: return static_cast<double *>(s.q);
   case LGPU_V: return s.v;
   case LGPU_VOLO: return s.volo;
   case LGPU_VNEW: return s.vnew;
   case LGPU_DELV: return s.delv;
   case LGPU_VDOV: return s.vdov;
   case LGPU_AREALG: return s.arealg;
   case LGPU_SS: return static_cast<double *>(s.ss);
   case LGPU_ELEMMASS: return s.elemMass;
   case LGPU_DELX_XI: return static_cast<double *>(s.delx_xi);
   case LGPU_DELX_ETA: return static_cast<double *>(s.delx_eta);
   case LGPU_DELX_ZETA: return static_cast<double *>(s.delx_zeta);
   case LGPU_DELV_XI: return static_cast<double *>(s.delv_xi);
   case LGPU_DELV_ETA: return static_cast<double *>(s.delv_eta);
   case LGPU_DELV_ZETA: return static_cast<double *>(s.delv_zeta);
   default: return nullptr;
   }
}

} // namespace

extern "C" {

int lgpu_ensemble_create(const lgpu_mesh_desc *members, int32_t nMembers, lgpu_ensemble **out)
{
   if (!members || nMembers < 1 || !out) return LGPU_EINVAL;
   *out = nullptr;
   const int32_t prec = members[0].precision;
   if (prec != LGPU_FP64 && prec != LGPU_MIXED && prec != LGPU_FP64_FMA) return LGPU_EINVAL;
   for (int32_t m = 0; m < nMembers; ++m) {
      const lgpu_mesh_desc &d = members[m];
      if (d.precision != prec || memcmp(&d.k, &members[0].k, sizeof(lgpu_constants)) != 0) {
         fprintf(stderr, "lgpu_ensemble_create: members must share precision and constants\n");
         return LGPU_EINVAL;
      }
      // Validate each member on its own (connectivity ranges, regions); the
      // size limit is lgpu_create's (8 corners per element index in int32).
      if (d.numElem < 1 || d.numNode < 1 || d.numElem > INT32_MAX / 8 || !d.nodelist ||
          !d.nodeElemStart || !d.nodeElemCornerList || d.numReg < 1 || d.cost < 0)
         return LGPU_EINVAL;
      for (int64_t i = 0; i < 8 * (int64_t)d.numElem; ++i)
         if (d.nodelist[i] < 0 || d.nodelist[i] >= d.numNode) return LGPU_EINVAL;
      const int32_t *nb[6] = {d.lxim, d.lxip, d.letam, d.letap, d.lzetam, d.lzetap};
      for (const int32_t *a : nb) {
         if (!a) return LGPU_EINVAL;
         for (int32_t i = 0; i < d.numElem; ++i)
            if (a[i] < 0 || a[i] >= d.numElem) return LGPU_EINVAL;
      }
      for (int32_t n = 0; n < d.numNode; ++n)
         if (d.nodeElemStart[n + 1] < d.nodeElemStart[n]) return LGPU_EINVAL;
      if (d.nodeElemStart[0] != 0 || d.nodeElemStart[d.numNode] > 8 * d.numElem) return LGPU_EINVAL;
      for (int32_t c = 0; c < d.nodeElemStart[d.numNode]; ++c)
         if (d.nodeElemCornerList[c] < 0 || d.nodeElemCornerList[c] >= 8 * d.numElem) return LGPU_EINVAL;
      int64_t tot = 0;
      std::vector<uint8_t> seen((size_t)d.numElem, 0);
      for (int32_t r = 0; r < d.numReg; ++r)
         for (int32_t i = 0; i < d.regElemSize[r]; ++i) {
            const int32_t el = d.regElemlist[r][i];
            if (el < 0 || el >= d.numElem || seen[(size_t)el]) return LGPU_EINVAL;
            seen[(size_t)el] = 1;
            ++tot;
         }
      if (tot != d.numElem) return LGPU_EINVAL;
      const int32_t *sy[3] = {d.symmX, d.symmY, d.symmZ};
      const int32_t ns[3] = {d.numSymmX, d.numSymmY, d.numSymmZ};
      for (int a = 0; a < 3; ++a)
         for (int32_t i = 0; i < ns[a]; ++i)
            if (!sy[a] || sy[a][i] < 0 || sy[a][i] >= d.numNode) return LGPU_EINVAL;
   }

   lgpu_ensemble *e = new (std::nothrow) lgpu_ensemble();
   if (!e) return LGPU_ENOMEM;
   auto bail = [&](int code) {
      lgpu_ensemble_destroy(e);
      return code;
   };
   e->nMembers = nMembers;
   e->precision = prec;
   // Brick path when every member is the reference's brick (and the generic
   // kernels are not asked for).
   e->brick = !(members[0].flags & LGPU_FLAG_GENERIC);
   for (int32_t m = 0; m < nMembers && e->brick; ++m) e->brick = is_reference_brick(&members[m]);
   for (int32_t m = 0; m < nMembers; ++m) e->memberSize.push_back(members[m].sizeX);
   const lgpu_constants &kk = members[0].k;
   e->k = Consts{kk.e_cut, kk.p_cut, kk.q_cut, kk.v_cut, kk.u_cut, kk.hgcoef, kk.ss4o3,
                 kk.qstop, kk.monoq_max_slope, kk.monoq_limiter_mult, kk.qlc_monoq,
                 kk.qqc_monoq, kk.qqc, kk.eosvmax, kk.eosvmin, kk.pmin, kk.emin, kk.dvovmax,
                 kk.refdens};
   e->k.fastEos = fast_eos_consts(e->k);
   if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(LGPU_EDEVICE);

   // The concatenated arrays are indexed with int32 (corners: 8 per element).
   {
      int64_t totE = 0, totN = 0;
      for (int32_t m = 0; m < nMembers; ++m) {
         totE += members[m].numElem;
         totN += members[m].numNode;
      }
      if (totE > INT32_MAX / 8 || totN > INT32_MAX) {
         fprintf(stderr, "lgpu_ensemble_create: %lld elements / %lld nodes in all: too many\n",
                 (long long)totE, (long long)totN);
         return bail(LGPU_EINVAL);
      }
   }
   // Offsets.
   e->nodeStart.assign((size_t)nMembers + 1, 0);
   e->elemStart.assign((size_t)nMembers + 1, 0);
   std::vector<int64_t> cornerStart((size_t)nMembers + 1, 0);
   std::vector<int32_t> blockStart((size_t)nMembers + 1, 0);
   int64_t numRegTot = 0;
   for (int32_t m = 0; m < nMembers; ++m) {
      const lgpu_mesh_desc &d = members[m];
      e->nodeStart[m + 1] = e->nodeStart[m] + d.numNode;
      e->elemStart[m + 1] = e->elemStart[m] + d.numElem;
      cornerStart[m + 1] = cornerStart[m] + d.nodeElemStart[d.numNode];
      blockStart[m + 1] = blockStart[m] + (int32_t)eos_blocks(d.numElem);
      e->nodeCount.push_back(d.numNode);
      e->elemCount.push_back(d.numElem);
      numRegTot += d.numReg;
   }
   const size_t E = (size_t)e->elemStart[nMembers], N = (size_t)e->nodeStart[nMembers];
   const size_t nPerm = (size_t)blockStart[nMembers] * kEosBlock;
   const size_t nCorner = (size_t)cornerStart[nMembers];
   const size_t cb = prec == LGPU_MIXED ? 4 : 8, gb = cb;
   const int64_t plane = (int64_t)(align_up(E * cb) / cb);
   e->planeStride = plane;

   // Arena: state first (snapshot-able), then connectivity and scratch.
   struct Slot { void **dst; size_t bytes; };
   std::vector<Slot> slots;
   DevState &s = e->s;
   auto add = [&](void **dst, size_t bytes) { slots.push_back({dst, align_up(bytes)}); };
   double **stateNode[] = {&s.x, &s.y, &s.z, &s.xd, &s.yd, &s.zd};
   // two doubles of padding, which the kinematics kernel's pair copies may
   // read past the last node
   for (double **p : stateNode) add((void **)p, 8 * (N + 2));
   if (e->brick) {
      // the second node buffer is state too: either may be current
      double **second[] = {&s.xo, &s.yo, &s.zo, &s.xdo, &s.ydo, &s.zdo};
      for (double **p : second) add((void **)p, 8 * (N + 2));
   }
   double **stateElem[] = {&s.e, &s.v};
   for (double **p : stateElem) add((void **)p, 8 * E);
   void **statePqs[] = {&s.p, &s.q, &s.ss};
   for (void **p : statePqs) add(p, gb * E);
   EnsTime &t = e->t;
   double **tf[] = {&t.time, &t.deltatime, &t.dtcourant, &t.dthydro, &t.stoptime, &t.dtfixed,
                    &t.dtmax, &t.multlb, &t.multub};
   for (double **p : tf) add((void **)p, 8 * (size_t)nMembers);
   add((void **)&t.cycle, 4 * (size_t)nMembers);
   add((void **)&t.active, 4 * (size_t)nMembers);
   add((void **)&t.maxCycle, 4);
   add((void **)&s.memberErr, 4 * (size_t)nMembers * kNumErrSites);
   add((void **)&s.memberDtOut, 16 * (size_t)nMembers);
   const size_t stateSlots = slots.size();
   add((void **)&s.nodalMass, 8 * N);
   double **elemFix[] = {&s.volo, &s.vnew, &s.delv, &s.elemMass, &s.arealg, &s.vdov};
   for (double **p : elemFix) add((void **)p, 8 * E);
   void **grads[] = {&s.delx_xi, &s.delx_eta, &s.delx_zeta, &s.delv_xi, &s.delv_eta, &s.delv_zeta};
   for (void **p : grads) add(p, gb * E);
   add((void **)&s.nodeBC, N);
   add((void **)&s.nodeElemStart, 4 * (N + 1));
   add((void **)&s.cornerList, 4 * std::max<size_t>(nCorner, 1));
   add((void **)&s.nodelist, 32 * E);
   const int32_t **nbr[] = {&s.lxim, &s.lxip, &s.letam, &s.letap, &s.lzetam, &s.lzetap, &s.elemBC};
   for (const int32_t **p : nbr) add((void **)p, 4 * E);
   if (!e->brick) add(&s.corner, cb * (size_t)plane * 48);
   add((void **)&s.perm, 4 * nPerm);
   add((void **)&s.regStart, 4 * ((size_t)numRegTot + 1));
   add((void **)&s.regRep, 4 * (size_t)numRegTot);
   add((void **)&s.err, 4 * kNumErrSites);
   add((void **)&s.dtPartial, 16 * (size_t)blockStart[nMembers]);
   add((void **)&s.dtOut, 16);
   add((void **)&s.memberNodeStart, 4 * ((size_t)nMembers + 1));
   add((void **)&s.memberElemStart, 4 * ((size_t)nMembers + 1));
   add((void **)&s.blockMember, 4 * (size_t)blockStart[nMembers]);
   add((void **)&s.eosSlot, sizeof(EosSlot) * (size_t)blockStart[nMembers]);
   add((void **)&s.memberBlockStart, 4 * ((size_t)nMembers + 1));
   size_t total = 0, stateBytes = 0;
   for (size_t i = 0; i < slots.size(); ++i) {
      if (i == stateSlots) stateBytes = total;
      total += slots[i].bytes;
   }
   if (hipMalloc((void **)&e->arena, total) != hipSuccess) {
      (void)hipGetLastError();
      return bail(LGPU_ENOMEM);
   }
   e->arenaBytes = total;
   e->stateBytes = stateBytes;
   if (hipMemset(e->arena, 0, total) != hipSuccess) return bail(LGPU_EDEVICE);
   size_t off = 0;
   for (const Slot &sl : slots) {
      *sl.dst = e->arena + off;
      off += sl.bytes;
   }
   s.numElem = (int32_t)E;
   s.numNode = (int32_t)N;
   s.numReg = (int32_t)numRegTot;
   s.numPerm = (int32_t)nPerm;
   s.xcdRemap = 1;
   s.isBrick = 0;
   s.nMembers = nMembers;
   s.memberDt = t.deltatime;
   s.memberActive = t.active;

   // Concatenated connectivity.
   std::vector<int32_t> nodelist(8 * E), nbrs[7], nodeElemStart(N + 1), corner(nCorner), perm(nPerm, -1);
   for (auto &v : nbrs) v.assign(E, 0);
   std::vector<int32_t> regStart((size_t)numRegTot + 1, 0), regRep((size_t)numRegTot, 1);
   std::vector<int32_t> blockMember((size_t)blockStart[nMembers], -1);
   std::vector<uint8_t> bc(N, 0);
   int64_t regBase = 0;
   for (int32_t m = 0; m < nMembers; ++m) {
      const lgpu_mesh_desc &d = members[m];
      const int32_t n0 = e->nodeStart[m], e0 = e->elemStart[m];
      for (int64_t i = 0; i < 8 * (int64_t)d.numElem; ++i) nodelist[8 * (size_t)e0 + i] = d.nodelist[i] + n0;
      const int32_t *src[7] = {d.lxim, d.lxip, d.letam, d.letap, d.lzetam, d.lzetap, d.elemBC};
      for (int a = 0; a < 7; ++a)
         for (int32_t i = 0; i < d.numElem; ++i)
            nbrs[a][(size_t)(e0 + i)] = (a < 6) ? src[a][i] + e0 : src[a][i];
      for (int32_t n = 0; n <= d.numNode; ++n)
         nodeElemStart[(size_t)(n0 + n)] = (int32_t)(cornerStart[m] + d.nodeElemStart[n]);
      for (int32_t c = 0; c < d.nodeElemStart[d.numNode]; ++c)
         corner[(size_t)(cornerStart[m] + c)] = d.nodeElemCornerList[c] + 8 * e0;
      const int32_t *sy[3] = {d.symmX, d.symmY, d.symmZ};
      const int32_t ns[3] = {d.numSymmX, d.numSymmY, d.numSymmZ};
      for (int a = 0; a < 3; ++a)
         for (int32_t i = 0; i < ns[a]; ++i) bc[(size_t)(n0 + sy[a][i])] |= (uint8_t)(1u << a);
      int64_t pos = (int64_t)blockStart[m] * kEosBlock;
      // As for a single domain: the member's elements grouped by EOS repeat
      // count, largest first, in mesh order within a group; the member's
      // remaining region slots stay empty at its end.
      std::vector<int32_t> repOf((size_t)d.numElem, 1), reps;
      for (int32_t r = 0; r < d.numReg; ++r) {
         const int32_t rep = region_rep(r, d.numReg, d.cost);
         if (d.regElemSize[r] > 0 && std::find(reps.begin(), reps.end(), rep) == reps.end())
            reps.push_back(rep);
         for (int32_t i = 0; i < d.regElemSize[r]; ++i) repOf[(size_t)d.regElemlist[r][i]] = rep;
      }
      std::sort(reps.begin(), reps.end(), std::greater<int32_t>());
      for (int32_t k = 0; k < d.numReg; ++k) {
         regStart[(size_t)(regBase + k)] = (int32_t)pos;
         regRep[(size_t)(regBase + k)] = k < (int32_t)reps.size() ? reps[(size_t)k] : 1;
         if (k < (int32_t)reps.size())
            for (int32_t el = 0; el < d.numElem; ++el)
               if (repOf[(size_t)el] == reps[(size_t)k]) perm[(size_t)pos++] = el + e0;
      }
      regBase += d.numReg;
      for (int32_t b = blockStart[m]; b < blockStart[m + 1]; ++b) blockMember[(size_t)b] = m;
   }
   regStart[(size_t)numRegTot] = (int32_t)nPerm;
   // k_eos dispatch slots: blocks in order, each with its repeat counts
   std::vector<int32_t> blockOrder((size_t)blockStart[nMembers]);
   for (size_t b = 0; b < blockOrder.size(); ++b) blockOrder[b] = (int32_t)b;
   const std::vector<EosSlot> eosSlots = eos_slots(blockOrder, regStart, regRep, (int64_t)nPerm);

   auto h2d = [&](void *dst, const void *src, size_t bytes) {
      return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
   };
   std::vector<double> ones(E, 1.0);
   bool ok = h2d((void *)s.nodelist, nodelist.data(), 4 * nodelist.size()) &&
             h2d((void *)s.lxim, nbrs[0].data(), 4 * E) && h2d((void *)s.lxip, nbrs[1].data(), 4 * E) &&
             h2d((void *)s.letam, nbrs[2].data(), 4 * E) && h2d((void *)s.letap, nbrs[3].data(), 4 * E) &&
             h2d((void *)s.lzetam, nbrs[4].data(), 4 * E) && h2d((void *)s.lzetap, nbrs[5].data(), 4 * E) &&
             h2d((void *)s.elemBC, nbrs[6].data(), 4 * E) &&
             h2d((void *)s.nodeElemStart, nodeElemStart.data(), 4 * (N + 1)) &&
             (nCorner == 0 || h2d((void *)s.cornerList, corner.data(), 4 * nCorner)) &&
             h2d((void *)s.perm, perm.data(), 4 * nPerm) &&
             h2d((void *)s.regStart, regStart.data(), 4 * regStart.size()) &&
             h2d((void *)s.regRep, regRep.data(), 4 * regRep.size()) &&
             h2d((void *)s.nodeBC, bc.data(), N) &&
             h2d((void *)s.memberNodeStart, e->nodeStart.data(), 4 * e->nodeStart.size()) &&
             h2d((void *)s.memberElemStart, e->elemStart.data(), 4 * e->elemStart.size()) &&
             h2d((void *)s.blockMember, blockMember.data(), 4 * blockMember.size()) &&
             h2d((void *)s.eosSlot, eosSlots.data(), sizeof(EosSlot) * eosSlots.size()) &&
             h2d((void *)s.memberBlockStart, blockStart.data(), 4 * blockStart.size()) &&
             h2d(s.v, ones.data(), 8 * E);
   if (!ok) return bail(LGPU_EDEVICE);
   if (e->brick && build_brick_tables(e) != LGPU_OK) return bail(LGPU_EDEVICE);
   if (hipHostMalloc((void **)&e->host, (size_t)nMembers * (9 * 8 + 8) + 64 +
                                            (size_t)nMembers * kNumErrSites * 4,
                     hipHostMallocDefault) != hipSuccess)
      return bail(LGPU_ENOMEM);
   *out = e;
   return LGPU_OK;
}

void lgpu_ensemble_destroy(lgpu_ensemble *e)
{
   if (!e) return;
   if (e->stream) (void)hipStreamSynchronize(e->stream);
   drop_graph(e);
   if (e->arena) (void)hipFree(e->arena);
   if (e->tables) (void)hipFree(e->tables);
   if (e->host) (void)hipHostFree(e->host);
   if (e->stream) (void)hipStreamDestroy(e->stream);
   delete e;
}

int lgpu_ensemble_upload(lgpu_ensemble *e, int32_t member, int field, const double *host)
{
   if (!e || !host || member < 0 || member >= e->nMembers) return LGPU_EINVAL;
   bool isElem, narrow;
   double *base = field_base(e, field, isElem, narrow);
   if (!base) return LGPU_EINVAL;
   const size_t n = isElem ? (size_t)e->elemCount[member] : (size_t)e->nodeCount[member];
   const size_t off = isElem ? (size_t)e->elemStart[member] : (size_t)e->nodeStart[member];
   ENS_HIP(e, hipStreamSynchronize(e->stream));
   if (narrow) {
      std::vector<float> tmp(n);
      for (size_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
      ENS_HIP(e, hipMemcpy(reinterpret_cast<float *>(base) + off, tmp.data(), 4 * n, hipMemcpyHostToDevice));
   } else {
      ENS_HIP(e, hipMemcpy(base + off, host, 8 * n, hipMemcpyHostToDevice));
   }
   return LGPU_OK;
}

int lgpu_ensemble_download(lgpu_ensemble *e, int32_t member, int field, double *host)
{
   if (!e || !host || member < 0 || member >= e->nMembers) return LGPU_EINVAL;
   bool isElem, narrow;
   double *base = field_base(e, field, isElem, narrow);
   if (!base) return LGPU_EINVAL;
   const size_t n = isElem ? (size_t)e->elemCount[member] : (size_t)e->nodeCount[member];
   const size_t off = isElem ? (size_t)e->elemStart[member] : (size_t)e->nodeStart[member];
   ENS_HIP(e, hipStreamSynchronize(e->stream));
   if (narrow) {
      std::vector<float> tmp(n);
      ENS_HIP(e, hipMemcpy(tmp.data(), reinterpret_cast<float *>(base) + off, 4 * n, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
   } else {
      ENS_HIP(e, hipMemcpy(host, base + off, 8 * n, hipMemcpyDeviceToHost));
   }
   return LGPU_OK;
}

int lgpu_ensemble_run(lgpu_ensemble *e, lgpu_time_state *ts, int32_t maxCycle, int32_t *status)
{
   if (!e || !ts) return LGPU_EINVAL;
   const int M = e->nMembers;
   // Pinned staging: 9 double arrays, cycle, active, err.
   double *hd = reinterpret_cast<double *>(e->host);
   int32_t *hcyc = reinterpret_cast<int32_t *>(hd + 9 * (size_t)M);
   int32_t *hact = hcyc + M;
   int32_t *herr = hact + M;
   double *cols[9];
   for (int i = 0; i < 9; ++i) cols[i] = hd + (size_t)i * M;
   for (int m = 0; m < M; ++m) {
      cols[0][m] = ts[m].time;
      cols[1][m] = ts[m].deltatime;
      cols[2][m] = ts[m].dtcourant;
      cols[3][m] = ts[m].dthydro;
      cols[4][m] = ts[m].stoptime;
      cols[5][m] = ts[m].dtfixed;
      cols[6][m] = ts[m].dtmax;
      cols[7][m] = ts[m].deltatimemultlb;
      cols[8][m] = ts[m].deltatimemultub;
      hcyc[m] = ts[m].cycle;
      hact[m] = 1;
   }
   EnsTime &t = e->t;
   double *dcols[9] = {t.time, t.deltatime, t.dtcourant, t.dthydro, t.stoptime, t.dtfixed,
                       t.dtmax, t.multlb, t.multub};
   for (int i = 0; i < 9; ++i)
      ENS_HIP(e, hipMemcpyAsync(dcols[i], cols[i], 8 * (size_t)M, hipMemcpyHostToDevice, e->stream));
   ENS_HIP(e, hipMemcpyAsync(t.cycle, hcyc, 4 * (size_t)M, hipMemcpyHostToDevice, e->stream));
   ENS_HIP(e, hipMemcpyAsync(t.active, hact, 4 * (size_t)M, hipMemcpyHostToDevice, e->stream));
   // dt results feed the first TimeIncrement of members past cycle 0.
   {
      std::vector<double> dtout(2 * (size_t)M);
      for (int m = 0; m < M; ++m) {
         dtout[2 * (size_t)m] = ts[m].dtcourant;
         dtout[2 * (size_t)m + 1] = ts[m].dthydro;
      }
      ENS_HIP(e, hipMemcpyAsync(e->s.memberDtOut, dtout.data(), 16 * (size_t)M,
                                hipMemcpyHostToDevice, e->stream));
      ENS_HIP(e, hipMemcpyAsync(t.maxCycle, &maxCycle, 4, hipMemcpyHostToDevice, e->stream));
      ENS_HIP(e, hipStreamSynchronize(e->stream));
   }
   if (e->useGraph && !e->timing && !e->exec) {
      const int r = build_graph(e);
      if (r) return r;
   }
   const int batch = 8;
   for (;;) {
      for (int b = 0; b < batch; ++b) {
         if (e->timing) {
            float ms[kENumKernels];
            const int r = timed_cycle_any(e, ms);
            if (r) return r;
            for (int i = 0; i < kENumKernels; ++i) {
               e->kLaunches[i] += 1;
               e->kMs[i] += ms[i];
            }
         } else if (e->useGraph) {
            const int r = launch_graph(e);
            if (r) return r;
         } else {
            enqueue_cycle(e);
            ENS_HIP(e, hipGetLastError());
         }
      }
      ENS_HIP(e, hipMemcpyAsync(hact, t.active, 4 * (size_t)M, hipMemcpyDeviceToHost, e->stream));
      ENS_HIP(e, hipStreamSynchronize(e->stream));
      bool any = false;
      for (int m = 0; m < M; ++m) any = any || hact[m] != 0;
      if (!any) break;
   }
   // Read back the time state; a member's last dt results are in
   // memberDtOut (inactive members leave them untouched).
   for (int i = 0; i < 9; ++i)
      ENS_HIP(e, hipMemcpyAsync(cols[i], dcols[i], 8 * (size_t)M, hipMemcpyDeviceToHost, e->stream));
   ENS_HIP(e, hipMemcpyAsync(hcyc, t.cycle, 4 * (size_t)M, hipMemcpyDeviceToHost, e->stream));
   ENS_HIP(e, hipMemcpyAsync(herr, e->s.memberErr, 4 * (size_t)M * kNumErrSites,
                             hipMemcpyDeviceToHost, e->stream));
   std::vector<double> dtout(2 * (size_t)M);
   ENS_HIP(e, hipMemcpyAsync(dtout.data(), e->s.memberDtOut, 16 * (size_t)M, hipMemcpyDeviceToHost,
                             e->stream));
   ENS_HIP(e, hipStreamSynchronize(e->stream));
   int worst = LGPU_OK;
   for (int m = 0; m < M; ++m) {
      ts[m].time = cols[0][m];
      ts[m].deltatime = cols[1][m];
      ts[m].cycle = hcyc[m];
      if (hcyc[m] > 0) {
         ts[m].dtcourant = dtout[2 * (size_t)m];
         ts[m].dthydro = dtout[2 * (size_t)m + 1];
      }
      const int32_t *er = herr + (size_t)m * kNumErrSites;
      int st = LGPU_OK;
      if (er[kErrForceDeterm] || er[kErrHourglassV] || er[kErrKinVnew]) st = LGPU_VOLUME_ERROR;
      else if (er[kErrQStop]) st = LGPU_QSTOP_ERROR;
      else if (er[kErrMaterialV]) st = LGPU_VOLUME_ERROR;
      if (status) status[m] = st;
      if (st != LGPU_OK && worst == LGPU_OK) worst = st;
   }
   return worst;
}

int lgpu_ensemble_set_option(lgpu_ensemble *e, int option, int64_t value)
{
   if (!e) return LGPU_EINVAL;
   switch (option) {
   case LGPU_ENS_OPT_GRAPH:
      e->useGraph = value != 0;
      drop_graph(e);
      return LGPU_OK;
   case LGPU_ENS_OPT_TIMING:
      e->timing = value != 0;
      return LGPU_OK;
   case LGPU_ENS_OPT_BLOCK_FORCE:
   case LGPU_ENS_OPT_BLOCK_NODE:
   case LGPU_ENS_OPT_BLOCK_KIN:
      if (value != 64 && value != 128 && value != 256) return LGPU_EINVAL;
      e->block[option - LGPU_ENS_OPT_BLOCK_FORCE] = (int)value;
      drop_graph(e);
      return LGPU_OK;
   case LGPU_ENS_OPT_XCD_REMAP:
      e->s.xcdRemap = value ? 1 : 0;
      drop_graph(e);
      return LGPU_OK;
   case LGPU_ENS_OPT_BRICK_CHUNK:
   case LGPU_ENS_OPT_KIN_CHUNK:
      if (!e->brick || value < 1 || value > 1024) return LGPU_EINVAL;
      (option == LGPU_ENS_OPT_BRICK_CHUNK ? e->nodalChunk : e->kinChunk) = (int)value;
      return build_brick_tables(e);
   default:
      return LGPU_EINVAL;
   }
}

int lgpu_ensemble_get_option(lgpu_ensemble *e, int option, int64_t *value)
{
   if (!e || !value) return LGPU_EINVAL;
   switch (option) {
   case LGPU_ENS_OPT_GRAPH: *value = e->useGraph; return LGPU_OK;
   case LGPU_ENS_OPT_TIMING: *value = e->timing; return LGPU_OK;
   case LGPU_ENS_OPT_BLOCK_FORCE:
   case LGPU_ENS_OPT_BLOCK_NODE:
   case LGPU_ENS_OPT_BLOCK_KIN: *value = e->block[option - LGPU_ENS_OPT_BLOCK_FORCE]; return LGPU_OK;
   case LGPU_ENS_OPT_XCD_REMAP: *value = e->s.xcdRemap; return LGPU_OK;
   case LGPU_ENS_OPT_BRICK_CHUNK:
      if (!e->brick) return LGPU_EINVAL;
      *value = e->nodalChunk;
      return LGPU_OK;
   case LGPU_ENS_OPT_KIN_CHUNK:
      if (!e->brick) return LGPU_EINVAL;
      *value = e->kinChunk;
      return LGPU_OK;
   case LGPU_ENS_OPT_PATH: *value = e->brick ? 1 : 0; return LGPU_OK;
   default: return LGPU_EINVAL;
   }
}

int lgpu_ensemble_autotune(lgpu_ensemble *e, const lgpu_time_state *ts, int32_t trialCycles)
{
   if (!e || !ts || trialCycles < 1) return LGPU_EINVAL;
   // The trials advance the state: snapshot the state part of the arena (on
   // the brick path both node buffers) and restore it, and the node-buffer
   // parity, after every trial.
   char *snap = nullptr;
   ENS_HIP(e, hipMalloc((void **)&snap, e->stateBytes));
   ENS_HIP(e, hipMemcpyAsync(snap, e->arena, e->stateBytes, hipMemcpyDeviceToDevice, e->stream));
   const int M = e->nMembers;
   const int parity0 = e->parity;
   std::vector<lgpu_time_state> tsv(ts, ts + M);
   const bool savedTiming = e->timing, savedGraph = e->useGraph;
   e->timing = true;
   auto trial = [&](float ms[kENumKernels]) -> int {
      for (int i = 0; i < kENumKernels; ++i) ms[i] = 0.f;
      for (int i = 0; i < kENumKernels; ++i) {
         e->kLaunches[i] = 0;
         e->kMs[i] = 0.0;
      }
      std::vector<lgpu_time_state> t2 = tsv;
      const int r = lgpu_ensemble_run(e, t2.data(), t2[0].cycle + trialCycles, nullptr);
      if (r <= -100) return r;
      for (int i = 0; i < kENumKernels; ++i)
         ms[i] = e->kLaunches[i] ? (float)(e->kMs[i] / e->kLaunches[i]) : 0.f;
      ENS_HIP(e, hipMemcpyAsync(e->arena, snap, e->stateBytes, hipMemcpyDeviceToDevice, e->stream));
      ENS_HIP(e, hipStreamSynchronize(e->stream));
      if (e->parity != parity0) swap_nodes(e);
      return LGPU_OK;
   };
   int r = LGPU_OK;
   if (e->brick) {
      // Chunk of each fused kernel over a few candidates, then the XCD mapping.
      const int nodalCands[] = {4, 6, 8, 11, 16, 22}, kinCands[] = {4, 6, 8, 12, 16};
      float best = 1e30f;
      int bestC = e->nodalChunk;
      for (int c : nodalCands) {
         e->nodalChunk = c;
         r = build_brick_tables(e);
         float ms[kENumKernels];
         if (!r) r = trial(ms);
         if (r) break;
         if (ms[kENodalBrick] < best) {
            best = ms[kENodalBrick];
            bestC = c;
         }
      }
      e->nodalChunk = bestC;
      best = 1e30f;
      bestC = e->kinChunk;
      for (int c : kinCands) {
         if (r) break;
         e->kinChunk = c;
         r = build_brick_tables(e);
         float ms[kENumKernels];
         if (!r) r = trial(ms);
         if (r) break;
         if (ms[kEKinBrick] < best) {
            best = ms[kEKinBrick];
            bestC = c;
         }
      }
      e->kinChunk = bestC;
      if (!r) r = build_brick_tables(e);
   } else {
      // Each of the three launch-shape-sensitive kernels is tuned over the
      // block sizes.
      const int cands[3] = {64, 128, 256};
      const int kern[3] = {kEForce, kENode, kEKin};
      for (int which = 0; which < 3 && !r; ++which) {
         float best = 1e30f;
         int bestB = e->block[which];
         for (int b : cands) {
            e->block[which] = b;
            float ms[kENumKernels];
            r = trial(ms);
            if (r) break;
            if (ms[kern[which]] < best) {
               best = ms[kern[which]];
               bestB = b;
            }
         }
         e->block[which] = bestB;
      }
   }
   if (!r) {
      float tot[2];
      for (int remap = 0; remap < 2 && !r; ++remap) {
         e->s.xcdRemap = remap;
         float ms[kENumKernels];
         r = trial(ms);
         tot[remap] = 0.f;
         for (int i = 0; i < kENumKernels; ++i)
            if (i != kETime && i != kEDt) tot[remap] += ms[i];
      }
      if (!r) e->s.xcdRemap = (tot[1] <= tot[0]) ? 1 : 0;
   }
   e->timing = savedTiming;
   e->useGraph = savedGraph;
   for (int i = 0; i < kENumKernels; ++i) {
      e->kLaunches[i] = 0;
      e->kMs[i] = 0.0;
   }
   // the cycle graphs of the chosen configuration, built here rather than
   // in the first run
   drop_graph(e);
   if (!r && e->useGraph && !e->timing) r = build_graph(e);
   (void)hipFree(snap);
   return r;
}

int lgpu_ensemble_kernel_stats(lgpu_ensemble *e, int kernel, int64_t *launches, double *total_ms)
{
   if (!e || kernel < 0 || kernel >= kENumKernels) return LGPU_EINVAL;
   if (launches) *launches = e->kLaunches[kernel];
   if (total_ms) *total_ms = e->kMs[kernel];
   return LGPU_OK;
}

const char *lgpu_ensemble_kernel_name(int kernel)
{
   return (kernel >= 0 && kernel < kENumKernels) ? kEnsKernelNames[kernel] : "";
}

int lgpu_ensemble_num_kernels(void) { return kENumKernels; }

int64_t lgpu_ensemble_num_elem(lgpu_ensemble *e) { return e ? e->s.numElem : 0; }

const char *lgpu_ensemble_last_error(lgpu_ensemble *e) { return e ? e->lastError.c_str() : ""; }

} // extern "C"
