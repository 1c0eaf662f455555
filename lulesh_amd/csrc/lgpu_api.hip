// lgpu_api.hip -- device context and C ABI (include/lulesh_gpu.h).
//
// One context owns one HIP stream and one device arena holding every array
// the step touches, laid out SoA with 256-byte aligned planes.  All
// connectivity is validated on the host before the first launch, so no
// kernel ever indexes outside the arena.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "lgpu_brick.h"
#include "lgpu_kernels.h"
#include "lgpu_launch.h"
#include "lgpu_setup.h"
#include "lgpu_time.h"
#include "lulesh_gpu.h"

using namespace lgpu;

namespace {

enum KernelId {
   kKForce = 0, kKNode, kKKin, kKEos, kKDtReduce, kKTimeIncrement, kKNodalBrick, kKKinBrick,
   kNumKernels
};
const char *const kKernelNames[kNumKernels] = {"k_force", "k_node", "k_kin", "k_eos",
                                               "k_dt_reduce", "k_time_increment",
                                               "k_nodal_brick", "k_kin_brick"};

constexpr size_t kAlign = 256;

inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

struct TimedLaunch {
   int kernel;
   hipEvent_t start, stop;
};

} // namespace

struct lgpu_ctx {
   int device = 0;
   hipStream_t stream = nullptr;
   int32_t numElem = 0, numNode = 0, numReg = 0, cost = 0, precision = 0;
   int32_t sizeX = 0;
   int block = 256;
   bool keepDiag = false;
   std::vector<int32_t> eosBlockRep; // repeat count of each k_eos block (region order)
   std::vector<int32_t> hostRegStart, hostRegRep; // the work-list groups (host copies)
   int eosInterleave = 0;
   int32_t brickPerCu = 0;
   bool timing = false;
   uint32_t timingMask = 0xffffffffu;
   Consts k{};
   char *arena = nullptr;
   size_t arenaBytes = 0;
   int64_t planeStride = 0; // corner plane length (elements, padded)
   DevState s{};
   // pinned readback: dt[2], err[kNumErrSites]
   double *hostDt = nullptr;
   int32_t *hostErr = nullptr;
   DevTime *devTime = nullptr;
   DevTime *hostTime = nullptr;
   // diagnostics storage (allocated on first use)
   double *diag = nullptr;
   std::vector<TimedLaunch> pending;
   std::vector<hipEvent_t> eventPool;
   int64_t kLaunches[kNumKernels] = {};
   double kMs[kNumKernels] = {};
   std::string lastError;
};

namespace {

int set_hip_error(lgpu_ctx *c, hipError_t e, const char *what)
{
   if (e == hipSuccess) return LGPU_OK;
   if (c) c->lastError = std::string(what) + ": " + hipGetErrorString(e);
   return (e == hipErrorOutOfMemory) ? LGPU_ENOMEM : LGPU_EDEVICE;
}

#define HIP_OR_RETURN(c, call)                                   \
   do {                                                          \
      const hipError_t e_ = (call);                              \
      if (e_ != hipSuccess) return set_hip_error((c), e_, #call); \
   } while (0)

hipEvent_t take_event(lgpu_ctx *c)
{
   if (!c->eventPool.empty()) {
      hipEvent_t e = c->eventPool.back();
      c->eventPool.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   if (hipEventCreate(&e) != hipSuccess) return nullptr;
   return e;
}

// Folds finished timed launches into the per-kernel totals.  Call only after
// the stream has been synchronized.
void harvest_timing(lgpu_ctx *c)
{
   for (TimedLaunch &t : c->pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
         c->kLaunches[t.kernel] += 1;
         c->kMs[t.kernel] += ms;
      }
      c->eventPool.push_back(t.start);
      c->eventPool.push_back(t.stop);
   }
   c->pending.clear();
}

template <class F>
int timed(lgpu_ctx *c, int kernel, F &&launch)
{
   TimedLaunch t{kernel, nullptr, nullptr};
   const bool on = c->timing && ((c->timingMask >> kernel) & 1u);
   if (on) {
      t.start = take_event(c);
      t.stop = take_event(c);
      if (t.start && t.stop) HIP_OR_RETURN(c, hipEventRecord(t.start, c->stream));
   }
   launch();
   HIP_OR_RETURN(c, hipGetLastError());
   if (on && t.start && t.stop) {
      HIP_OR_RETURN(c, hipEventRecord(t.stop, c->stream));
      c->pending.push_back(t);
   }
   return LGPU_OK;
}

bool in_range(const int32_t *a, int64_t n, int32_t lo, int32_t hi)
{
   for (int64_t i = 0; i < n; ++i)
      if (a[i] < lo || a[i] >= hi) return false;
   return true;
}

int validate(const lgpu_mesh_desc *d, std::string &why)
{
   if (!d) { why = "null description"; return LGPU_EINVAL; }
   const int64_t E = d->numElem, N = d->numNode;
   if (E < 1 || N < 1 || E > (INT32_MAX / 8)) { why = "numElem/numNode out of range"; return LGPU_EINVAL; }
   if (!d->nodelist || !d->lxim || !d->lxip || !d->letam || !d->letap || !d->lzetam ||
       !d->lzetap || !d->elemBC || !d->nodeElemStart || !d->nodeElemCornerList ||
       !d->regElemSize || !d->regElemlist) {
      why = "missing array";
      return LGPU_EINVAL;
   }
   if (d->precision != LGPU_FP64 && d->precision != LGPU_MIXED && d->precision != LGPU_FP64_FMA) {
      why = "precision";
      return LGPU_EINVAL;
   }
   if (d->numReg < 1 || d->cost < 0) { why = "numReg/cost"; return LGPU_EINVAL; }
   if (!in_range(d->nodelist, 8 * E, 0, (int32_t)N)) { why = "nodelist index out of range"; return LGPU_EINVAL; }
   const int32_t *nb[6] = {d->lxim, d->lxip, d->letam, d->letap, d->lzetam, d->lzetap};
   for (const int32_t *a : nb)
      if (!in_range(a, E, 0, (int32_t)E)) { why = "face neighbour out of range"; return LGPU_EINVAL; }
   if (d->nodeElemStart[0] != 0) { why = "nodeElemStart[0] != 0"; return LGPU_EINVAL; }
   for (int64_t n = 0; n < N; ++n)
      if (d->nodeElemStart[n + 1] < d->nodeElemStart[n]) { why = "nodeElemStart not monotone"; return LGPU_EINVAL; }
   const int64_t nc = d->nodeElemStart[N];
   if (nc > 8 * E) { why = "corner list too long"; return LGPU_EINVAL; }
   if (!in_range(d->nodeElemCornerList, nc, 0, (int32_t)(8 * E))) { why = "corner index out of range"; return LGPU_EINVAL; }
   const int32_t *sy[3] = {d->symmX, d->symmY, d->symmZ};
   const int32_t ns[3] = {d->numSymmX, d->numSymmY, d->numSymmZ};
   for (int a = 0; a < 3; ++a) {
      if (ns[a] < 0 || (ns[a] > 0 && !sy[a])) { why = "symmetry set"; return LGPU_EINVAL; }
      if (ns[a] > 0 && !in_range(sy[a], ns[a], 0, (int32_t)N)) { why = "symmetry node out of range"; return LGPU_EINVAL; }
   }
   // regions must partition the elements
   std::vector<uint8_t> seen((size_t)E, 0);
   int64_t total = 0;
   for (int r = 0; r < d->numReg; ++r) {
      const int32_t len = d->regElemSize[r];
      if (len < 0 || (len > 0 && !d->regElemlist[r])) { why = "region list"; return LGPU_EINVAL; }
      for (int32_t i = 0; i < len; ++i) {
         const int32_t e = d->regElemlist[r][i];
         if (e < 0 || e >= E || seen[(size_t)e]) { why = "regions do not partition the elements"; return LGPU_EINVAL; }
         seen[(size_t)e] = 1;
      }
      total += len;
   }
   if (total != E) { why = "regions do not cover every element"; return LGPU_EINVAL; }
   return LGPU_OK;
}

// EOS repeat count of region r, lulesh.cc:2390-2400.
int32_t region_rep(int32_t r, int32_t numReg, int32_t cost)
{
   if (r < numReg / 2) return 1;
   if (r < numReg - (numReg + 15) / 20) return 1 + cost;
   return 10 * (1 + cost);
}

// k_eos dispatch order.  k = 0: work-list order (largest repeat count
// first).  k > 0: the workgroups of the largest repeat count are spread out,
// one every k + 1 dispatches, with the others (in list order) between them,
// so that compute-bound and memory-bound workgroups share the CUs.  The
// dispatcher deals workgroups to the 8 XCDs in turn, so an even period would
// put every heavy workgroup on the same few XCDs (measured: period 4 nearly
// doubles the EOS time); an even k + 1 is raised to the next odd period.
int set_eos_order(lgpu_ctx *c, int k)
{
   if (k > 0 && (k + 1) % 2 == 0) ++k;
   const size_t nb = c->eosBlockRep.size();
   std::vector<int32_t> order;
   order.reserve(nb);
   if (k <= 0) {
      for (size_t b = 0; b < nb; ++b) order.push_back((int32_t)b);
   } else {
      const int32_t top = *std::max_element(c->eosBlockRep.begin(), c->eosBlockRep.end());
      std::vector<int32_t> heavy, rest;
      for (size_t b = 0; b < nb; ++b) (c->eosBlockRep[b] == top ? heavy : rest).push_back((int32_t)b);
      // The FP64 forms walk each group from the top of the mesh down: the
      // kinematics pass before k_eos sweeps the mesh bottom to top, so its
      // last outputs are still in the caches when k_eos starts (s=192, 21
      // regions: k_eos 367 -> 357 us; the mixed form gains nothing and its
      // next nodal pass loses 2 us, so it keeps the ascending walk).
      if (c->precision != LGPU_MIXED) {
         std::reverse(heavy.begin(), heavy.end());
         std::reverse(rest.begin(), rest.end());
      }
      size_t h = 0, r = 0;
      while (h < heavy.size() || r < rest.size()) {
         if (h < heavy.size()) order.push_back(heavy[h++]);
         for (int j = 0; j < k && r < rest.size(); ++j) order.push_back(rest[r++]);
      }
   }
   const std::vector<EosSlot> slots = eos_slots(order, c->hostRegStart, c->hostRegRep, c->numElem);
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   HIP_OR_RETURN(c, hipMemcpy(const_cast<EosSlot *>(c->s.eosSlot), slots.data(),
                              sizeof(EosSlot) * nb, hipMemcpyHostToDevice));
   c->eosInterleave = std::max(0, k);
   return LGPU_OK;
}

// Device array of a field.  `narrow`: stored as FP32 in the mixed form (the
// Q gradients, p, q and ss).
double *real_field_ptr(lgpu_ctx *c, int f, bool &isElem, bool &narrow)
{
   isElem = f >= LGPU_E;
   narrow = c->precision == LGPU_MIXED &&
            (f >= LGPU_DELX_XI || f == LGPU_P || f == LGPU_Q || f == LGPU_SS);
   DevState &s = c->s;
   switch (f) {
   case LGPU_X: return s.x;
   case LGPU_Y: return s.y;
   case LGPU_Z: return s.z;
   case LGPU_XD: return s.xd;
   case LGPU_YD: return s.yd;
   case LGPU_ZD: return s.zd;
   case LGPU_XDD: return s.xdd;
   case LGPU_YDD: return s.ydd;
   case LGPU_ZDD: return s.zdd;
   case LGPU_FX: return s.fx;
   case LGPU_FY: return s.fy;
   case LGPU_FZ: return s.fz;
   case LGPU_NODALMASS: return s.nodalMass;
   case LGPU_E: return s.e;
   case LGPU_P: return static_cast<double *>(s.p);
   case LGPU_Q: return static_cast<double *>(s.q);
   case LGPU_QL: return s.ql;
   case LGPU_QQ: return s.qq;
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

int enable_diagnostics(lgpu_ctx *c)
{
   if (c->diag) return LGPU_OK;
   const size_t nn = align_up(sizeof(double) * (size_t)c->numNode);
   const size_t ne = align_up(sizeof(double) * (size_t)c->numElem);
   char *p = nullptr;
   HIP_OR_RETURN(c, hipMalloc((void **)&p, 6 * nn + 2 * ne));
   HIP_OR_RETURN(c, hipMemset(p, 0, 6 * nn + 2 * ne));
   c->diag = reinterpret_cast<double *>(p);
   double **node[6] = {&c->s.fx, &c->s.fy, &c->s.fz, &c->s.xdd, &c->s.ydd, &c->s.zdd};
   for (int i = 0; i < 6; ++i) *node[i] = reinterpret_cast<double *>(p + i * nn);
   c->s.ql = reinterpret_cast<double *>(p + 6 * nn);
   c->s.qq = reinterpret_cast<double *>(p + 6 * nn + ne);
   c->s.delvOut = c->s.delv;
   return LGPU_OK;
}

template <class P>
int launch_nodal(lgpu_ctx *c, double dt)
{
   if (c->s.isBrick) {
      const int r = timed(c, kKNodalBrick, [&] { Launch<P>::nodal_brick(c->s, c->k, dt, c->stream); });
      if (r) return r;
      // The new node state is in the second buffer: make it current.
      DevState &s = c->s;
      std::swap(s.x, s.xo);
      std::swap(s.y, s.yo);
      std::swap(s.z, s.zo);
      std::swap(s.xd, s.xdo);
      std::swap(s.yd, s.ydo);
      std::swap(s.zd, s.zdo);
      return LGPU_OK;
   }
   int r = timed(c, kKForce, [&] { Launch<P>::force(c->s, c->k, c->planeStride, c->block, c->stream); });
   if (r) return r;
   return timed(c, kKNode, [&] { Launch<P>::node(c->s, c->k, c->planeStride, dt, c->block, c->stream); });
}

template <class P>
int launch_elements(lgpu_ctx *c, double dt)
{
   int r = c->s.isBrick
              ? timed(c, kKKinBrick, [&] { Launch<P>::kin_brick(c->s, c->k, dt, c->stream); })
              : timed(c, kKKin, [&] { Launch<P>::kin(c->s, c->k, dt, c->block, c->stream); });
   if (r) return r;
   r = timed(c, kKEos, [&] { Launch<P>::eos(c->s, c->k, c->stream); });
   if (r) return r;
   return timed(c, kKDtReduce, [&] { Launch<P>::dt_reduce(c->s, c->stream); });
}

int resolve_status(const int32_t *err)
{
   if (err[kErrForceDeterm] || err[kErrHourglassV] || err[kErrKinVnew]) return LGPU_VOLUME_ERROR;
   if (err[kErrQStop]) return LGPU_QSTOP_ERROR;
   if (err[kErrMaterialV]) return LGPU_VOLUME_ERROR;
   return LGPU_OK;
}

// TimeIncrement for the first cycle of a device-side loop; later cycles'
// increments are folded into the dt reduction, k_dt_slots (lgpu_time.h).
__global__ void k_time_increment(DevTime *t, const double *dtOut, const int32_t *err)
{
   if (threadIdx.x != 0 || blockIdx.x != 0) return;
   time_increment(t, dtOut[0], dtOut[1], err);
}

// Folds the last cycle's dt results into the time state once the loop ends.
__global__ void k_time_final(DevTime *t, const double *dtOut)
{
   if (threadIdx.x != 0 || blockIdx.x != 0) return;
   if (t->cycle > 0) {
      t->dtcourant = dtOut[0];
      t->dthydro = dtOut[1];
   }
}

// What the context's arena and kernel path depend on.
struct CtxSpec {
   int32_t numElem = 0, numNode = 0, sizeX = 0, numReg = 0, cost = 0, precision = 0;
   bool brick = false, brickFaces = false;
   int64_t nCorner = 0; // length of the node-to-corner list
   lgpu_constants k{};
};

// Builds the context: stream, one zeroed arena with every array of the step,
// the brick tiling.  Tables and state are filled by the caller.
int alloc_context(const CtxSpec &sp, lgpu_ctx **out)
{
   lgpu_ctx *c = new (std::nothrow) lgpu_ctx();
   if (!c) return LGPU_ENOMEM;
   auto fail = [&](int code) {
      lgpu_destroy(c);
      return code;
   };
   int dev = 0;
   if (hipGetDevice(&dev) != hipSuccess) return fail(LGPU_EDEVICE);
   c->device = dev;
   c->numElem = sp.numElem;
   c->numNode = sp.numNode;
   c->numReg = sp.numReg;
   c->cost = sp.cost;
   c->precision = sp.precision;
   c->sizeX = sp.sizeX;
   const lgpu_constants &kk = sp.k;
   c->k = Consts{kk.e_cut, kk.p_cut, kk.q_cut, kk.v_cut, kk.u_cut, kk.hgcoef, kk.ss4o3,
                 kk.qstop, kk.monoq_max_slope, kk.monoq_limiter_mult, kk.qlc_monoq,
                 kk.qqc_monoq, kk.qqc, kk.eosvmax, kk.eosvmin, kk.pmin, kk.emin, kk.dvovmax,
                 kk.refdens};
   c->k.fastEos = fast_eos_consts(c->k);
   if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
      return fail(LGPU_EDEVICE);

   const bool brick = sp.brick;
   const size_t E = (size_t)sp.numElem, N = (size_t)sp.numNode;
   const size_t gradBytes = (sp.precision == LGPU_MIXED) ? 4 : 8;
   const size_t cornerBytes = (sp.precision == LGPU_MIXED) ? 4 : 8;
   const int64_t plane = (int64_t)(align_up(E * cornerBytes) / cornerBytes);
   c->planeStride = plane;
   const int64_t nCorner = sp.nCorner;
   const unsigned eosBlocks = eos_blocks(sp.numElem);

   // Arena layout.
   struct Slot { void **dst; size_t bytes; };
   std::vector<Slot> slots;
   DevState &s = c->s;
   auto add = [&](void **dst, size_t bytes) { slots.push_back({dst, align_up(bytes)}); };
   // node arrays: two doubles of padding, which the kinematics kernel's
   // pair copies may read past the last node
   double **nodeReal[] = {&s.x, &s.y, &s.z, &s.xd, &s.yd, &s.zd, &s.nodalMass};
   for (double **p : nodeReal) add((void **)p, sizeof(double) * (N + 2));
   if (brick) {
      double **second[] = {&s.xo, &s.yo, &s.zo, &s.xdo, &s.ydo, &s.zdo};
      for (double **p : second) add((void **)p, sizeof(double) * (N + 2));
   }
   add((void **)&s.nodeBC, N);
   add((void **)&s.nodeElemStart, sizeof(int32_t) * (N + 1));
   add((void **)&s.cornerList, sizeof(int32_t) * (size_t)std::max<int64_t>(nCorner, 1));
   add((void **)&s.nodelist, sizeof(int32_t) * 8 * E);
   const int32_t **nbr[] = {&s.lxim, &s.lxip, &s.letam, &s.letap, &s.lzetam, &s.lzetap, &s.elemBC};
   for (const int32_t **p : nbr) add((void **)p, sizeof(int32_t) * E);
   void **pqsArrays[] = {&s.p, &s.q, &s.ss};
   for (void **p : pqsArrays) add(p, gradBytes * E);
   double **elemReal[] = {&s.e, &s.v, &s.volo, &s.vnew, &s.delv, &s.elemMass, &s.arealg, &s.vdov};
   for (double **p : elemReal) add((void **)p, sizeof(double) * E);
   void **grads[] = {&s.delx_xi, &s.delx_eta, &s.delx_zeta, &s.delv_xi, &s.delv_eta, &s.delv_zeta};
   for (void **p : grads) add(p, gradBytes * E);
   if (!brick) add(&s.corner, cornerBytes * (size_t)plane * 48);
   add((void **)&s.perm, sizeof(int32_t) * E);
   add((void **)&s.eosSlot, sizeof(EosSlot) * eosBlocks);
   add((void **)&s.regStart, sizeof(int32_t) * (size_t)(sp.numReg + 1));
   add((void **)&s.regRep, sizeof(int32_t) * (size_t)sp.numReg);
   add((void **)&s.err, sizeof(int32_t) * kNumErrSites);
   add((void **)&s.dtPartial, sizeof(double) * 2 * kDtSlots);
   add((void **)&s.dtOut, sizeof(double) * 2);
   add((void **)&c->devTime, sizeof(DevTime));
   size_t total = 0;
   for (const Slot &sl : slots) total += sl.bytes;
   if (hipMalloc((void **)&c->arena, total) != hipSuccess) {
      (void)hipGetLastError();
      fprintf(stderr, "lgpu_create: device arena of %zu bytes failed\n", total);
      return fail(LGPU_ENOMEM);
   }
   c->arenaBytes = total;
   if (hipMemset(c->arena, 0, total) != hipSuccess) return fail(LGPU_EDEVICE);
   size_t off = 0;
   for (const Slot &sl : slots) {
      *sl.dst = c->arena + off;
      off += sl.bytes;
   }
   // the dt slot pairs start at the key of 1.0e+20 (k_dt_slots resets them
   // to it after every read): sign bit set on the value's bits
   {
      const double big = 1.0e+20;
      uint64_t key = 0;
      std::memcpy(&key, &big, sizeof key);
      key |= 0x8000000000000000ull;
      const std::vector<uint64_t> init(2 * (size_t)kDtSlots, key);
      if (hipMemcpy(s.dtPartial, init.data(), sizeof(uint64_t) * init.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
         return fail(LGPU_EDEVICE);
   }
   s.numElem = sp.numElem;
   s.numNode = sp.numNode;
   s.numReg = sp.numReg;
   s.numPerm = sp.numElem;
   s.xcdRemap = 1;
   s.isBrick = brick ? 1 : 0;
   s.brickFaces = (brick && sp.brickFaces) ? 1 : 0;
   if (brick) {
      Brick &g = s.brick;
      init_brick(g, sp.sizeX);
      int cus = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
         cus = 0;
      int perCu = 0;
      switch (sp.precision) {
      case LGPU_MIXED: perCu = Launch<PolicyMixed>::nodal_brick_blocks_per_cu(); break;
      case LGPU_FP64_FMA: perCu = Launch<PolicyFp64Fma>::nodal_brick_blocks_per_cu(); break;
      default: perCu = Launch<PolicyFp64>::nodal_brick_blocks_per_cu(); break;
      }
      c->brickPerCu = perCu;
      set_brick_chunk(g, pick_brick_chunk(g, cus * perCu));
      int kinPerCu = 0;
      switch (sp.precision) {
      case LGPU_MIXED: kinPerCu = Launch<PolicyMixed>::kin_brick_blocks_per_cu(); break;
      case LGPU_FP64_FMA: kinPerCu = Launch<PolicyFp64Fma>::kin_brick_blocks_per_cu(); break;
      default: kinPerCu = Launch<PolicyFp64>::kin_brick_blocks_per_cu(); break;
      }
      set_kin_chunk(g, pick_kin_chunk(g, cus * kinPerCu));
   }
   *out = c;
   return LGPU_OK;
}

// The EOS work-list groups: one per distinct repeat count, largest first,
// with their sizes; the elements of group i are perm[regStart[i] ..
// regStart[i+1]).  Uploads regStart/regRep and derives the repeat count of
// each k_eos block (the largest over the groups it holds).
int set_groups(lgpu_ctx *c, const std::vector<int32_t> &reps, const std::vector<int64_t> &sizes)
{
   DevState &s = c->s;
   const int numGroups = std::max<int>(1, (int)reps.size());
   std::vector<int32_t> regStart((size_t)numGroups + 1, 0), regRep((size_t)numGroups, 1);
   int64_t at = 0;
   for (size_t k = 0; k < reps.size(); ++k) {
      regStart[k] = (int32_t)at;
      regRep[k] = reps[k];
      at += sizes[k];
   }
   regStart[(size_t)numGroups] = (int32_t)at;
   if (at != c->numElem) return LGPU_EINVAL;
   s.numReg = numGroups;
   const unsigned eosBlocks = eos_blocks(c->numElem);
   c->eosBlockRep.assign(eosBlocks, 1);
   for (int i = 0; i < numGroups; ++i)
      for (int64_t p = regStart[(size_t)i]; p < regStart[(size_t)i + 1]; p += kEosBlock) {
         const size_t b = (size_t)(p / kEosBlock);
         c->eosBlockRep[b] = std::max(c->eosBlockRep[b], regRep[(size_t)i]);
      }
   for (int i = 0; i < numGroups; ++i)
      if (regStart[(size_t)i + 1] > regStart[(size_t)i]) {
         const size_t b = (size_t)((regStart[(size_t)i + 1] - 1) / kEosBlock);
         c->eosBlockRep[b] = std::max(c->eosBlockRep[b], regRep[(size_t)i]);
      }
   c->hostRegStart = regStart;
   c->hostRegRep = regRep;
   HIP_OR_RETURN(c, hipMemcpy(const_cast<int32_t *>(s.regStart), regStart.data(),
                              sizeof(int32_t) * regStart.size(), hipMemcpyHostToDevice));
   HIP_OR_RETURN(c, hipMemcpy(const_cast<int32_t *>(s.regRep), regRep.data(),
                              sizeof(int32_t) * regRep.size(), hipMemcpyHostToDevice));
   return LGPU_OK;
}

// Last steps of both constructors: the EOS dispatch order and the pinned
// read-back buffers.
int finish_context(lgpu_ctx *c)
{
   // One workgroup of the largest repeat count every 5 dispatches: measured
   // 7-10% faster than plain work-list order at s=192 (FP64, 11 and 21
   // regions).
   const int st = set_eos_order(c, 4);
   if (st != LGPU_OK) return st;
   if (hipHostMalloc((void **)&c->hostDt, sizeof(double) * 2 + sizeof(int32_t) * kNumErrSites +
                                               sizeof(DevTime) + 64, hipHostMallocDefault) != hipSuccess)
      return LGPU_ENOMEM;
   c->hostErr = reinterpret_cast<int32_t *>(c->hostDt + 2);
   c->hostTime = reinterpret_cast<DevTime *>(reinterpret_cast<char *>(c->hostDt) + 64);
   return LGPU_OK;
}

// The EOS's FP64 reciprocal and square root on caller data (lgpu_fp64_routines).
__global__ void __launch_bounds__(256)
k_fp64_routines(const double *__restrict__ x, int64_t n, double *__restrict__ rc,
                double *__restrict__ rt)
{
   const int64_t stride = (int64_t)gridDim.x * 256;
   for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
      const double v = x[i];
      if (rc) rc[i] = recip_rn_normal(v);
      if (rt) rt[i] = sqrt_rn_normal(v);
   }
}

} // namespace

extern "C" {

int lgpu_fp64_routines(const double *x, int64_t n, double *recip, double *root, void *stream)
{
   if (!x || n < 0) return LGPU_EINVAL;
   if (n == 0) return LGPU_OK;
   const int64_t want = (n + 255) / 256;
   const int grid = (int)(want < 4096 ? want : 4096);
   hipLaunchKernelGGL(k_fp64_routines, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream),
                      x, n, recip, root);
   if (hipGetLastError() != hipSuccess) return LGPU_EDEVICE;
   return LGPU_OK;
}

void lgpu_default_constants(lgpu_constants *k)
{
   if (!k) return;
   k->e_cut = 1.0e-7;
   k->p_cut = 1.0e-7;
   k->q_cut = 1.0e-7;
   k->v_cut = 1.0e-10;
   k->u_cut = 1.0e-7;
   k->hgcoef = 3.0;
   k->ss4o3 = 4.0 / 3.0;
   k->qstop = 1.0e+12;
   k->monoq_max_slope = 1.0;
   k->monoq_limiter_mult = 2.0;
   k->qlc_monoq = 0.5;
   k->qqc_monoq = 2.0 / 3.0;
   k->qqc = 2.0;
   k->eosvmax = 1.0e+9;
   k->eosvmin = 1.0e-9;
   k->pmin = 0.0;
   k->emin = -1.0e+15;
   k->dvovmax = 0.1;
   k->refdens = 1.0;
}

int lgpu_create(const lgpu_mesh_desc *d, lgpu_ctx **out)
{
   if (!out) return LGPU_EINVAL;
   *out = nullptr;
   std::string why;
   if (validate(d, why) != LGPU_OK) {
      fprintf(stderr, "lgpu_create: invalid mesh description: %s\n", why.c_str());
      return LGPU_EINVAL;
   }
   CtxSpec sp;
   sp.numElem = d->numElem;
   sp.numNode = d->numNode;
   sp.sizeX = d->sizeX;
   sp.numReg = d->numReg;
   sp.cost = d->cost;
   sp.precision = d->precision;
   sp.brick = !(d->flags & LGPU_FLAG_GENERIC) && is_reference_brick(d);
   sp.brickFaces = sp.brick && brick_faces_computable(d);
   sp.nCorner = d->nodeElemStart[d->numNode];
   sp.k = d->k;
   lgpu_ctx *c = nullptr;
   int st = alloc_context(sp, &c);
   if (st != LGPU_OK) return st;
   auto fail = [&](int code) {
      lgpu_destroy(c);
      return code;
   };
   DevState &s = c->s;
   const size_t E = (size_t)d->numElem, N = (size_t)d->numNode;
   const int64_t nCorner = sp.nCorner;

   // Host-built tables.
   std::vector<uint8_t> bc(N, 0);
   const int32_t *sy[3] = {d->symmX, d->symmY, d->symmZ};
   const int32_t ns[3] = {d->numSymmX, d->numSymmY, d->numSymmZ};
   for (int a = 0; a < 3; ++a)
      for (int32_t i = 0; i < ns[a]; ++i) bc[(size_t)sy[a][i]] |= (uint8_t)(1u << a);
   // The EOS work list: elements grouped by their region's EOS repeat count,
   // largest first, and in mesh order within a group.  An element's EOS
   // depends only on the element and its repeat count (lulesh.cc:2387-2401
   // runs the same code for every region), so merging the regions of equal
   // count into one ascending list is exact; it turns the regions' short runs
   // (~38 elements on average) into long ones, so k_eos reads whole cache
   // lines.  regStart/regRep then describe the groups (k_eos reads a
   // block's counts from its dispatch slot, built from them).
   std::vector<int32_t> perm;
   perm.reserve(E);
   std::vector<int32_t> repOf(E, 1);
   std::vector<int32_t> reps;
   for (int r = 0; r < d->numReg; ++r) {
      const int32_t rep = region_rep(r, d->numReg, d->cost);
      if (d->regElemSize[r] > 0 && std::find(reps.begin(), reps.end(), rep) == reps.end())
         reps.push_back(rep);
      for (int32_t i = 0; i < d->regElemSize[r]; ++i) repOf[(size_t)d->regElemlist[r][i]] = rep;
   }
   std::sort(reps.begin(), reps.end(), std::greater<int32_t>());
   std::vector<int64_t> sizes;
   for (int32_t rep : reps) {
      const size_t before = perm.size();
      for (size_t el = 0; el < E; ++el)
         if (repOf[el] == rep) perm.push_back((int32_t)el);
      sizes.push_back((int64_t)(perm.size() - before));
   }
   std::vector<double> ones(E, 1.0);

   auto h2d = [&](const void *dst, const void *src, size_t bytes) {
      return hipMemcpy(const_cast<void *>(dst), src, bytes, hipMemcpyHostToDevice) == hipSuccess;
   };
   bool ok = h2d(s.nodeBC, bc.data(), N) &&
             h2d(s.nodeElemStart, d->nodeElemStart, sizeof(int32_t) * (N + 1)) &&
             (nCorner == 0 || h2d(s.cornerList, d->nodeElemCornerList, sizeof(int32_t) * (size_t)nCorner)) &&
             h2d(s.nodelist, d->nodelist, sizeof(int32_t) * 8 * E) &&
             h2d(s.lxim, d->lxim, sizeof(int32_t) * E) && h2d(s.lxip, d->lxip, sizeof(int32_t) * E) &&
             h2d(s.letam, d->letam, sizeof(int32_t) * E) && h2d(s.letap, d->letap, sizeof(int32_t) * E) &&
             h2d(s.lzetam, d->lzetam, sizeof(int32_t) * E) && h2d(s.lzetap, d->lzetap, sizeof(int32_t) * E) &&
             h2d(s.elemBC, d->elemBC, sizeof(int32_t) * E) &&
             h2d(s.perm, perm.data(), sizeof(int32_t) * E) &&
             h2d(s.v, ones.data(), sizeof(double) * E);
   if (!ok) return fail(LGPU_EDEVICE);
   if ((st = set_groups(c, reps, sizes)) != LGPU_OK) return fail(st);
   if ((st = finish_context(c)) != LGPU_OK) return fail(st);
   *out = c;
   return LGPU_OK;
}

int lgpu_create_sedov(const lgpu_sedov_desc *d, lgpu_ctx **out, lgpu_time_state *ts)
{
   if (!out) return LGPU_EINVAL;
   *out = nullptr;
   auto invalid = [](const char *why) {
      fprintf(stderr, "lgpu_create_sedov: invalid description: %s\n", why);
      return LGPU_EINVAL;
   };
   if (!d || !ts) return invalid("null argument");
   // 8 s^3 corner indices must fit int32 (s <= 645); the bound is tested
   // before any product is formed
   if (d->size < 1 || d->size > 645) return invalid("size out of range");
   const int64_t sz = d->size, E = sz * sz * sz, N = (sz + 1) * (sz + 1) * (sz + 1);
   if (d->precision != LGPU_FP64 && d->precision != LGPU_MIXED && d->precision != LGPU_FP64_FMA)
      return invalid("precision");
   if (d->numReg < 1 || d->cost < 0 || d->balance < 0) return invalid("numReg/balance/cost");
   if (!(d->emult > 0.0)) return invalid("emult");
   if (d->numRuns < 1 || !d->runReg || !d->runLen) return invalid("region runs");
   // Runs: region of each, start of each, and the groups of equal repeat count.
   std::vector<int32_t> runStart((size_t)d->numRuns + 1, 0);
   std::vector<int64_t> regSize((size_t)d->numReg, 0);
   int64_t at = 0;
   for (int32_t r = 0; r < d->numRuns; ++r) {
      if (d->runReg[r] < 1 || d->runReg[r] > d->numReg || d->runLen[r] < 1)
         return invalid("region run");
      runStart[(size_t)r] = (int32_t)at;
      at += d->runLen[r];
      if (at > E) return invalid("region runs exceed the element count");
      regSize[(size_t)(d->runReg[r] - 1)] += d->runLen[r];
   }
   if (at != E) return invalid("region runs do not cover every element");
   runStart[(size_t)d->numRuns] = (int32_t)E;
   std::vector<int32_t> reps;
   for (int r = 0; r < d->numReg; ++r) {
      const int32_t rep = region_rep(r, d->numReg, d->cost);
      if (regSize[(size_t)r] > 0 && std::find(reps.begin(), reps.end(), rep) == reps.end())
         reps.push_back(rep);
   }
   std::sort(reps.begin(), reps.end(), std::greater<int32_t>());
   std::vector<int64_t> groupSize(reps.size(), 0), groupAt(reps.size(), 0);
   auto group_of = [&](int32_t reg) {
      const int32_t rep = region_rep(reg - 1, d->numReg, d->cost);
      return (size_t)(std::find(reps.begin(), reps.end(), rep) - reps.begin());
   };
   for (int r = 0; r < d->numReg; ++r)
      if (regSize[(size_t)r] > 0) groupSize[group_of(r + 1)] += regSize[(size_t)r];
   for (size_t g = 1; g < reps.size(); ++g) groupAt[g] = groupAt[g - 1] + groupSize[g - 1];
   std::vector<int32_t> runDest((size_t)d->numRuns);
   for (int32_t r = 0; r < d->numRuns; ++r) {
      const size_t g = group_of(d->runReg[r]);
      runDest[(size_t)r] = (int32_t)groupAt[g];
      groupAt[g] += d->runLen[r];
   }

   CtxSpec sp;
   sp.numElem = (int32_t)E;
   sp.numNode = (int32_t)N;
   sp.sizeX = (int32_t)sz;
   sp.numReg = d->numReg;
   sp.cost = d->cost;
   sp.precision = d->precision;
   sp.brick = !(d->flags & LGPU_FLAG_GENERIC);
   sp.brickFaces = sp.brick;
   sp.nCorner = 8 * E;
   sp.k = d->k;
   lgpu_ctx *c = nullptr;
   int st = alloc_context(sp, &c);
   if (st != LGPU_OK) return st;
   int32_t *runTables = nullptr;
   auto fail = [&](int code) {
      if (runTables) (void)hipFree(runTables);
      lgpu_destroy(c);
      return code;
   };
   const size_t runBytes = sizeof(int32_t) * ((size_t)d->numRuns + 1);
   if (hipMalloc((void **)&runTables, 2 * runBytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(LGPU_ENOMEM);
   }
   int32_t *dStart = runTables, *dDest = runTables + d->numRuns + 1;
   if (hipMemcpy(dStart, runStart.data(), runBytes, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(dDest, runDest.data(), sizeof(int32_t) * (size_t)d->numRuns,
                 hipMemcpyHostToDevice) != hipSuccess)
      return fail(LGPU_EDEVICE);
   SedovSetup t;
   t.size = (int32_t)sz;
   t.runStart = dStart;
   t.runDest = dDest;
   t.numRuns = d->numRuns;
   if (launch_sedov_setup(c->s, t, c->stream) != hipSuccess) return fail(LGPU_EDEVICE);
   if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(LGPU_EDEVICE);
   (void)hipFree(runTables);
   runTables = nullptr;
   if ((st = set_groups(c, reps, groupSize)) != LGPU_OK) return fail(st);
   if ((st = finish_context(c)) != LGPU_OK) return fail(st);

   // Energy deposit and first dt (lulesh-init.cc:180-192).
   const double ebase = 3.948746e+7;
   const double scale = (double)sz / 45.0;
   double einit = ebase * scale * scale * scale;
   if (d->emult != 1.0) einit = einit * d->emult;
   double volo0 = 0.0;
   if (hipMemcpy(c->s.e, &einit, sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(&volo0, c->s.volo, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(LGPU_EDEVICE);
   // Time controls, lulesh-init.cc:146-156.
   ts->time = 0.0;
   ts->deltatime = (0.5 * cbrt(volo0)) / sqrt(2.0 * einit);
   ts->dtcourant = 1.0e+20;
   ts->dthydro = 1.0e+20;
   ts->stoptime = 1.0e-2;
   ts->dtfixed = -1.0e-6;
   ts->dtmax = 1.0e-2;
   ts->deltatimemultlb = 1.1;
   ts->deltatimemultub = 1.2;
   ts->cycle = 0;
   *out = c;
   return LGPU_OK;
}

void lgpu_destroy(lgpu_ctx *c)
{
   if (!c) return;
   if (c->stream) (void)hipStreamSynchronize(c->stream);
   for (TimedLaunch &t : c->pending) {
      (void)hipEventDestroy(t.start);
      (void)hipEventDestroy(t.stop);
   }
   for (hipEvent_t e : c->eventPool) (void)hipEventDestroy(e);
   if (c->diag) (void)hipFree(c->diag);
   if (c->arena) (void)hipFree(c->arena);
   if (c->hostDt) (void)hipHostFree(c->hostDt);
   if (c->stream) (void)hipStreamDestroy(c->stream);
   delete c;
}

int lgpu_upload(lgpu_ctx *c, int field, const double *host)
{
   if (!c || !host || field < 0 || field >= LGPU_NUM_FIELDS) return LGPU_EINVAL;
   bool isElem, narrow;
   if ((field >= LGPU_XDD && field <= LGPU_FZ) || field == LGPU_QL || field == LGPU_QQ) {
      int r = enable_diagnostics(c);
      if (r) return r;
   }
   double *dst = real_field_ptr(c, field, isElem, narrow);
   if (!dst) return LGPU_EINVAL;
   const size_t n = isElem ? (size_t)c->numElem : (size_t)c->numNode;
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   if (narrow) {
      std::vector<float> tmp(n);
      for (size_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
      HIP_OR_RETURN(c, hipMemcpy(dst, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice));
   } else {
      HIP_OR_RETURN(c, hipMemcpy(dst, host, sizeof(double) * n, hipMemcpyHostToDevice));
   }
   return LGPU_OK;
}

int lgpu_download(lgpu_ctx *c, int field, double *host)
{
   if (!c || !host || field < 0 || field >= LGPU_NUM_FIELDS) return LGPU_EINVAL;
   bool isElem, narrow;
   double *src = real_field_ptr(c, field, isElem, narrow);
   const size_t n = isElem ? (size_t)c->numElem : (size_t)c->numNode;
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   if (!src) {
      std::fill(host, host + n, 0.0);
      return LGPU_OK;
   }
   if (narrow) {
      std::vector<float> tmp(n);
      HIP_OR_RETURN(c, hipMemcpy(tmp.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) host[i] = (double)tmp[i];
   } else {
      HIP_OR_RETURN(c, hipMemcpy(host, src, sizeof(double) * n, hipMemcpyDeviceToHost));
   }
   return LGPU_OK;
}

int lgpu_lagrange_nodal(lgpu_ctx *c, double dt)
{
   if (!c) return LGPU_EINVAL;
   switch (c->precision) {
   case LGPU_MIXED: return launch_nodal<PolicyMixed>(c, dt);
   case LGPU_FP64_FMA: return launch_nodal<PolicyFp64Fma>(c, dt);
   default: return launch_nodal<PolicyFp64>(c, dt);
   }
}

int lgpu_lagrange_elements(lgpu_ctx *c, double dt)
{
   if (!c) return LGPU_EINVAL;
   switch (c->precision) {
   case LGPU_MIXED: return launch_elements<PolicyMixed>(c, dt);
   case LGPU_FP64_FMA: return launch_elements<PolicyFp64Fma>(c, dt);
   default: return launch_elements<PolicyFp64>(c, dt);
   }
}

int lgpu_time_constraints(lgpu_ctx *c, double *dtcourant, double *dthydro)
{
   if (!c) return LGPU_EINVAL;
   HIP_OR_RETURN(c, hipMemcpyAsync(c->hostDt, c->s.dtOut, sizeof(double) * 2,
                                   hipMemcpyDeviceToHost, c->stream));
   HIP_OR_RETURN(c, hipMemcpyAsync(c->hostErr, c->s.err, sizeof(int32_t) * kNumErrSites,
                                   hipMemcpyDeviceToHost, c->stream));
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   if (c->timing) harvest_timing(c);
   const int st = resolve_status(c->hostErr);
   if (st == LGPU_OK) {
      if (dtcourant) *dtcourant = c->hostDt[0];
      if (dthydro) *dthydro = c->hostDt[1];
   }
   return st;
}

int lgpu_run(lgpu_ctx *c, lgpu_time_state *ts, int32_t maxCycle)
{
   if (!c || !ts) return LGPU_EINVAL;
   DevTime *h = c->hostTime;
   h->time = ts->time;
   h->deltatime = ts->deltatime;
   h->dtcourant = ts->dtcourant;
   h->dthydro = ts->dthydro;
   h->stoptime = ts->stoptime;
   h->dtfixed = ts->dtfixed;
   h->dtmax = ts->dtmax;
   h->multlb = ts->deltatimemultlb;
   h->multub = ts->deltatimemultub;
   h->cycle = ts->cycle;
   h->maxCycle = maxCycle;
   h->active = 1;
   h->pad = 0;
   HIP_OR_RETURN(c, hipMemcpyAsync(c->devTime, h, sizeof(DevTime), hipMemcpyHostToDevice, c->stream));
   // dtOut feeds the first TimeIncrement only when cycle > 0: seed it with
   // the caller's constraints.
   c->hostDt[0] = ts->dtcourant;
   c->hostDt[1] = ts->dthydro;
   HIP_OR_RETURN(c, hipMemcpyAsync(c->s.dtOut, c->hostDt, sizeof(double) * 2,
                                   hipMemcpyHostToDevice, c->stream));
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));

   c->s.dtDevice = &c->devTime->deltatime;
   c->s.active = &c->devTime->active;
   c->s.time = c->devTime; // k_dt_slots runs the next cycle's TimeIncrement
   // Cycles are enqueued in batches between host checks of the loop flag.
   // No more cycles are enqueued than the cycle limit allows, so only a run
   // that ends at stoptime (or on an error) enqueues no-op cycles.
   int64_t remaining = (int64_t)maxCycle - (int64_t)ts->cycle;
   int st = LGPU_OK;
   bool first = true;
   for (;;) {
      const int batch = (int)std::max<int64_t>(1, std::min<int64_t>(16, remaining));
      remaining -= batch;
      for (int b = 0; b < batch; ++b) {
         if (first) {
            st = timed(c, kKTimeIncrement, [&] {
               hipLaunchKernelGGL(k_time_increment, dim3(1), dim3(64), 0, c->stream, c->devTime,
                                  c->s.dtOut, c->s.err);
            });
            if (st) break;
            first = false;
         }
         st = lgpu_lagrange_nodal(c, 0.0);
         if (st) break;
         st = lgpu_lagrange_elements(c, 0.0);
         if (st) break;
      }
      if (st) break;
      HIP_OR_RETURN(c, hipMemcpyAsync(h, c->devTime, sizeof(DevTime), hipMemcpyDeviceToHost, c->stream));
      HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
      if (c->timing) harvest_timing(c);
      if (!h->active) break;
   }
   // Back to host-driven dt (the node buffers keep whatever swaps happened).
   c->s.dtDevice = nullptr;
   c->s.active = nullptr;
   c->s.time = nullptr;
   if (st) return st;
   hipLaunchKernelGGL(k_time_final, dim3(1), dim3(64), 0, c->stream, c->devTime, c->s.dtOut);
   HIP_OR_RETURN(c, hipGetLastError());
   HIP_OR_RETURN(c, hipMemcpyAsync(h, c->devTime, sizeof(DevTime), hipMemcpyDeviceToHost, c->stream));
   HIP_OR_RETURN(c, hipMemcpyAsync(c->hostErr, c->s.err, sizeof(int32_t) * kNumErrSites,
                                   hipMemcpyDeviceToHost, c->stream));
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   ts->time = h->time;
   ts->deltatime = h->deltatime;
   ts->dtcourant = h->dtcourant;
   ts->dthydro = h->dthydro;
   ts->cycle = h->cycle;
   return resolve_status(c->hostErr);
}

int lgpu_set_option(lgpu_ctx *c, int option, int64_t value)
{
   if (!c) return LGPU_EINVAL;
   switch (option) {
   case LGPU_OPT_KEEP_DIAGNOSTICS:
      if (value) {
         const int r = enable_diagnostics(c);
         if (r) return r;
      } else {
         c->s.fx = c->s.fy = c->s.fz = c->s.xdd = c->s.ydd = c->s.zdd = nullptr;
         c->s.ql = c->s.qq = c->s.delvOut = nullptr;
         HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
         if (c->diag) (void)hipFree(c->diag);
         c->diag = nullptr;
      }
      c->keepDiag = value != 0;
      return LGPU_OK;
   case LGPU_OPT_KERNEL_TIMING:
      c->timing = value != 0;
      return LGPU_OK;
   case LGPU_OPT_BLOCK_SIZE:
      if (value != 64 && value != 128 && value != 256) return LGPU_EINVAL;
      c->block = (int)value;
      return LGPU_OK;
   case LGPU_OPT_BRICK_CHUNK:
      if (!c->s.isBrick || value < 1) return LGPU_EINVAL;
      set_brick_chunk(c->s.brick, (int)std::min<int64_t>(value, 1 << 20));
      return LGPU_OK;
   case LGPU_OPT_XCD_REMAP:
      c->s.xcdRemap = value ? 1 : 0;
      return LGPU_OK;
   case LGPU_OPT_KIN_CHUNK:
      if (!c->s.isBrick || value < 1) return LGPU_EINVAL;
      set_kin_chunk(c->s.brick, (int)std::min<int64_t>(value, 1 << 20));
      return LGPU_OK;
   case LGPU_OPT_BRICK_ROWS:
      if (!c->s.isBrick || (value != 8 && value != 16)) return LGPU_EINVAL;
      set_brick_tile(c->s.brick, 16, (int)value);
      return LGPU_OK;
   case LGPU_OPT_EOS_INTERLEAVE:
      if (value < 0 || value > 1024) return LGPU_EINVAL;
      return set_eos_order(c, (int)value);
   case LGPU_OPT_KERNEL_TIMING_MASK:
      c->timingMask = (uint32_t)value;
      return LGPU_OK;
   default:
      return LGPU_EINVAL;
   }
}

int lgpu_get_option(lgpu_ctx *c, int option, int64_t *value)
{
   if (!c || !value) return LGPU_EINVAL;
   switch (option) {
   case LGPU_OPT_KEEP_DIAGNOSTICS: *value = c->keepDiag ? 1 : 0; return LGPU_OK;
   case LGPU_OPT_KERNEL_TIMING: *value = c->timing ? 1 : 0; return LGPU_OK;
   case LGPU_OPT_BLOCK_SIZE: *value = c->block; return LGPU_OK;
   case LGPU_OPT_BRICK_CHUNK:
      if (!c->s.isBrick) return LGPU_EINVAL;
      *value = c->s.brick.chunkZ;
      return LGPU_OK;
   case LGPU_OPT_XCD_REMAP: *value = c->s.xcdRemap; return LGPU_OK;
   case LGPU_OPT_KERNEL_TIMING_MASK: *value = c->timingMask; return LGPU_OK;
   case LGPU_OPT_KIN_CHUNK:
      if (!c->s.isBrick) return LGPU_EINVAL;
      *value = c->s.brick.kinChunk;
      return LGPU_OK;
   case LGPU_OPT_BRICK_ROWS:
      if (!c->s.isBrick) return LGPU_EINVAL;
      *value = c->s.brick.rows;
      return LGPU_OK;
   case LGPU_OPT_EOS_INTERLEAVE: *value = c->eosInterleave; return LGPU_OK;
   default: return LGPU_EINVAL;
   }
}

int lgpu_num_kernels(void) { return kNumKernels; }

const char *lgpu_kernel_name(int kernel)
{
   return (kernel >= 0 && kernel < kNumKernels) ? kKernelNames[kernel] : "";
}

int lgpu_kernel_stats(lgpu_ctx *c, int kernel, int64_t *launches, double *total_ms)
{
   if (!c || kernel < 0 || kernel >= kNumKernels) return LGPU_EINVAL;
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   harvest_timing(c);
   if (launches) *launches = c->kLaunches[kernel];
   if (total_ms) *total_ms = c->kMs[kernel];
   return LGPU_OK;
}

int lgpu_reset_kernel_stats(lgpu_ctx *c)
{
   if (!c) return LGPU_EINVAL;
   HIP_OR_RETURN(c, hipStreamSynchronize(c->stream));
   harvest_timing(c);
   for (int i = 0; i < kNumKernels; ++i) {
      c->kLaunches[i] = 0;
      c->kMs[i] = 0.0;
   }
   return LGPU_OK;
}

int lgpu_get_info(lgpu_ctx *c, lgpu_info *info)
{
   if (!c || !info) return LGPU_EINVAL;
   info->path = c->s.isBrick ? 1 : 0;
   info->precision = c->precision;
   info->numElem = c->numElem;
   info->numNode = c->numNode;
   info->numReg = c->numReg;
   info->block = c->block;
   info->arena_bytes = (int64_t)c->arenaBytes;
   info->brick_faces = c->s.brickFaces;
   info->brick_blocks_per_cu = c->brickPerCu;
   return LGPU_OK;
}

const char *lgpu_last_error(lgpu_ctx *c) { return c ? c->lastError.c_str() : ""; }

} // extern "C"
