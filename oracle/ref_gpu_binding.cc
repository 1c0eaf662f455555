// The reference's own program with its three phase functions bound to the
// MI355X library (include/lulesh_gpu.h), built from the reference's sources
// where they lie (oracle/Makefile, target ref-gpu -> _ref/lulesh2.0-gpu).
//
// Test infrastructure: it checks that the binding a maintainer adds to
// lulesh.cc ("Reference-side binding") drives the library
// correctly from the reference's own Domain, main loop (lulesh.cc:2745-2748),
// TimeIncrement (lulesh.cc:166-222), LagrangeLeapFrog (lulesh.cc:2600-2645)
// and end-of-run report (lulesh-util.cc:175-230).  The library itself links
// nothing built from the reference.
//
// The reference's lulesh.cc is included unchanged.  While it is included,
// three function-like macros split each phase function's name by its
// argument list: the definitions, written `LagrangeNodal(Domain& domain)`
// (lulesh.cc:1224, 2432, 2577), become `*_reference` (the CPU bodies, now
// unused), and the calls in LagrangeLeapFrog, written `LagrangeNodal(domain)`
// (lulesh.cc:2609, 2617, 2638), become calls of the `*_bound` functions
// below.  The token after the opening parenthesis (`Domain` or `domain`) is
// pasted onto a prefix that selects the name.  VerifyAndWriteFinalOutput is
// wrapped so that the final state is copied back before the report.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lulesh_gpu.h"

class Domain;
typedef int Index_t;
static void LagrangeNodal_bound(Domain &domain);
static void LagrangeElements_bound(Domain &domain, Index_t numElem);
static void CalcTimeConstraintsForElems_bound(Domain &domain);

#define LagrangeNodal(arg) LAGRANGE_NODAL_##arg)
#define LAGRANGE_NODAL_Domain LagrangeNodal_reference(Domain
#define LAGRANGE_NODAL_domain LagrangeNodal_bound(domain
#define LagrangeElements(arg, n) LAGRANGE_ELEMENTS_##arg, n)
#define LAGRANGE_ELEMENTS_Domain LagrangeElements_reference(Domain
#define LAGRANGE_ELEMENTS_domain LagrangeElements_bound(domain
#define CalcTimeConstraintsForElems(arg) TIME_CONSTRAINTS_##arg)
#define TIME_CONSTRAINTS_Domain CalcTimeConstraintsForElems_reference(Domain
#define TIME_CONSTRAINTS_domain CalcTimeConstraintsForElems_bound(domain
#define VerifyAndWriteFinalOutput VerifyAndWriteFinalOutput_bound
#include "lulesh.cc"
#undef VerifyAndWriteFinalOutput
#undef LagrangeNodal
#undef LagrangeElements
#undef CalcTimeConstraintsForElems

void VerifyAndWriteFinalOutput(Real_t elapsed_time, Domain &locDom, Int_t nx, Int_t numRanks);

// ---------------------------------------------------------------------------
// The binding of INTEGRATION.md.  AttachGpu runs at the first phase call
// (the reference's main builds the Domain and goes straight into its loop).

static lgpu_ctx *g_gpu = nullptr;

static void AttachGpu(Domain &d, int precision)
{
   const Index_t N = d.numNode(), E = d.numElem(), nbc = (d.sizeX() + 1) * (d.sizeX() + 1);
   std::vector<Index_t> start(N + 1, 0), symX(nbc), symY(nbc), symZ(nbc);
   for (Index_t n = 0; n < N; ++n) start[n + 1] = start[n] + d.nodeElemCount(n);
   for (Index_t i = 0; i < nbc; ++i) { symX[i] = d.symmX(i); symY[i] = d.symmY(i); symZ[i] = d.symmZ(i); }
   std::vector<const Index_t *> lists(d.numReg());
   for (Int_t r = 0; r < d.numReg(); ++r) lists[r] = d.regElemlist(r);

   lgpu_mesh_desc m = {};
   m.numElem = E; m.numNode = N; m.sizeX = d.sizeX(); m.sizeY = d.sizeY(); m.sizeZ = d.sizeZ();
   m.nodelist = d.nodelist(0);
   m.lxim = &d.lxim(0); m.lxip = &d.lxip(0); m.letam = &d.letam(0); m.letap = &d.letap(0);
   m.lzetam = &d.lzetam(0); m.lzetap = &d.lzetap(0); m.elemBC = &d.elemBC(0);
   m.nodeElemStart = start.data(); m.nodeElemCornerList = d.nodeElemCornerList(0);
   m.symmX = symX.data(); m.symmY = symY.data(); m.symmZ = symZ.data();
   m.numSymmX = m.numSymmY = m.numSymmZ = nbc;
   m.numReg = d.numReg(); m.regElemSize = &d.regElemSize(0); m.regElemlist = lists.data();
   m.cost = d.cost(); m.precision = precision;
   lgpu_default_constants(&m.k);
   if (lgpu_create(&m, &g_gpu) != LGPU_OK) exit(-1);

   const struct { int f; Real_t &(Domain::*a)(Index_t); } up[] = {
      {LGPU_X, &Domain::x}, {LGPU_Y, &Domain::y}, {LGPU_Z, &Domain::z},
      {LGPU_XD, &Domain::xd}, {LGPU_YD, &Domain::yd}, {LGPU_ZD, &Domain::zd},
      {LGPU_NODALMASS, &Domain::nodalMass}, {LGPU_E, &Domain::e}, {LGPU_P, &Domain::p},
      {LGPU_Q, &Domain::q}, {LGPU_V, &Domain::v}, {LGPU_VOLO, &Domain::volo},
      {LGPU_SS, &Domain::ss}, {LGPU_ELEMMASS, &Domain::elemMass}};
   for (auto &u : up) lgpu_upload(g_gpu, u.f, &(d.*u.a)(0));
}

static void LagrangeNodal_bound(Domain &domain)
{
   if (!g_gpu) AttachGpu(domain, getenv("LGPU_MIXED") ? LGPU_MIXED : LGPU_FP64);
   if (int st = lgpu_lagrange_nodal(g_gpu, domain.deltatime())) exit(st);
}

static void LagrangeElements_bound(Domain &domain, Index_t)
{
   if (int st = lgpu_lagrange_elements(g_gpu, domain.deltatime())) exit(st);
}

static void CalcTimeConstraintsForElems_bound(Domain &domain)
{
   if (int st = lgpu_time_constraints(g_gpu, &domain.dtcourant(), &domain.dthydro())) exit(st);
}

// The final state back into the Domain, the origin energy at full
// precision, then the reference's own report.
void VerifyAndWriteFinalOutput_bound(Real_t elapsed_time, Domain &locDom, Int_t nx, Int_t numRanks)
{
   if (g_gpu) {
      const struct { int f; Real_t &(Domain::*a)(Index_t); } down[] = {
         {LGPU_X, &Domain::x}, {LGPU_Y, &Domain::y}, {LGPU_Z, &Domain::z},
         {LGPU_XD, &Domain::xd}, {LGPU_YD, &Domain::yd}, {LGPU_ZD, &Domain::zd},
         {LGPU_E, &Domain::e}, {LGPU_P, &Domain::p}, {LGPU_Q, &Domain::q},
         {LGPU_V, &Domain::v}, {LGPU_SS, &Domain::ss}};
      for (auto &f : down) lgpu_download(g_gpu, f.f, &(locDom.*f.a)(0));
      lgpu_destroy(g_gpu);
      g_gpu = nullptr;
   }
   printf("Bound run: cycles %d, final origin energy %.17e\n", (int)locDom.cycle(), locDom.e(0));
   fflush(stdout);
   VerifyAndWriteFinalOutput(elapsed_time, locDom, nx, numRanks);
}
