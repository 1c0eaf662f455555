// domain.cc -- set-up of the host Domain, the per-cycle call structure over
// the device path, and the flat C API (include/lulesh_host.h).
#include "host/domain.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <memory>
#include <random>

#include "host/glibc_rand.h"
#include "lulesh_host.h"

namespace {

// elemBC bits, lulesh.h:59-87.
enum : Int_t {
   XI_M_SYMM = 0x00001, XI_P_FREE = 0x00010, ETA_M_SYMM = 0x00040, ETA_P_FREE = 0x00400,
   ZETA_M_SYMM = 0x01000, ZETA_P_FREE = 0x10000
};

} // namespace

namespace lhost {

// The region draws of CreateRegionIndexSets (lulesh-init.cc:401-480) as runs
// of consecutive elements: region numbers (1-based) drawn with weights
// (i+1)^balance and never twice in a row, run lengths from the reference's
// bins, the last run cut at numElem.  One region: a single run.
void draw_region_runs(Index_t numElem, Int_t nr, Int_t balance, std::vector<Int_t> &runReg,
                      std::vector<Int_t> &runLen)
{
   runReg.clear();
   runLen.clear();
   if (nr == 1) {
      runReg.push_back(1);
      runLen.push_back(numElem);
      return;
   }
   lhost::GlibcRand rng(0);
   // Cumulative weights (i+1)^balance.
   std::vector<Int_t> binEnd((size_t)nr);
   Int_t total = 0;
   for (Int_t i = 0; i < nr; ++i) {
      total += pow(i + 1, balance);
      binEnd[(size_t)i] = total;
   }
   auto draw_region = [&]() {
      const Int_t v = rng.next() % total;
      Int_t i = 0;
      while (v >= binEnd[(size_t)i]) ++i;
      return (i % nr) + 1;
   };
   Index_t next = 0;
   Int_t last = -1;
   while (next < numElem) {
      Int_t reg = draw_region();
      while (reg == last) reg = draw_region();
      // Run length by bin (lulesh-init.cc:460-480).
      const Int_t bin = rng.next() % 1000;
      Index_t run;
      if (bin < 773) run = rng.next() % 15 + 1;
      else if (bin < 937) run = rng.next() % 16 + 16;
      else if (bin < 970) run = rng.next() % 32 + 32;
      else if (bin < 974) run = rng.next() % 64 + 64;
      else if (bin < 978) run = rng.next() % 128 + 128;
      else if (bin < 981) run = rng.next() % 256 + 256;
      else run = rng.next() % 1537 + 512;
      const Index_t end = std::min(next + run, numElem);
      runReg.push_back(reg);
      runLen.push_back(end - next);
      next = end;
      last = reg;
   }
}

} // namespace lhost

Domain::Domain(Int_t numRanks, Index_t colLoc, Index_t rowLoc, Index_t planeLoc, Index_t nx,
               Int_t tp, Int_t nr, Int_t balance, Int_t cost)
{
   if (numRanks != 1 || tp != 1 || colLoc != 0 || rowLoc != 0 || planeLoc != 0 || nx < 1 || nr < 1) {
      fprintf(stderr, "Domain: only a single domain (numRanks = tp = 1) is supported\n");
      exit(-1);
   }
   lgpu_default_constants(&m_k);
   const Index_t edgeElems = nx, edgeNodes = nx + 1;
   m_cost = cost;
   m_sizeX = m_sizeY = m_sizeZ = edgeElems;
   m_numElem = edgeElems * edgeElems * edgeElems;
   m_numNode = edgeNodes * edgeNodes * edgeNodes;
   const size_t E = (size_t)m_numElem, N = (size_t)m_numNode;

   for (auto *a : {&m_x, &m_y, &m_z, &m_xd, &m_yd, &m_zd, &m_xdd, &m_ydd, &m_zdd, &m_fx, &m_fy,
                   &m_fz, &m_nodalMass})
      a->assign(N, 0.0);
   for (auto *a : {&m_e, &m_p, &m_q, &m_ql, &m_qq, &m_volo, &m_vnew, &m_delv, &m_vdov, &m_arealg,
                   &m_ss, &m_elemMass, &m_delx_xi, &m_delx_eta, &m_delx_zeta, &m_delv_xi,
                   &m_delv_eta, &m_delv_zeta})
      a->assign(E, 0.0);
   m_v.assign(E, 1.0); // v starts at 1 (lulesh-init.cc:97-100)
   m_nodelist.assign(8 * E, 0);
   for (auto *a : {&m_lxim, &m_lxip, &m_letam, &m_letap, &m_lzetam, &m_lzetap, &m_elemBC})
      a->assign(E, 0);
   m_regNumList.assign(E, 0);
   m_symmX.assign((size_t)edgeNodes * edgeNodes, 0);
   m_symmY.assign((size_t)edgeNodes * edgeNodes, 0);
   m_symmZ.assign((size_t)edgeNodes * edgeNodes, 0);

   BuildMesh(nx, edgeNodes, edgeElems);
   SetupThreadSupportStructures();
   CreateRegionIndexSets(nr, balance);
   SetupSymmetryPlanes(edgeNodes);
   SetupElementConnectivities(edgeElems);
   SetupBoundaryConditions(edgeElems);

   // Time controls, lulesh-init.cc:146-156.
   m_dtfixed = -1.0e-6;
   m_stoptime = 1.0e-2;
   m_deltatimemultlb = 1.1;
   m_deltatimemultub = 1.2;
   m_dtcourant = 1.0e+20;
   m_dthydro = 1.0e+20;
   m_dtmax = 1.0e-2;
   m_time = 0.;
   m_cycle = 0;

   // Reference volume and masses, lulesh-init.cc:158-178.
   for (Index_t i = 0; i < m_numElem; ++i) {
      const Index_t *nl = nodelist(i);
      double xl[8], yl[8], zl[8];
      for (int j = 0; j < 8; ++j) {
         xl[j] = m_x[nl[j]];
         yl[j] = m_y[nl[j]];
         zl[j] = m_z[nl[j]];
      }
      // Hexahedron volume, lulesh.cc:1273-1356.
      auto d = [](const double *c, int a, int b) { return c[a] - c[b]; };
      auto trip = [](double ux, double uy, double uz, double vx, double vy, double vz, double wx,
                     double wy, double wz) {
         return ux * (vy * wz - vz * wy) + uy * (vz * wx - vx * wz) + uz * (vx * wy - vy * wx);
      };
      const double t1 = trip(d(xl, 3, 1) + d(xl, 7, 2), d(yl, 3, 1) + d(yl, 7, 2), d(zl, 3, 1) + d(zl, 7, 2),
                             d(xl, 6, 3), d(yl, 6, 3), d(zl, 6, 3), d(xl, 2, 0), d(yl, 2, 0), d(zl, 2, 0));
      const double t2 = trip(d(xl, 4, 3) + d(xl, 5, 7), d(yl, 4, 3) + d(yl, 5, 7), d(zl, 4, 3) + d(zl, 5, 7),
                             d(xl, 6, 4), d(yl, 6, 4), d(zl, 6, 4), d(xl, 7, 0), d(yl, 7, 0), d(zl, 7, 0));
      const double t3 = trip(d(xl, 1, 4) + d(xl, 2, 5), d(yl, 1, 4) + d(yl, 2, 5), d(zl, 1, 4) + d(zl, 2, 5),
                             d(xl, 6, 1), d(yl, 6, 1), d(zl, 6, 1), d(xl, 5, 0), d(yl, 5, 0), d(zl, 5, 0));
      double volume = t1 + t2 + t3;
      volume *= 1.0 / 12.0;
      m_volo[i] = volume;
      m_elemMass[i] = volume;
      for (int j = 0; j < 8; ++j) m_nodalMass[nl[j]] += volume / 8.0;
   }
   ScaleInitialEnergy(1.0);
}

Domain::~Domain()
{
   if (m_gpu) lgpu_destroy(m_gpu);
}

void Domain::ScaleInitialEnergy(Real_t m)
{
   // Energy deposit and first dt, lulesh-init.cc:180-192.
   const Real_t ebase = 3.948746e+7;
   const Real_t scale = m_sizeX / 45.0;
   Real_t einit = ebase * scale * scale * scale;
   if (m != 1.0) einit = einit * m;
   m_e[0] = einit;
   m_deltatime = (0.5 * cbrt(m_volo[0])) / sqrt(2.0 * einit);
}

// Node lattice and element-to-node map, lulesh-init.cc:217-267.
void Domain::BuildMesh(Int_t nx, Int_t edgeNodes, Int_t edgeElems)
{
   // Coordinate of lattice line i: 1.125 * i / nx (0 for i = 0).
   std::vector<Real_t> c((size_t)edgeNodes);
   for (Index_t i = 0; i < edgeNodes; ++i) c[i] = 1.125 * (double)i / (double)nx;
   Index_t n = 0;
   for (Index_t k = 0; k < edgeNodes; ++k)
      for (Index_t j = 0; j < edgeNodes; ++j)
         for (Index_t i = 0; i < edgeNodes; ++i, ++n) {
            m_x[n] = c[i];
            m_y[n] = c[j];
            m_z[n] = c[k];
         }
   const Index_t en = edgeNodes, en2 = edgeNodes * edgeNodes;
   Index_t e = 0;
   for (Index_t k = 0; k < edgeElems; ++k)
      for (Index_t j = 0; j < edgeElems; ++j)
         for (Index_t i = 0; i < edgeElems; ++i, ++e) {
            const Index_t b = k * en2 + j * en + i;
            Index_t *nl = nodelist(e);
            nl[0] = b;
            nl[1] = b + 1;
            nl[2] = b + en + 1;
            nl[3] = b + en;
            nl[4] = b + en2;
            nl[5] = b + en2 + 1;
            nl[6] = b + en2 + en + 1;
            nl[7] = b + en2 + en;
         }
}

// Node-to-corner lists, lulesh-init.cc:271-337: corners of each node in
// ascending (element, corner) order.  Always built, since the device gather
// uses them.
void Domain::SetupThreadSupportStructures()
{
   const size_t N = (size_t)m_numNode;
   std::vector<Index_t> count(N, 0);
   for (Index_t n : m_nodelist) ++count[(size_t)n];
   m_nodeElemStart.assign(N + 1, 0);
   for (size_t n = 0; n < N; ++n) m_nodeElemStart[n + 1] = m_nodeElemStart[n] + count[n];
   m_nodeElemCornerList.assign((size_t)m_nodeElemStart[N], 0);
   std::fill(count.begin(), count.end(), 0);
   for (size_t c = 0; c < m_nodelist.size(); ++c) {
      const Index_t n = m_nodelist[c];
      m_nodeElemCornerList[(size_t)(m_nodeElemStart[n] + count[(size_t)n]++)] = (Index_t)c;
   }
}

// Region index sets, lulesh-init.cc:400-510 (single rank: seed 0, no
// rotation), drawn from the glibc sequence.
void Domain::CreateRegionIndexSets(Int_t nr, Int_t balance)
{
   m_numReg = nr;
   m_regElemSize.assign((size_t)nr, 0);
   m_regElemlist.assign((size_t)nr, {});
   std::vector<Int_t> runReg, runLen;
   lhost::draw_region_runs(m_numElem, nr, balance, runReg, runLen);
   Index_t next = 0;
   for (size_t r = 0; r < runReg.size(); ++r)
      for (Index_t i = 0; i < runLen[r]; ++i) m_regNumList[(size_t)next++] = runReg[r];
   for (Index_t i = 0; i < m_numElem; ++i) m_regElemSize[(size_t)(m_regNumList[(size_t)i] - 1)]++;
   for (Int_t r = 0; r < nr; ++r) m_regElemlist[(size_t)r].reserve((size_t)m_regElemSize[(size_t)r]);
   for (Index_t i = 0; i < m_numElem; ++i) m_regElemlist[(size_t)(m_regNumList[(size_t)i] - 1)].push_back(i);
}

// Symmetry-plane node sets, lulesh-init.cc:513-533.
void Domain::SetupSymmetryPlanes(Int_t edgeNodes)
{
   Index_t n = 0;
   for (Index_t a = 0; a < edgeNodes; ++a)
      for (Index_t b = 0; b < edgeNodes; ++b, ++n) {
         m_symmZ[(size_t)n] = a * edgeNodes + b;
         m_symmY[(size_t)n] = a * edgeNodes * edgeNodes + b;
         m_symmX[(size_t)n] = a * edgeNodes * edgeNodes + b * edgeNodes;
      }
}

// Face neighbours, lulesh-init.cc:538-565: the neighbouring element across
// each face, or the element itself on the mesh boundary.
void Domain::SetupElementConnectivities(Int_t s)
{
   const Index_t s2 = s * s;
   for (Index_t e = 0; e < m_numElem; ++e) {
      // Like the reference, xi neighbours run along the flat element index
      // (wrapping across rows); the boundary masks keep the wrapped ones
      // from being read.
      m_lxim[(size_t)e] = (e > 0) ? e - 1 : 0;
      m_lxip[(size_t)e] = (e + 1 < m_numElem) ? e + 1 : e;
      m_letam[(size_t)e] = (e >= s) ? e - s : e;
      m_letap[(size_t)e] = (e + s < m_numElem) ? e + s : e;
      m_lzetam[(size_t)e] = (e >= s2) ? e - s2 : e;
      m_lzetap[(size_t)e] = (e + s2 < m_numElem) ? e + s2 : e;
   }
}

// Boundary masks of a single domain, lulesh-init.cc:568-673: symmetry on the
// three planes through the origin, free surface on the other three.
void Domain::SetupBoundaryConditions(Int_t s)
{
   for (Index_t e = 0; e < m_numElem; ++e) {
      const Index_t i = e % s, j = (e / s) % s, k = e / (s * s);
      Int_t bc = 0;
      if (i == 0) bc |= XI_M_SYMM;
      if (i == s - 1) bc |= XI_P_FREE;
      if (j == 0) bc |= ETA_M_SYMM;
      if (j == s - 1) bc |= ETA_P_FREE;
      if (k == 0) bc |= ZETA_M_SYMM;
      if (k == s - 1) bc |= ZETA_P_FREE;
      m_elemBC[(size_t)e] = bc;
   }
}

Real_t *Domain::realField(int f)
{
   std::vector<Real_t> *t[LGPU_NUM_FIELDS] = {
      &m_x, &m_y, &m_z, &m_xd, &m_yd, &m_zd, &m_xdd, &m_ydd, &m_zdd, &m_fx, &m_fy, &m_fz,
      &m_nodalMass, &m_e, &m_p, &m_q, &m_ql, &m_qq, &m_v, &m_volo, &m_vnew, &m_delv, &m_vdov,
      &m_arealg, &m_ss, &m_elemMass, &m_delx_xi, &m_delx_eta, &m_delx_zeta, &m_delv_xi,
      &m_delv_eta, &m_delv_zeta};
   return (f >= 0 && f < LGPU_NUM_FIELDS) ? t[f]->data() : nullptr;
}

namespace {
// Fields that carry state from one cycle to the next (plus the set-up
// constants volo, elemMass, nodalMass).
const int kStateFields[] = {LGPU_X, LGPU_Y, LGPU_Z, LGPU_XD, LGPU_YD, LGPU_ZD, LGPU_NODALMASS,
                            LGPU_E, LGPU_P, LGPU_Q, LGPU_V, LGPU_VOLO, LGPU_SS, LGPU_ELEMMASS};
// What a cycle writes: the fields a sync brings back.  nodalMass, volo and
// elemMass stay the host's FP64 inputs (the mixed form keeps elemMass in FP32
// on the device; its download would be the rounded copy).  delv, like ql and
// qq, is formed and used inside one cycle and not kept.
const int kStepFields[] = {LGPU_X, LGPU_Y, LGPU_Z, LGPU_XD, LGPU_YD, LGPU_ZD, LGPU_E, LGPU_P,
                           LGPU_Q, LGPU_V, LGPU_SS, LGPU_VNEW, LGPU_VDOV, LGPU_AREALG};
} // namespace

int Domain::AttachDevice(int precision, int flags)
{
   if (m_gpu) {
      lgpu_destroy(m_gpu);
      m_gpu = nullptr;
   }
   std::vector<const Index_t *> lists((size_t)m_numReg);
   for (Int_t r = 0; r < m_numReg; ++r) lists[(size_t)r] = m_regElemlist[(size_t)r].data();
   lgpu_mesh_desc d{};
   d.numElem = m_numElem;
   d.numNode = m_numNode;
   d.sizeX = m_sizeX;
   d.sizeY = m_sizeY;
   d.sizeZ = m_sizeZ;
   d.nodelist = m_nodelist.data();
   d.lxim = m_lxim.data();
   d.lxip = m_lxip.data();
   d.letam = m_letam.data();
   d.letap = m_letap.data();
   d.lzetam = m_lzetam.data();
   d.lzetap = m_lzetap.data();
   d.elemBC = m_elemBC.data();
   d.nodeElemStart = m_nodeElemStart.data();
   d.nodeElemCornerList = m_nodeElemCornerList.data();
   d.symmX = m_symmX.data();
   d.symmY = m_symmY.data();
   d.symmZ = m_symmZ.data();
   d.numSymmX = (int32_t)m_symmX.size();
   d.numSymmY = (int32_t)m_symmY.size();
   d.numSymmZ = (int32_t)m_symmZ.size();
   d.numReg = m_numReg;
   d.regElemSize = m_regElemSize.data();
   d.regElemlist = lists.data();
   d.cost = m_cost;
   d.precision = precision;
   d.flags = flags;
   d.k = m_k;
   int st = lgpu_create(&d, &m_gpu);
   if (st) return st;
   for (int f : kStateFields) {
      st = lgpu_upload(m_gpu, f, realField(f));
      if (st) return st;
   }
   return LGPU_OK;
}

int Domain::SyncFromDevice()
{
   if (!m_gpu) return LGPU_EINVAL;
   for (int f : kStepFields) {
      const int st = lgpu_download(m_gpu, f, realField(f));
      if (st) return st;
   }
   return LGPU_OK;
}

// ---------------------------------------------------------------------------
// Per-cycle call structure.

void TimeIncrement(Domain &domain)
{
   // lulesh.cc:166-222 (no MPI reduction: one domain).
   Real_t targetdt = domain.stoptime() - domain.time();
   if ((domain.dtfixed() <= 0.0) && (domain.cycle() != 0)) {
      const Real_t olddt = domain.deltatime();
      Real_t newdt = 1.0e+20;
      if (domain.dtcourant() < newdt) newdt = domain.dtcourant() / 2.0;
      if (domain.dthydro() < newdt) newdt = domain.dthydro() * 2.0 / 3.0;
      const Real_t ratio = newdt / olddt;
      if (ratio >= 1.0) {
         if (ratio < domain.deltatimemultlb()) newdt = olddt;
         else if (ratio > domain.deltatimemultub()) newdt = olddt * domain.deltatimemultub();
      }
      if (newdt > domain.dtmax()) newdt = domain.dtmax();
      domain.deltatime() = newdt;
   }
   if ((targetdt > domain.deltatime()) && (targetdt < (4.0 * domain.deltatime() / 3.0)))
      targetdt = 2.0 * domain.deltatime() / 3.0;
   if (targetdt < domain.deltatime()) domain.deltatime() = targetdt;
   domain.time() += domain.deltatime();
   ++domain.cycle();
}

static void exit_on(int st)
{
   if (st == LGPU_OK) return;
   if (st > -100) exit(st);
   fprintf(stderr, "device path failed (%d)\n", st);
   exit(st);
}

void LagrangeNodal(Domain &domain)
{
   exit_on(lgpu_lagrange_nodal(domain.device(), domain.deltatime()));
}

void LagrangeElements(Domain &domain, Index_t)
{
   exit_on(lgpu_lagrange_elements(domain.device(), domain.deltatime()));
}

void CalcTimeConstraintsForElems(Domain &domain)
{
   exit_on(lgpu_time_constraints(domain.device(), &domain.dtcourant(), &domain.dthydro()));
}

void LagrangeLeapFrog(Domain &domain)
{
   LagrangeNodal(domain);
   LagrangeElements(domain, domain.numElem());
   CalcTimeConstraintsForElems(domain);
}

int LagrangeLeapFrogStatus(Domain &domain)
{
   lgpu_ctx *g = domain.device();
   if (!g) return LGPU_EINVAL;
   int st = lgpu_lagrange_nodal(g, domain.deltatime());
   if (st) return st;
   st = lgpu_lagrange_elements(g, domain.deltatime());
   if (st) return st;
   return lgpu_time_constraints(g, &domain.dtcourant(), &domain.dthydro());
}

// ---------------------------------------------------------------------------
// C API.

struct lh_domain {
   Domain *d;
   bool owned;
};

struct lh_ensemble {
   std::vector<std::unique_ptr<Domain>> members;
   std::vector<lh_domain> views;
   lgpu_ensemble *dev = nullptr;
};

namespace {

// The lgpu_mesh_desc of a host Domain (arrays borrowed from it; the region
// pointer table lives in *lists).
lgpu_mesh_desc describe(Domain &d, std::vector<const Index_t *> &lists, int precision)
{
   lists.resize((size_t)d.numReg());
   for (Int_t r = 0; r < d.numReg(); ++r) lists[(size_t)r] = d.regElemlist(r);
   lgpu_mesh_desc m{};
   m.numElem = d.numElem();
   m.numNode = d.numNode();
   m.sizeX = d.sizeX();
   m.sizeY = d.sizeY();
   m.sizeZ = d.sizeZ();
   m.nodelist = d.nodelist(0);
   m.lxim = &d.lxim(0);
   m.lxip = &d.lxip(0);
   m.letam = &d.letam(0);
   m.letap = &d.letap(0);
   m.lzetam = &d.lzetam(0);
   m.lzetap = &d.lzetap(0);
   m.elemBC = &d.elemBC(0);
   m.nodeElemStart = d.nodeElemStartVec().data();
   m.nodeElemCornerList = d.cornerListVec().data();
   m.symmX = d.symmXdata();
   m.symmY = d.symmYdata();
   m.symmZ = d.symmZdata();
   m.numSymmX = m.numSymmY = m.numSymmZ = (d.sizeX() + 1) * (d.sizeX() + 1);
   m.numReg = d.numReg();
   m.regElemSize = &d.regElemSize(0);
   m.regElemlist = lists.data();
   m.cost = d.cost();
   m.precision = precision;
   m.k = d.constants();
   return m;
}

lgpu_time_state time_state(Domain &d)
{
   lgpu_time_state ts;
   ts.time = d.time();
   ts.deltatime = d.deltatime();
   ts.dtcourant = d.dtcourant();
   ts.dthydro = d.dthydro();
   ts.stoptime = d.stoptime();
   ts.dtfixed = d.dtfixed();
   ts.dtmax = d.dtmax();
   ts.deltatimemultlb = d.deltatimemultlb();
   ts.deltatimemultub = d.deltatimemultub();
   ts.cycle = d.cycle();
   return ts;
}

void set_time_state(Domain &d, const lgpu_time_state &ts)
{
   d.time() = ts.time;
   d.deltatime() = ts.deltatime;
   d.dtcourant() = ts.dtcourant;
   d.dthydro() = ts.dthydro;
   d.cycle() = ts.cycle;
}

} // namespace

extern "C" {

lh_domain *lh_create(int32_t s, int32_t nr, int32_t balance, int32_t cost, double emult)
{
   if (s < 1 || nr < 1 || cost < 0) return nullptr;
   lh_domain *h = new lh_domain{new Domain(1, 0, 0, 0, s, 1, nr, balance, cost), true};
   if (emult != 1.0) h->d->ScaleInitialEnergy(emult);
   return h;
}

void lh_destroy(lh_domain *h)
{
   if (!h || !h->owned) return;
   delete h->d;
   delete h;
}

int lh_attach_device(lh_domain *h, int32_t precision)
{
   return h ? h->d->AttachDevice(precision) : LGPU_EINVAL;
}

int lh_attach_device_flags(lh_domain *h, int32_t precision, int32_t flags)
{
   return h ? h->d->AttachDevice(precision, flags) : LGPU_EINVAL;
}

lgpu_ctx *lh_device(lh_domain *h) { return h ? h->d->device() : nullptr; }

int lh_step(lh_domain *h)
{
   if (!h || !h->d->device()) return LGPU_EINVAL;
   TimeIncrement(*h->d);
   return LagrangeLeapFrogStatus(*h->d);
}

int lh_run(lh_domain *h, int32_t its, int32_t device_loop)
{
   if (!h || !h->d->device()) return LGPU_EINVAL;
   Domain &d = *h->d;
   if (device_loop) {
      lgpu_time_state ts;
      ts.time = d.time();
      ts.deltatime = d.deltatime();
      ts.dtcourant = d.dtcourant();
      ts.dthydro = d.dthydro();
      ts.stoptime = d.stoptime();
      ts.dtfixed = d.dtfixed();
      ts.dtmax = d.dtmax();
      ts.deltatimemultlb = d.deltatimemultlb();
      ts.deltatimemultub = d.deltatimemultub();
      ts.cycle = d.cycle();
      const int st = lgpu_run(d.device(), &ts, its);
      d.time() = ts.time;
      d.deltatime() = ts.deltatime;
      d.dtcourant() = ts.dtcourant;
      d.dthydro() = ts.dthydro;
      d.cycle() = ts.cycle;
      return st;
   }
   while (d.time() < d.stoptime() && d.cycle() < its) {
      TimeIncrement(d);
      const int st = LagrangeLeapFrogStatus(d);
      if (st) return st;
   }
   return LGPU_OK;
}

int lh_sync_from_device(lh_domain *h) { return h ? h->d->SyncFromDevice() : LGPU_EINVAL; }
int32_t lh_num_elem(lh_domain *h) { return h ? h->d->numElem() : 0; }
int32_t lh_num_node(lh_domain *h) { return h ? h->d->numNode() : 0; }
int32_t lh_num_reg(lh_domain *h) { return h ? h->d->numReg() : 0; }
double *lh_real(lh_domain *h, int32_t f) { return h ? h->d->realField(f) : nullptr; }

int32_t *lh_int(lh_domain *h, int32_t which, int32_t region, int64_t *len)
{
   if (!h || !len) return nullptr;
   Domain &d = *h->d;
   const int64_t E = d.numElem(), N = d.numNode(), en = d.sizeX() + 1;
   switch (which) {
   case 0: *len = 8 * E; return d.nodelist(0);
   case 1: *len = E; return &d.lxim(0);
   case 2: *len = E; return &d.lxip(0);
   case 3: *len = E; return &d.letam(0);
   case 4: *len = E; return &d.letap(0);
   case 5: *len = E; return &d.lzetam(0);
   case 6: *len = E; return &d.lzetap(0);
   case 7: *len = E; return &d.elemBC(0);
   case 8: *len = N + 1; return const_cast<Index_t *>(d.nodeElemStartVec().data());
   case 9: *len = (int64_t)d.cornerListVec().size(); return const_cast<Index_t *>(d.cornerListVec().data());
   case 10: *len = en * en; return d.symmXempty() ? nullptr : d.symmXdata();
   case 11: *len = en * en; return d.symmYempty() ? nullptr : d.symmYdata();
   case 12: *len = en * en; return d.symmZempty() ? nullptr : d.symmZdata();
   case 13: *len = E; return d.regNumList();
   case 14: *len = d.numReg(); return &d.regElemSize(0);
   case 15:
      if (region < 0 || region >= d.numReg()) break;
      *len = d.regElemSize(region);
      return d.regElemlist(region);
   default: break;
   }
   *len = 0;
   return nullptr;
}

double lh_get(lh_domain *h, int32_t which)
{
   if (!h) return 0.0;
   Domain &d = *h->d;
   switch (which) {
   case 0: return d.time();
   case 1: return d.deltatime();
   case 2: return d.dtcourant();
   case 3: return d.dthydro();
   case 4: return (double)d.cycle();
   case 5: return d.stoptime();
   case 6: return d.dtfixed();
   case 7: return d.dtmax();
   case 8: return d.deltatimemultlb();
   case 9: return d.deltatimemultub();
   default: return 0.0;
   }
}

void lh_set(lh_domain *h, int32_t which, double v)
{
   if (!h) return;
   Domain &d = *h->d;
   switch (which) {
   case 0: d.time() = v; break;
   case 1: d.deltatime() = v; break;
   case 2: d.dtcourant() = v; break;
   case 3: d.dthydro() = v; break;
   case 4: d.cycle() = (Int_t)v; break;
   case 5: d.stoptime() = v; break;
   case 6: d.dtfixed() = v; break;
   case 7: d.dtmax() = v; break;
   case 8: d.deltatimemultlb() = v; break;
   case 9: d.deltatimemultub() = v; break;
   default: break;
   }
}

int lh_write_vtk(lh_domain *h, const char *path)
{
   if (!h || !path) return -1;
   return WriteVtk(*h->d, path);
}

void lh_symmetry(lh_domain *h, double out[3])
{
   // lulesh-util.cc:197-212 over the host copy of e.
   Domain &d = *h->d;
   const Index_t nx = d.sizeX();
   Real_t maxAbs = 0.0, tot = 0.0, maxRel = 0.0;
   for (Index_t j = 0; j < nx; ++j)
      for (Index_t k = j + 1; k < nx; ++k) {
         const Real_t ad = fabs(d.e(j * nx + k) - d.e(k * nx + j));
         tot += ad;
         if (maxAbs < ad) maxAbs = ad;
         const Real_t rd = ad / d.e(k * nx + j);
         if (maxRel < rd) maxRel = rd;
      }
   out[0] = maxAbs;
   out[1] = tot;
   out[2] = maxRel;
}

void lh_glibc_rand(uint32_t seed, int32_t n, int32_t *out)
{
   lhost::GlibcRand g(seed);
   for (int32_t i = 0; i < n; ++i) out[i] = g.next();
}

int32_t lh_region_runs(int32_t s, int32_t nr, int32_t balance, int32_t *reg, int32_t *len,
                       int32_t capacity)
{
   if (s < 1 || nr < 1 || balance < 0 || (int64_t)s * s * s > INT32_MAX || capacity < 0 ||
       (capacity > 0 && (!reg || !len)))
      return -1;
   std::vector<Int_t> runReg, runLen;
   lhost::draw_region_runs(s * s * s, nr, balance, runReg, runLen);
   const int32_t n = (int32_t)runReg.size();
   for (int32_t i = 0; i < std::min(n, capacity); ++i) {
      reg[i] = runReg[(size_t)i];
      len[i] = runLen[(size_t)i];
   }
   return n;
}

void lh_ensemble_draw(uint64_t seed, int32_t n, int32_t *s, double *emult, int32_t *balance,
                      int32_t *cost)
{
   std::mt19937_64 g(seed);
   const double lo = log(0.5), span = log(2.0) - log(0.5);
   for (int32_t i = 0; i < n; ++i) {
      const uint64_t x0 = g(), x1 = g(), x2 = g(), x3 = g();
      if (s) s[i] = 16 + (int32_t)(x0 % 49u);
      if (emult) emult[i] = exp(lo + span * ((double)(x1 >> 11) * 0x1.0p-53));
      if (balance) balance[i] = 1 + (int32_t)(x2 % 2u);
      if (cost) cost[i] = 1 + (int32_t)(x3 % 2u);
   }
}

lh_ensemble *lh_ensemble_create(int32_t n, const int32_t *s, const double *emult, const int32_t *nr,
                                const int32_t *balance, const int32_t *cost, int32_t precision)
{
   return lh_ensemble_create_flags(n, s, emult, nr, balance, cost, precision, 0);
}

lh_ensemble *lh_ensemble_create_flags(int32_t n, const int32_t *s, const double *emult,
                                      const int32_t *nr, const int32_t *balance, const int32_t *cost,
                                      int32_t precision, int32_t flags)
{
   if (n < 1 || !s || !emult || !nr || !balance || !cost) return nullptr;
   for (int32_t i = 0; i < n; ++i)
      if (s[i] < 1 || nr[i] < 1 || cost[i] < 0) return nullptr;
   auto *e = new lh_ensemble();
   for (int32_t i = 0; i < n; ++i) {
      e->members.emplace_back(new Domain(1, 0, 0, 0, s[i], 1, nr[i], balance[i], cost[i]));
      if (emult[i] != 1.0) e->members.back()->ScaleInitialEnergy(emult[i]);
   }
   for (auto &m : e->members) e->views.push_back(lh_domain{m.get(), false});
   std::vector<std::vector<const Index_t *>> lists((size_t)n);
   std::vector<lgpu_mesh_desc> descs;
   for (int32_t i = 0; i < n; ++i) {
      descs.push_back(describe(*e->members[(size_t)i], lists[(size_t)i], precision));
      descs.back().flags = flags;
   }
   if (lgpu_ensemble_create(descs.data(), n, &e->dev) != LGPU_OK) {
      delete e;
      return nullptr;
   }
   static const int kUp[] = {LGPU_X, LGPU_Y, LGPU_Z, LGPU_XD, LGPU_YD, LGPU_ZD, LGPU_NODALMASS,
                             LGPU_E, LGPU_P, LGPU_Q, LGPU_V, LGPU_VOLO, LGPU_SS, LGPU_ELEMMASS};
   for (int32_t i = 0; i < n; ++i)
      for (int f : kUp)
         if (lgpu_ensemble_upload(e->dev, i, f, e->members[(size_t)i]->realField(f)) != LGPU_OK) {
            lh_ensemble_destroy(e);
            return nullptr;
         }
   return e;
}

void lh_ensemble_destroy(lh_ensemble *e)
{
   if (!e) return;
   if (e->dev) lgpu_ensemble_destroy(e->dev);
   delete e;
}

int lh_ensemble_run(lh_ensemble *e, int32_t its, int32_t *status)
{
   if (!e) return LGPU_EINVAL;
   std::vector<lgpu_time_state> ts;
   for (auto &m : e->members) ts.push_back(time_state(*m));
   const int st = lgpu_ensemble_run(e->dev, ts.data(), its, status);
   for (size_t i = 0; i < ts.size(); ++i) set_time_state(*e->members[i], ts[i]);
   return st;
}

int lh_ensemble_sync(lh_ensemble *e)
{
   if (!e) return LGPU_EINVAL;
   for (size_t i = 0; i < e->members.size(); ++i)
      for (int f : kStepFields) {
         const int st = lgpu_ensemble_download(e->dev, (int32_t)i, f, e->members[i]->realField(f));
         if (st) return st;
      }
   return LGPU_OK;
}

lh_domain *lh_ensemble_member(lh_ensemble *e, int32_t m)
{
   if (!e || m < 0 || (size_t)m >= e->views.size()) return nullptr;
   return &e->views[(size_t)m];
}

lgpu_ensemble *lh_ensemble_device(lh_ensemble *e) { return e ? e->dev : nullptr; }
int32_t lh_ensemble_size(lh_ensemble *e) { return e ? (int32_t)e->members.size() : 0; }

} // extern "C"
