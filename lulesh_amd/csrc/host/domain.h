// domain.h -- host-side Domain of the MI355X LULESH step.
//
// Keeps the reference's data contract (class Domain, lulesh.h:148-595): the
// same constructor arguments, the same per-index accessors and the same
// set-up (mesh, node-to-corner lists, region sets, symmetry sets, face
// neighbours, boundary masks, initial state).  The physics arrays also live
// on the device (lgpu_ctx); the host copies are the set-up source and are
// refreshed from the device on request (SyncFromDevice).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "lulesh_gpu.h"

typedef int32_t Index_t;
typedef double Real_t;
typedef int32_t Int_t;

enum { VolumeError = -1, QStopError = -2 };

class Domain {
 public:
   // Domain::Domain, lulesh-init.cc:16-194.  One domain per process:
   // numRanks = tp = 1 and loc = 0 (the reference's MPI decomposition is out
   // of scope).
   Domain(Int_t numRanks, Index_t colLoc, Index_t rowLoc, Index_t planeLoc, Index_t nx, Int_t tp,
          Int_t nr, Int_t balance, Int_t cost);
   ~Domain();
   Domain(const Domain &) = delete;
   Domain &operator=(const Domain &) = delete;

   // Scales the deposited energy by m and re-derives the first dt as
   // lulesh-init.cc:183-192 does for the unscaled energy.
   void ScaleInitialEnergy(Real_t m);

   // Device mirror.
   int AttachDevice(int precision, int flags = 0); // lgpu_create + upload of the state
   lgpu_ctx *device() { return m_gpu; }
   int SyncFromDevice();            // download every state field

   // Node-centred accessors (lulesh.h:266-294).
   Real_t &x(Index_t i) { return m_x[i]; }
   Real_t &y(Index_t i) { return m_y[i]; }
   Real_t &z(Index_t i) { return m_z[i]; }
   Real_t &xd(Index_t i) { return m_xd[i]; }
   Real_t &yd(Index_t i) { return m_yd[i]; }
   Real_t &zd(Index_t i) { return m_zd[i]; }
   Real_t &xdd(Index_t i) { return m_xdd[i]; }
   Real_t &ydd(Index_t i) { return m_ydd[i]; }
   Real_t &zdd(Index_t i) { return m_zdd[i]; }
   Real_t &fx(Index_t i) { return m_fx[i]; }
   Real_t &fy(Index_t i) { return m_fy[i]; }
   Real_t &fz(Index_t i) { return m_fz[i]; }
   Real_t &nodalMass(Index_t i) { return m_nodalMass[i]; }
   Index_t symmX(Index_t i) { return m_symmX[i]; }
   Index_t symmY(Index_t i) { return m_symmY[i]; }
   Index_t symmZ(Index_t i) { return m_symmZ[i]; }
   bool symmXempty() { return m_symmX.empty(); }
   bool symmYempty() { return m_symmY.empty(); }
   bool symmZempty() { return m_symmZ.empty(); }
   Index_t *symmXdata() { return m_symmX.data(); }
   Index_t *symmYdata() { return m_symmY.data(); }
   Index_t *symmZdata() { return m_symmZ.data(); }

   // Element-centred accessors (lulesh.h:299-367).
   Index_t &regElemSize(Index_t r) { return m_regElemSize[r]; }
   Index_t &regNumList(Index_t i) { return m_regNumList[i]; }
   Index_t *regNumList() { return m_regNumList.data(); }
   Index_t *regElemlist(Int_t r) { return m_regElemlist[r].data(); }
   Index_t &regElemlist(Int_t r, Index_t i) { return m_regElemlist[r][i]; }
   Index_t *nodelist(Index_t i) { return &m_nodelist[(size_t)8 * i]; }
   Index_t &lxim(Index_t i) { return m_lxim[i]; }
   Index_t &lxip(Index_t i) { return m_lxip[i]; }
   Index_t &letam(Index_t i) { return m_letam[i]; }
   Index_t &letap(Index_t i) { return m_letap[i]; }
   Index_t &lzetam(Index_t i) { return m_lzetam[i]; }
   Index_t &lzetap(Index_t i) { return m_lzetap[i]; }
   Int_t &elemBC(Index_t i) { return m_elemBC[i]; }
   Real_t &e(Index_t i) { return m_e[i]; }
   Real_t &p(Index_t i) { return m_p[i]; }
   Real_t &q(Index_t i) { return m_q[i]; }
   Real_t &ql(Index_t i) { return m_ql[i]; }
   Real_t &qq(Index_t i) { return m_qq[i]; }
   Real_t &v(Index_t i) { return m_v[i]; }
   Real_t &delv(Index_t i) { return m_delv[i]; }
   Real_t &volo(Index_t i) { return m_volo[i]; }
   Real_t &vnew(Index_t i) { return m_vnew[i]; }
   Real_t &vdov(Index_t i) { return m_vdov[i]; }
   Real_t &arealg(Index_t i) { return m_arealg[i]; }
   Real_t &ss(Index_t i) { return m_ss[i]; }
   Real_t &elemMass(Index_t i) { return m_elemMass[i]; }
   Index_t nodeElemCount(Index_t n) { return m_nodeElemStart[n + 1] - m_nodeElemStart[n]; }
   Index_t *nodeElemCornerList(Index_t n) { return &m_nodeElemCornerList[m_nodeElemStart[n]]; }

   // Constants (lulesh.h:377-399).
   Real_t u_cut() const { return m_k.u_cut; }
   Real_t e_cut() const { return m_k.e_cut; }
   Real_t p_cut() const { return m_k.p_cut; }
   Real_t q_cut() const { return m_k.q_cut; }
   Real_t v_cut() const { return m_k.v_cut; }
   Real_t hgcoef() const { return m_k.hgcoef; }
   Real_t qstop() const { return m_k.qstop; }
   Real_t qqc() const { return m_k.qqc; }
   Real_t dvovmax() const { return m_k.dvovmax; }
   const lgpu_constants &constants() const { return m_k; }

   // Time-step controls (lulesh.h:402-411).
   Real_t &time() { return m_time; }
   Real_t &deltatime() { return m_deltatime; }
   Real_t &deltatimemultlb() { return m_deltatimemultlb; }
   Real_t &deltatimemultub() { return m_deltatimemultub; }
   Real_t &stoptime() { return m_stoptime; }
   Real_t &dtcourant() { return m_dtcourant; }
   Real_t &dthydro() { return m_dthydro; }
   Real_t &dtmax() { return m_dtmax; }
   Real_t &dtfixed() { return m_dtfixed; }
   Int_t &cycle() { return m_cycle; }

   Index_t &sizeX() { return m_sizeX; }
   Index_t &sizeY() { return m_sizeY; }
   Index_t &sizeZ() { return m_sizeZ; }
   Index_t &numReg() { return m_numReg; }
   Int_t &cost() { return m_cost; }
   Index_t &numElem() { return m_numElem; }
   Index_t &numNode() { return m_numNode; }

   // Raw arrays for the C API.
   Real_t *realField(int field);
   const std::vector<Index_t> &nodeElemStartVec() const { return m_nodeElemStart; }
   const std::vector<Index_t> &cornerListVec() const { return m_nodeElemCornerList; }

 private:
   void BuildMesh(Int_t nx, Int_t edgeNodes, Int_t edgeElems);
   void SetupThreadSupportStructures();
   void CreateRegionIndexSets(Int_t nreg, Int_t balance);
   void SetupSymmetryPlanes(Int_t edgeNodes);
   void SetupElementConnectivities(Int_t edgeElems);
   void SetupBoundaryConditions(Int_t edgeElems);

   lgpu_constants m_k;
   std::vector<Real_t> m_x, m_y, m_z, m_xd, m_yd, m_zd, m_xdd, m_ydd, m_zdd, m_fx, m_fy, m_fz,
      m_nodalMass;
   std::vector<Index_t> m_symmX, m_symmY, m_symmZ;
   Int_t m_numReg = 0, m_cost = 0;
   std::vector<Index_t> m_regElemSize, m_regNumList;
   std::vector<std::vector<Index_t>> m_regElemlist;
   std::vector<Index_t> m_nodelist, m_lxim, m_lxip, m_letam, m_letap, m_lzetam, m_lzetap;
   std::vector<Int_t> m_elemBC;
   std::vector<Real_t> m_e, m_p, m_q, m_ql, m_qq, m_v, m_volo, m_vnew, m_delv, m_vdov, m_arealg,
      m_ss, m_elemMass;
   std::vector<Real_t> m_delx_xi, m_delx_eta, m_delx_zeta, m_delv_xi, m_delv_eta, m_delv_zeta;
   Real_t m_dtcourant = 0, m_dthydro = 0, m_dtfixed = 0, m_time = 0, m_deltatime = 0;
   Real_t m_deltatimemultlb = 0, m_deltatimemultub = 0, m_dtmax = 0, m_stoptime = 0;
   Int_t m_cycle = 0;
   Index_t m_sizeX = 0, m_sizeY = 0, m_sizeZ = 0, m_numElem = 0, m_numNode = 0;
   std::vector<Index_t> m_nodeElemStart, m_nodeElemCornerList;
   lgpu_ctx *m_gpu = nullptr;
};

// The reference's per-cycle call structure (lulesh.cc:166-222, 1223-1269,
// 2431-2443, 2576-2645) over the device path.  Device errors are reported
// the reference's way: exit(VolumeError) / exit(QStopError).
void TimeIncrement(Domain &domain);
void LagrangeNodal(Domain &domain);
void LagrangeElements(Domain &domain, Index_t numElem);
namespace lhost {
// The region draws of CreateRegionIndexSets (lulesh-init.cc:401-480) as runs
// of consecutive elements: region numbers (1-based) and run lengths.
void draw_region_runs(Index_t numElem, Int_t nr, Int_t balance, std::vector<Int_t> &runReg,
                      std::vector<Int_t> &runLen);
} // namespace lhost

// End-of-run visualization dump (DumpToVisit, lulesh-viz.cc) as legacy VTK
// (host/viz.cc); writes the host copies.  0 on success.
int WriteVtk(Domain &domain, const char *path);
void CalcTimeConstraintsForElems(Domain &domain);
void LagrangeLeapFrog(Domain &domain);

// Same loop, returning the status instead of exiting (for embedding).
int LagrangeLeapFrogStatus(Domain &domain);
