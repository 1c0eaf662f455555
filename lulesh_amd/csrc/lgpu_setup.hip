// lgpu_setup.hip -- Domain::Domain (lulesh-init.cc:16-194) for the
// reference's Sedov problem as device kernels: every table and initial field
// is a closed-form function of the element or node index, except the region
// layout (a sequential glibc rand() draw the host makes) and the nodal mass
// (a sum over the node's elements, taken in the reference's element order).
// Built without FMA contraction: the reference volumes and nodal masses are
// bit-identical to the host's.
#include <algorithm>

#include "lgpu_geom.h"
#include "lgpu_setup.h"

namespace lgpu {
namespace {

// Lattice line i of an s^3 brick: BuildMesh (lulesh-init.cc:223-243) sets
// tx = 1.125 * i / s (FP64, in this order).
__device__ __forceinline__ double lattice(int i, int s) { return 1.125 * (double)i / (double)s; }

// Elements along one lattice direction that touch node line t (0..s), and
// their count over the lines before t.
__device__ __forceinline__ int64_t touching(int t, int s) { return (t == 0 || t == s) ? 1 : 2; }
__device__ __forceinline__ int64_t touching_before(int t) { return t == 0 ? 0 : 2 * (int64_t)t - 1; }

// Corner of node (i+di, j+dj, k+dk) in element (i, j, k): the node order of
// BuildMesh (lulesh-init.cc:252-259).
__device__ __forceinline__ int corner_of(int di, int dj, int dk)
{
   return 4 * dk + (dj ? (di ? 2 : 3) : (di ? 1 : 0));
}

// Nodes: coordinates (BuildMesh :223-243), symmetry-plane membership
// (SetupSymmetryPlanes :514-535, as the bit mask the kernels test), and the
// node-to-corner lists (SetupThreadSupportStructures :272-337): node
// (i, j, k) lies in ca(i) ca(j) ca(k) elements, ca = 1 on the two outer
// lattice lines and 2 inside, so its list starts at the closed-form count of
// the corners of every node before it, and holds its elements' corners in
// ascending element order.
__global__ void __launch_bounds__(256) k_setup_nodes(DevState s, int sz)
{
   const int64_t N = s.numNode, en = sz + 1, w = 2 * (int64_t)sz;
   int32_t *start = const_cast<int32_t *>(s.nodeElemStart);
   int32_t *corners = const_cast<int32_t *>(s.cornerList);
   uint8_t *bc = const_cast<uint8_t *>(s.nodeBC);
   for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N;
        n += (int64_t)gridDim.x * blockDim.x) {
      const int i = (int)(n % en), j = (int)((n / en) % en), k = (int)(n / (en * en));
      s.x[n] = lattice(i, sz);
      s.y[n] = lattice(j, sz);
      s.z[n] = lattice(k, sz);
      bc[n] = (uint8_t)((i == 0 ? 1 : 0) | (j == 0 ? 2 : 0) | (k == 0 ? 4 : 0));
      const int64_t first = touching_before(k) * w * w + touching(k, sz) * touching_before(j) * w +
                            touching(k, sz) * touching(j, sz) * touching_before(i);
      start[n] = (int32_t)first;
      if (n == N - 1) start[N] = (int32_t)(8 * (int64_t)s.numElem);
      int64_t at = first;
      for (int dk = 1; dk >= 0; --dk)
         for (int dj = 1; dj >= 0; --dj)
            for (int di = 1; di >= 0; --di) {
               const int a = i - di, b = j - dj, c = k - dk;
               if (a < 0 || b < 0 || c < 0 || a >= sz || b >= sz || c >= sz) continue;
               const int64_t e = ((int64_t)c * sz + b) * sz + a;
               corners[at++] = (int32_t)(8 * e + corner_of(di, dj, dk));
            }
   }
}

// Elements: node list (BuildMesh :246-265), face neighbours
// (SetupElementConnectivities :539-565), boundary masks
// (SetupBoundaryConditions :569-673), reference volume and element mass
// (:161-174, CalcElemVolume lulesh.cc:1273-1356) and v = 1 (:97-100).
__global__ void __launch_bounds__(256) k_setup_elems(DevState s, int sz)
{
   const int64_t E = s.numElem, en = sz + 1, s2 = (int64_t)sz * sz;
   int32_t *nl = const_cast<int32_t *>(s.nodelist);
   int32_t *lxim = const_cast<int32_t *>(s.lxim), *lxip = const_cast<int32_t *>(s.lxip);
   int32_t *letam = const_cast<int32_t *>(s.letam), *letap = const_cast<int32_t *>(s.letap);
   int32_t *lzetam = const_cast<int32_t *>(s.lzetam), *lzetap = const_cast<int32_t *>(s.lzetap);
   int32_t *ebc = const_cast<int32_t *>(s.elemBC);
   for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
        e += (int64_t)gridDim.x * blockDim.x) {
      const int a = (int)(e % sz), b = (int)((e / sz) % sz), c = (int)(e / s2);
      const int64_t n0 = ((int64_t)c * en + b) * en + a;
      const int64_t off[8] = {0, 1, en + 1, en, en * en, en * en + 1, en * en + en + 1, en * en + en};
#pragma unroll
      for (int j = 0; j < 8; ++j) nl[8 * e + j] = (int32_t)(n0 + off[j]);
      lxim[e] = (int32_t)(e > 0 ? e - 1 : 0);
      lxip[e] = (int32_t)(e + 1 < E ? e + 1 : e);
      letam[e] = (int32_t)(e >= sz ? e - sz : e);
      letap[e] = (int32_t)(e + sz < E ? e + sz : e);
      lzetam[e] = (int32_t)(e >= s2 ? e - s2 : e);
      lzetap[e] = (int32_t)(e + s2 < E ? e + s2 : e);
      ebc[e] = (a == 0 ? 0x00001 : 0) | (a == sz - 1 ? 0x00010 : 0) | (b == 0 ? 0x00040 : 0) |
               (b == sz - 1 ? 0x00400 : 0) | (c == 0 ? 0x01000 : 0) | (c == sz - 1 ? 0x10000 : 0);
      const double xa[2] = {lattice(a, sz), lattice(a + 1, sz)};
      const double yb[2] = {lattice(b, sz), lattice(b + 1, sz)};
      const double zc[2] = {lattice(c, sz), lattice(c + 1, sz)};
      const double x[8] = {xa[0], xa[1], xa[1], xa[0], xa[0], xa[1], xa[1], xa[0]};
      const double y[8] = {yb[0], yb[0], yb[1], yb[1], yb[0], yb[0], yb[1], yb[1]};
      const double z[8] = {zc[0], zc[0], zc[0], zc[0], zc[1], zc[1], zc[1], zc[1]};
      const double volume = hex_volume<double>(x, y, z);
      s.volo[e] = volume;
      s.elemMass[e] = volume;
      s.v[e] = 1.0;
   }
}

// Nodal mass (lulesh-init.cc:175-177): the reference adds volume/8 of each
// element to its nodes in element order, so a node sums its elements'
// contributions in ascending element order, starting from 0.
__global__ void __launch_bounds__(256) k_setup_mass(DevState s, int sz)
{
   const int64_t N = s.numNode, en = sz + 1;
   for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N;
        n += (int64_t)gridDim.x * blockDim.x) {
      const int i = (int)(n % en), j = (int)((n / en) % en), k = (int)(n / (en * en));
      double m = 0.0;
      for (int dk = 1; dk >= 0; --dk)
         for (int dj = 1; dj >= 0; --dj)
            for (int di = 1; di >= 0; --di) {
               const int a = i - di, b = j - dj, c = k - dk;
               if (a < 0 || b < 0 || c < 0 || a >= sz || b >= sz || c >= sz) continue;
               m += s.volo[((int64_t)c * sz + b) * sz + a] / 8.0;
            }
      s.nodalMass[n] = m;
   }
}

// EOS work list: element e of region run r goes to runDest[r] + (e -
// runStart[r]); the runs of one repeat-count group are laid out in element
// order, so each group lists its elements ascending.
__global__ void __launch_bounds__(256) k_setup_perm(DevState s, const int32_t *runStart,
                                                    const int32_t *runDest, int numRuns)
{
   int32_t *perm = const_cast<int32_t *>(s.perm);
   const int64_t E = s.numElem;
   for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
        e += (int64_t)gridDim.x * blockDim.x) {
      int lo = 0, hi = numRuns; // runStart[lo] <= e < runStart[hi]
      while (hi - lo > 1) {
         const int mid = (lo + hi) >> 1;
         if (runStart[mid] <= e) lo = mid;
         else hi = mid;
      }
      perm[runDest[lo] + (e - runStart[lo])] = (int32_t)e;
   }
}

unsigned grid_for(int64_t n)
{
   const int64_t blocks = (n + 255) / 256;
   return (unsigned)std::min<int64_t>(blocks, 65536);
}

} // namespace

hipError_t launch_sedov_setup(const DevState &s, const SedovSetup &t, hipStream_t st)
{
   hipLaunchKernelGGL(k_setup_nodes, dim3(grid_for(s.numNode)), dim3(256), 0, st, s, t.size);
   hipLaunchKernelGGL(k_setup_elems, dim3(grid_for(s.numElem)), dim3(256), 0, st, s, t.size);
   hipLaunchKernelGGL(k_setup_mass, dim3(grid_for(s.numNode)), dim3(256), 0, st, s, t.size);
   hipLaunchKernelGGL(k_setup_perm, dim3(grid_for(s.numElem)), dim3(256), 0, st, s, t.runStart,
                      t.runDest, t.numRuns);
   return hipGetLastError();
}

} // namespace lgpu
