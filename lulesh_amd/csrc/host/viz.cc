// viz.cc -- the end-of-run visualization dump (DumpToVisit, lulesh-viz.cc:
// 97-269 and 279-364) in the legacy VTK format, since the reference's SILO
// library is not available: the hexahedral mesh (node coordinates as FP32,
// as the reference writes them), the region of each element, the element
// fields e, p, v, q and the node fields speed, xd, yd, zd.  Binary legacy VTK
// is big-endian.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "host/domain.h"

namespace {

template <typename T>
void put_array(FILE *f, const T *a, size_t n)
{
   std::vector<unsigned char> buf(n * sizeof(T));
   for (size_t i = 0; i < n; ++i) {
      const unsigned char *s = reinterpret_cast<const unsigned char *>(&a[i]);
      for (size_t k = 0; k < sizeof(T); ++k) buf[i * sizeof(T) + k] = s[sizeof(T) - 1 - k];
   }
   fwrite(buf.data(), 1, buf.size(), f);
}

void scalars(FILE *f, const char *name, const double *a, size_t n)
{
   fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
   put_array(f, a, n);
   fputc('\n', f);
}

} // namespace

// Writes the domain's current host state (call SyncFromDevice first).
int WriteVtk(Domain &d, const char *path)
{
   FILE *f = fopen(path, "wb");
   if (!f) return -1;
   const size_t N = (size_t)d.numNode(), E = (size_t)d.numElem();
   fprintf(f, "# vtk DataFile Version 3.0\nLULESH cycle %d time %.17e\nBINARY\n", d.cycle(),
           d.time());
   fprintf(f, "DATASET UNSTRUCTURED_GRID\nPOINTS %zu float\n", N);
   std::vector<float> xyz(3 * N);
   for (size_t n = 0; n < N; ++n) {
      xyz[3 * n] = (float)d.x((Index_t)n);
      xyz[3 * n + 1] = (float)d.y((Index_t)n);
      xyz[3 * n + 2] = (float)d.z((Index_t)n);
   }
   put_array(f, xyz.data(), xyz.size());
   // The reference's node order (lulesh-init.cc:252-259) is VTK's hexahedron
   // order: the k face counter-clockwise, then the k+1 face.
   fprintf(f, "\nCELLS %zu %zu\n", E, 9 * E);
   std::vector<int32_t> cells(9 * E);
   for (size_t e = 0; e < E; ++e) {
      cells[9 * e] = 8;
      const Index_t *nl = d.nodelist((Index_t)e);
      for (int j = 0; j < 8; ++j) cells[9 * e + 1 + j] = nl[j];
   }
   put_array(f, cells.data(), cells.size());
   fprintf(f, "\nCELL_TYPES %zu\n", E);
   std::vector<int32_t> types(E, 12); // VTK_HEXAHEDRON
   put_array(f, types.data(), types.size());
   fprintf(f, "\nCELL_DATA %zu\nSCALARS region int 1\nLOOKUP_TABLE default\n", E);
   put_array(f, d.regNumList(), E);
   fputc('\n', f);
   std::vector<double> a(E);
   for (size_t e = 0; e < E; ++e) a[e] = d.e((Index_t)e);
   scalars(f, "e", a.data(), E);
   for (size_t e = 0; e < E; ++e) a[e] = d.p((Index_t)e);
   scalars(f, "p", a.data(), E);
   for (size_t e = 0; e < E; ++e) a[e] = d.v((Index_t)e);
   scalars(f, "v", a.data(), E);
   for (size_t e = 0; e < E; ++e) a[e] = d.q((Index_t)e);
   scalars(f, "q", a.data(), E);
   fprintf(f, "POINT_DATA %zu\n", N);
   std::vector<double> b(N);
   for (size_t n = 0; n < N; ++n) {
      const Index_t i = (Index_t)n;
      b[n] = sqrt(d.xd(i) * d.xd(i) + d.yd(i) * d.yd(i) + d.zd(i) * d.zd(i));
   }
   scalars(f, "speed", b.data(), N);
   for (size_t n = 0; n < N; ++n) b[n] = d.xd((Index_t)n);
   scalars(f, "xd", b.data(), N);
   for (size_t n = 0; n < N; ++n) b[n] = d.yd((Index_t)n);
   scalars(f, "yd", b.data(), N);
   for (size_t n = 0; n < N; ++n) b[n] = d.zd((Index_t)n);
   scalars(f, "zd", b.data(), N);
   const bool ok = ferror(f) == 0;
   return (fclose(f) == 0 && ok) ? 0 : -1;
}
