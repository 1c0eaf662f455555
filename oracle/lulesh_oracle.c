/* lulesh_oracle.c -- CPU restatement of the LULESH 2.x Lagrange leapfrog step.
 *
 * TEST INFRASTRUCTURE ONLY (see lulesh_oracle.h).  Single-threaded plain C.
 *
 * Parity contract.  Every floating-point expression below performs the same
 * IEEE operations on the same operands, in the same association order, as the
 * reference line it cites.  Operand order inside one commutative operation
 * and the placement of exact negations may differ, which IEEE arithmetic
 * leaves bit-identical.  Compiled with -ffp-contract=off (oracle/Makefile),
 * the results match the reference's OpenMP build bit for bit; the pinning
 * tests in tests/test_oracle_vs_reference.py check exactly that.
 */
#include "lulesh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Boundary-condition bits of elemBC (lulesh.h:59-87). */
enum {
   BC_XI_M = 0x00007, BC_XI_M_SYMM = 0x00001, BC_XI_M_FREE = 0x00002, BC_XI_M_COMM = 0x00004,
   BC_XI_P = 0x00038, BC_XI_P_SYMM = 0x00008, BC_XI_P_FREE = 0x00010, BC_XI_P_COMM = 0x00020,
   BC_ETA_M = 0x001c0, BC_ETA_M_SYMM = 0x00040, BC_ETA_M_FREE = 0x00080, BC_ETA_M_COMM = 0x00100,
   BC_ETA_P = 0x00e00, BC_ETA_P_SYMM = 0x00200, BC_ETA_P_FREE = 0x00400, BC_ETA_P_COMM = 0x00800,
   BC_ZETA_M = 0x07000, BC_ZETA_M_SYMM = 0x01000, BC_ZETA_M_FREE = 0x02000, BC_ZETA_M_COMM = 0x04000,
   BC_ZETA_P = 0x38000, BC_ZETA_P_SYMM = 0x08000, BC_ZETA_P_FREE = 0x10000, BC_ZETA_P_COMM = 0x20000
};

enum { VOLUME_ERROR = -1, QSTOP_ERROR = -2 };

struct orc_domain {
   int s, numElem, numNode, numReg, cost;
   int scatter;
   double *f[ORC_NUM_REAL_FIELDS];
   int32_t *nodelist, *lxim, *lxip, *letam, *letap, *lzetam, *lzetap, *elemBC;
   int32_t *nodeElemStart, *cornerList, *symmX, *symmY, *symmZ;
   int32_t *regNumList, *regElemSize, **regElemlist;
   double sc[ORC_NUM_SCALARS];
   /* constants, lulesh-init.cc:20-38 */
   double e_cut, p_cut, q_cut, v_cut, u_cut, hgcoef, ss4o3, qstop, monoq_max_slope,
      monoq_limiter_mult, qlc_monoq, qqc_monoq, qqc, eosvmax, eosvmin, pmin, emin,
      dvovmax, refdens;
   /* per-cycle scratch */
   double *sig, *determ, *cfx, *cfy, *cfz, *vnewc;
   double *dxx, *dyy, *dzz;
};

#define F(d, id) ((d)->f[(id)])

/* ------------------------------------------------------------------ */
/* Element geometry                                                     */
/* ------------------------------------------------------------------ */

static void gather8(const double *a, const int32_t *nl, double out[8])
{
   for (int j = 0; j < 8; ++j) out[j] = a[nl[j]];
}

/* u . (v x w), lulesh.cc:1337-1338 (TRIPLE_PRODUCT as used at :1341-1349). */
static double triple(double ux, double uy, double uz, double vx, double vy, double vz,
                     double wx, double wy, double wz)
{
   return ux * (vy * wz - vz * wy) + uy * (vz * wx - vx * wz) + uz * (vx * wy - vy * wx);
}

/* Hexahedron volume, lulesh.cc:1273-1356. */
double orc_elem_volume(const double x[8], const double y[8], const double z[8])
{
#define D(c, a, b) (c[a] - c[b])
   const double t1 = triple(D(x, 3, 1) + D(x, 7, 2), D(y, 3, 1) + D(y, 7, 2), D(z, 3, 1) + D(z, 7, 2),
                            D(x, 6, 3), D(y, 6, 3), D(z, 6, 3),
                            D(x, 2, 0), D(y, 2, 0), D(z, 2, 0));
   const double t2 = triple(D(x, 4, 3) + D(x, 5, 7), D(y, 4, 3) + D(y, 5, 7), D(z, 4, 3) + D(z, 5, 7),
                            D(x, 6, 4), D(y, 6, 4), D(z, 6, 4),
                            D(x, 7, 0), D(y, 7, 0), D(z, 7, 0));
   const double t3 = triple(D(x, 1, 4) + D(x, 2, 5), D(y, 1, 4) + D(y, 2, 5), D(z, 1, 4) + D(z, 2, 5),
                            D(x, 6, 1), D(y, 6, 1), D(z, 6, 1),
                            D(x, 5, 0), D(y, 5, 0), D(z, 5, 0));
#undef D
   double volume = t1 + t2 + t3;
   volume *= 1.0 / 12.0;
   return volume;
}

/* Centre Jacobian columns of one coordinate, lulesh.cc:319-329. */
static void jac_cols(const double c[8], double *dxi, double *deta, double *dzeta)
{
   const double a = c[6] - c[0], b = c[5] - c[3], g = c[7] - c[1], h = c[4] - c[2];
   *dxi = 0.125 * (a + b - g - h);
   *deta = 0.125 * (a - b + g - h);
   *dzeta = 0.125 * (a + b + g + h);
}

/* Shape-function derivatives (first four nodes; the other four are their
 * negatives) and Jacobian determinant, lulesh.cc:290-377. */
static void shape_derivs(const double x[8], const double y[8], const double z[8],
                         double b[3][4], double *volume)
{
   double xxi, xet, xze, yxi, yet, yze, zxi, zet, zze;
   jac_cols(x, &xxi, &xet, &xze);
   jac_cols(y, &yxi, &yet, &yze);
   jac_cols(z, &zxi, &zet, &zze);

   /* cofactors, lulesh.cc:332-342 */
   const double cxxi = (yet * zze) - (zet * yze);
   const double cxet = (zxi * yze) - (yxi * zze);
   const double cxze = (yxi * zet) - (zxi * yet);
   const double cyxi = (zet * xze) - (xet * zze);
   const double cyet = (xxi * zze) - (zxi * xze);
   const double cyze = (zxi * xet) - (xxi * zet);
   const double czxi = (xet * yze) - (yet * xze);
   const double czet = (yxi * xze) - (xxi * yze);
   const double czze = (xxi * yet) - (yxi * xet);

   const double c[3][3] = {{cxxi, cxet, cxze}, {cyxi, cyet, cyze}, {czxi, czet, czze}};
   for (int r = 0; r < 3; ++r) {
      const double ci = c[r][0], ce = c[r][1], cz = c[r][2];
      b[r][0] = -ci - ce - cz;
      b[r][1] = ci - ce - cz;
      b[r][2] = ci + ce - cz;
      b[r][3] = -ci + ce - cz;
   }
   *volume = 8.0 * (xet * cxet + yet * cyet + zet * czet);
}

/* Quadrilateral faces, in the order lulesh.cc:432-473 visits them. */
static const int kNormalFaces[6][4] = {
   {0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}, {4, 7, 6, 5}};

/* Area-weighted node normals, lulesh.cc:381-474. */
void orc_elem_node_normals(const double x[8], const double y[8], const double z[8],
                           double nx[8], double ny[8], double nz[8])
{
   for (int i = 0; i < 8; ++i) nx[i] = ny[i] = nz[i] = 0.0;
   for (int f = 0; f < 6; ++f) {
      const int *n = kNormalFaces[f];
      const double b0x = 0.5 * (x[n[3]] + x[n[2]] - x[n[1]] - x[n[0]]);
      const double b0y = 0.5 * (y[n[3]] + y[n[2]] - y[n[1]] - y[n[0]]);
      const double b0z = 0.5 * (z[n[3]] + z[n[2]] - z[n[1]] - z[n[0]]);
      const double b1x = 0.5 * (x[n[2]] + x[n[1]] - x[n[3]] - x[n[0]]);
      const double b1y = 0.5 * (y[n[2]] + y[n[1]] - y[n[3]] - y[n[0]]);
      const double b1z = 0.5 * (z[n[2]] + z[n[1]] - z[n[3]] - z[n[0]]);
      const double ax = 0.25 * (b0y * b1z - b0z * b1y);
      const double ay = 0.25 * (b0z * b1x - b0x * b1z);
      const double az = 0.25 * (b0x * b1y - b0y * b1x);
      for (int k = 0; k < 4; ++k) {
         nx[n[k]] += ax;
         ny[n[k]] += ay;
         nz[n[k]] += az;
      }
   }
}

/* The six-node bilinear form behind VoluDer, lulesh.cc:602-614:
 * dV/dx = P(y,z)/12, dV/dy = -P(x,z)/12, dV/dz = -P(y,x)/12. */
static double voluder_form(const double a[6], const double b[6])
{
   return (a[1] + a[2]) * (b[0] + b[1]) - (a[0] + a[1]) * (b[1] + b[2]) +
          (a[0] + a[4]) * (b[3] + b[4]) - (a[3] + a[4]) * (b[0] + b[4]) -
          (a[2] + a[5]) * (b[3] + b[5]) + (a[3] + a[5]) * (b[2] + b[5]);
}

/* Node permutation handed to VoluDer for each node's derivative,
 * lulesh.cc:631-662. */
static const int kVoluDerNodes[8][6] = {
   {1, 2, 3, 4, 5, 7}, {2, 3, 0, 5, 6, 4}, {3, 0, 1, 6, 7, 5}, {0, 1, 2, 7, 4, 6},
   {7, 6, 5, 0, 3, 1}, {4, 7, 6, 1, 0, 2}, {5, 4, 7, 2, 1, 3}, {6, 5, 4, 3, 2, 0}};

/* Volume derivative with respect to each node, lulesh.cc:591-663. */
void orc_elem_volume_derivative(const double x[8], const double y[8], const double z[8],
                                double dx[8], double dy[8], double dz[8])
{
   const double twelfth = 1.0 / 12.0;
   for (int n = 0; n < 8; ++n) {
      double px[6], py[6], pz[6];
      for (int k = 0; k < 6; ++k) {
         px[k] = x[kVoluDerNodes[n][k]];
         py[k] = y[kVoluDerNodes[n][k]];
         pz[k] = z[kVoluDerNodes[n][k]];
      }
      dx[n] = voluder_form(py, pz) * twelfth;
      dy[n] = -(voluder_form(px, pz) * twelfth);
      dz[n] = -(voluder_form(py, px) * twelfth);
   }
}

/* Squared-area form of one face, lulesh.cc:1370-1390. */
static double face_area2(const double x[8], const double y[8], const double z[8],
                         int a, int b, int c, int e)
{
   const double fx = (x[c] - x[a]) - (x[e] - x[b]);
   const double fy = (y[c] - y[a]) - (y[e] - y[b]);
   const double fz = (z[c] - z[a]) - (z[e] - z[b]);
   const double gx = (x[c] - x[a]) + (x[e] - x[b]);
   const double gy = (y[c] - y[a]) + (y[e] - y[b]);
   const double gz = (z[c] - z[a]) + (z[e] - z[b]);
   const double fg = fx * gx + fy * gy + fz * gz;
   return (fx * fx + fy * fy + fz * fz) * (gx * gx + gy * gy + gz * gz) - fg * fg;
}

static const int kAreaFaces[6][4] = {
   {0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};

/* Characteristic length, lulesh.cc:1394-1435. */
double orc_elem_char_length(const double x[8], const double y[8], const double z[8],
                            double volume)
{
   double cl = 0.0;
   for (int f = 0; f < 6; ++f) {
      const int *n = kAreaFaces[f];
      const double a = face_area2(x, y, z, n[0], n[1], n[2], n[3]);
      cl = (a < cl) ? cl : a; /* std::max(a, charLength) */
   }
   return 4.0 * volume / sqrt(cl);
}

/* ------------------------------------------------------------------ */
/* Set-up, lulesh-init.cc                                               */
/* ------------------------------------------------------------------ */

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

/* Region index sets, lulesh-init.cc:400-510 (one rank: srand(0), no rotation). */
static void make_regions(orc_domain *d, int nr, int balance)
{
   const int ne = d->numElem;
   srand(0);
   d->numReg = nr;
   d->regElemSize = (int32_t *)zalloc(sizeof(int32_t) * (size_t)nr);
   d->regElemlist = (int32_t **)zalloc(sizeof(int32_t *) * (size_t)nr);
   int next = 0;
   if (nr == 1) {
      for (; next < ne; ++next) d->regNumList[next] = 1;
   } else {
      int *binEnd = (int *)zalloc(sizeof(int) * (size_t)nr);
      int denom = 0, last = -1;
      for (int i = 0; i < nr; ++i) {
         denom += pow(i + 1, balance);
         binEnd[i] = denom;
      }
      while (next < ne) {
         int i = 0, v = rand() % denom;
         while (v >= binEnd[i]) ++i;
         int reg = (i % nr) + 1;
         while (reg == last) {
            v = rand() % denom;
            i = 0;
            while (v >= binEnd[i]) ++i;
            reg = (i % nr) + 1;
         }
         const int bin = rand() % 1000;
         int count;
         if (bin < 773) count = rand() % 15 + 1;
         else if (bin < 937) count = rand() % 16 + 16;
         else if (bin < 970) count = rand() % 32 + 32;
         else if (bin < 974) count = rand() % 64 + 64;
         else if (bin < 978) count = rand() % 128 + 128;
         else if (bin < 981) count = rand() % 256 + 256;
         else count = rand() % 1537 + 512;
         const int end = next + count;
         while (next < end && next < ne) d->regNumList[next++] = reg;
         last = reg;
      }
      free(binEnd);
   }
   for (int i = 0; i < ne; ++i) d->regElemSize[d->regNumList[i] - 1]++;
   for (int r = 0; r < nr; ++r) {
      d->regElemlist[r] = (int32_t *)zalloc(sizeof(int32_t) * (size_t)d->regElemSize[r]);
      d->regElemSize[r] = 0;
   }
   for (int i = 0; i < ne; ++i) {
      const int r = d->regNumList[i] - 1;
      d->regElemlist[r][d->regElemSize[r]++] = i;
   }
}

orc_domain *orc_create(int s, int nr, int balance, int cost, double emult)
{
   if (s < 1 || nr < 1) return NULL;
   orc_domain *d = (orc_domain *)zalloc(sizeof(*d));
   const int en = s + 1;
   d->s = s;
   d->numElem = s * s * s;
   d->numNode = en * en * en;
   d->cost = cost;
   const size_t ne = (size_t)d->numElem, nn = (size_t)d->numNode;

   d->e_cut = 1.0e-7;  d->p_cut = 1.0e-7;  d->q_cut = 1.0e-7;  d->v_cut = 1.0e-10;
   d->u_cut = 1.0e-7;  d->hgcoef = 3.0;    d->ss4o3 = 4.0 / 3.0; d->qstop = 1.0e+12;
   d->monoq_max_slope = 1.0; d->monoq_limiter_mult = 2.0; d->qlc_monoq = 0.5;
   d->qqc_monoq = 2.0 / 3.0; d->qqc = 2.0; d->eosvmax = 1.0e+9; d->eosvmin = 1.0e-9;
   d->pmin = 0.0; d->emin = -1.0e+15; d->dvovmax = 0.1; d->refdens = 1.0;

   for (int k = 0; k < ORC_NUM_REAL_FIELDS; ++k) {
      const size_t n = (k <= ORC_NODALMASS) ? nn : ne;
      d->f[k] = (double *)zalloc(sizeof(double) * n);
   }
   d->nodelist = (int32_t *)zalloc(sizeof(int32_t) * 8 * ne);
   d->lxim = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->lxip = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->letam = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->letap = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->lzetam = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->lzetap = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->elemBC = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->regNumList = (int32_t *)zalloc(sizeof(int32_t) * ne);
   d->symmX = (int32_t *)zalloc(sizeof(int32_t) * (size_t)(en * en));
   d->symmY = (int32_t *)zalloc(sizeof(int32_t) * (size_t)(en * en));
   d->symmZ = (int32_t *)zalloc(sizeof(int32_t) * (size_t)(en * en));
   d->sig = (double *)zalloc(sizeof(double) * ne);
   d->determ = (double *)zalloc(sizeof(double) * ne);
   d->cfx = (double *)zalloc(sizeof(double) * 8 * ne);
   d->cfy = (double *)zalloc(sizeof(double) * 8 * ne);
   d->cfz = (double *)zalloc(sizeof(double) * 8 * ne);
   d->vnewc = (double *)zalloc(sizeof(double) * ne);
   d->dxx = (double *)zalloc(sizeof(double) * ne);
   d->dyy = (double *)zalloc(sizeof(double) * ne);
   d->dzz = (double *)zalloc(sizeof(double) * ne);

   /* v starts at 1 (lulesh-init.cc:97-100); the rest are zero. */
   for (size_t i = 0; i < ne; ++i) F(d, ORC_V)[i] = 1.0;

   /* Node lattice, lulesh-init.cc:217-242. */
   double *X = F(d, ORC_X), *Y = F(d, ORC_Y), *Z = F(d, ORC_Z);
   {
      size_t n = 0;
      double tz = 1.125 * 0.0 / s;
      for (int k = 0; k < en; ++k) {
         double ty = 1.125 * 0.0 / s;
         for (int j = 0; j < en; ++j) {
            double tx = 1.125 * 0.0 / s;
            for (int i = 0; i < en; ++i) {
               X[n] = tx; Y[n] = ty; Z[n] = tz;
               ++n;
               tx = 1.125 * (double)(i + 1) / (double)s;
            }
            ty = 1.125 * (double)(j + 1) / (double)s;
         }
         tz = 1.125 * (double)(k + 1) / (double)s;
      }
   }
   /* Element-to-node map, lulesh-init.cc:245-266. */
   for (int k = 0, e = 0; k < s; ++k)
      for (int j = 0; j < s; ++j)
         for (int i = 0; i < s; ++i, ++e) {
            const int32_t n0 = (k * en + j) * en + i;
            int32_t *nl = d->nodelist + 8 * (size_t)e;
            nl[0] = n0;                 nl[1] = n0 + 1;
            nl[2] = n0 + en + 1;        nl[3] = n0 + en;
            nl[4] = n0 + en * en;       nl[5] = n0 + en * en + 1;
            nl[6] = n0 + en * en + en + 1; nl[7] = n0 + en * en + en;
         }

   /* Node-to-corner CSR, lulesh-init.cc:271-337. */
   d->nodeElemStart = (int32_t *)zalloc(sizeof(int32_t) * (nn + 1));
   {
      int32_t *cnt = (int32_t *)zalloc(sizeof(int32_t) * nn);
      for (size_t c = 0; c < 8 * ne; ++c) cnt[d->nodelist[c]]++;
      d->nodeElemStart[0] = 0;
      for (size_t n = 0; n < nn; ++n) d->nodeElemStart[n + 1] = d->nodeElemStart[n] + cnt[n];
      d->cornerList = (int32_t *)zalloc(sizeof(int32_t) * (size_t)d->nodeElemStart[nn]);
      memset(cnt, 0, sizeof(int32_t) * nn);
      for (size_t c = 0; c < 8 * ne; ++c) {
         const int32_t n = d->nodelist[c];
         d->cornerList[d->nodeElemStart[n] + cnt[n]++] = (int32_t)c;
      }
      free(cnt);
   }

   make_regions(d, nr, balance);

   /* Symmetry node sets, lulesh-init.cc:513-533. */
   for (int a = 0, n = 0; a < en; ++a)
      for (int b = 0; b < en; ++b, ++n) {
         d->symmZ[n] = a * en + b;
         d->symmY[n] = a * en * en + b;
         d->symmX[n] = a * en * en + b * en;
      }

   /* Face neighbours, lulesh-init.cc:538-565. */
   {
      const int E = d->numElem, s2 = s * s;
      d->lxim[0] = 0;
      for (int i = 1; i < E; ++i) { d->lxim[i] = i - 1; d->lxip[i - 1] = i; }
      d->lxip[E - 1] = E - 1;
      for (int i = 0; i < s; ++i) { d->letam[i] = i; d->letap[E - s + i] = E - s + i; }
      for (int i = s; i < E; ++i) { d->letam[i] = i - s; d->letap[i - s] = i; }
      for (int i = 0; i < s2; ++i) { d->lzetam[i] = i; d->lzetap[E - s2 + i] = E - s2 + i; }
      for (int i = s2; i < E; ++i) { d->lzetam[i] = i - s2; d->lzetap[i - s2] = i; }
   }

   /* Boundary masks of a single domain, lulesh-init.cc:568-673. */
   for (int a = 0; a < s; ++a)
      for (int b = 0; b < s; ++b) {
         const int plane = a * s * s, row = a * s;
         d->elemBC[row + b] |= BC_ZETA_M_SYMM;
         d->elemBC[row + b + d->numElem - s * s] |= BC_ZETA_P_FREE;
         d->elemBC[plane + b] |= BC_ETA_M_SYMM;
         d->elemBC[plane + b + s * s - s] |= BC_ETA_P_FREE;
         d->elemBC[plane + b * s] |= BC_XI_M_SYMM;
         d->elemBC[plane + b * s + s - 1] |= BC_XI_P_FREE;
      }

   /* Time controls, lulesh-init.cc:146-156. */
   d->sc[ORC_DTFIXED] = -1.0e-6;
   d->sc[ORC_STOPTIME] = 1.0e-2;
   d->sc[ORC_DELTATIMEMULTLB] = 1.1;
   d->sc[ORC_DELTATIMEMULTUB] = 1.2;
   d->sc[ORC_DTCOURANT] = 1.0e+20;
   d->sc[ORC_DTHYDRO] = 1.0e+20;
   d->sc[ORC_DTMAX] = 1.0e-2;
   d->sc[ORC_TIME] = 0.0;
   d->sc[ORC_CYCLE] = 0.0;

   /* Reference volume and masses, lulesh-init.cc:158-178. */
   for (int e = 0; e < d->numElem; ++e) {
      double xl[8], yl[8], zl[8];
      const int32_t *nl = d->nodelist + 8 * (size_t)e;
      gather8(X, nl, xl); gather8(Y, nl, yl); gather8(Z, nl, zl);
      const double vol = orc_elem_volume(xl, yl, zl);
      F(d, ORC_VOLO)[e] = vol;
      F(d, ORC_ELEMMASS)[e] = vol;
      for (int j = 0; j < 8; ++j) F(d, ORC_NODALMASS)[nl[j]] += vol / 8.0;
   }

   /* Energy deposit and first dt, lulesh-init.cc:180-192. */
   {
      const double ebase = 3.948746e+7;
      const double scale = s / 45.0;
      double einit = ebase * scale * scale * scale;
      if (emult != 1.0) einit = einit * emult;
      F(d, ORC_E)[0] = einit;
      d->sc[ORC_DELTATIME] = (0.5 * cbrt(F(d, ORC_VOLO)[0])) / sqrt(2.0 * einit);
   }
   return d;
}

void orc_destroy(orc_domain *d)
{
   if (!d) return;
   for (int k = 0; k < ORC_NUM_REAL_FIELDS; ++k) free(d->f[k]);
   int32_t *ints[] = {d->nodelist, d->lxim, d->lxip, d->letam, d->letap, d->lzetam,
                      d->lzetap, d->elemBC, d->nodeElemStart, d->cornerList, d->symmX,
                      d->symmY, d->symmZ, d->regNumList, d->regElemSize};
   for (size_t k = 0; k < sizeof(ints) / sizeof(ints[0]); ++k) free(ints[k]);
   for (int r = 0; r < d->numReg; ++r) free(d->regElemlist[r]);
   free(d->regElemlist);
   free(d->sig); free(d->determ); free(d->cfx); free(d->cfy); free(d->cfz);
   free(d->vnewc); free(d->dxx); free(d->dyy); free(d->dzz);
   free(d);
}

int orc_num_elem(const orc_domain *d) { return d->numElem; }
int orc_num_node(const orc_domain *d) { return d->numNode; }
int orc_num_reg(const orc_domain *d) { return d->numReg; }

/* EOS repetition count of region r, lulesh.cc:2390-2400. */
int orc_region_rep(const orc_domain *d, int r)
{
   if (r < d->numReg / 2) return 1;
   if (r < d->numReg - (d->numReg + 15) / 20) return 1 + d->cost;
   return 10 * (1 + d->cost);
}

double *orc_real(orc_domain *d, int field)
{
   return (field >= 0 && field < ORC_NUM_REAL_FIELDS) ? d->f[field] : NULL;
}

int32_t *orc_int(orc_domain *d, int field, int region)
{
   switch (field) {
   case ORC_NODELIST: return d->nodelist;
   case ORC_LXIM: return d->lxim;
   case ORC_LXIP: return d->lxip;
   case ORC_LETAM: return d->letam;
   case ORC_LETAP: return d->letap;
   case ORC_LZETAM: return d->lzetam;
   case ORC_LZETAP: return d->lzetap;
   case ORC_ELEMBC: return d->elemBC;
   case ORC_NODEELEMSTART: return d->nodeElemStart;
   case ORC_NODEELEMCORNERLIST: return d->cornerList;
   case ORC_SYMMX: return d->symmX;
   case ORC_SYMMY: return d->symmY;
   case ORC_SYMMZ: return d->symmZ;
   case ORC_REGNUMLIST: return d->regNumList;
   case ORC_REGELEMSIZE: return d->regElemSize;
   case ORC_REGELEMLIST:
      return (region >= 0 && region < d->numReg) ? d->regElemlist[region] : NULL;
   default: return NULL;
   }
}

int64_t orc_int_len(const orc_domain *d, int field, int region)
{
   const int64_t ne = d->numElem, en = d->s + 1;
   switch (field) {
   case ORC_NODELIST: return 8 * ne;
   case ORC_LXIM: case ORC_LXIP: case ORC_LETAM: case ORC_LETAP: case ORC_LZETAM:
   case ORC_LZETAP: case ORC_ELEMBC: case ORC_REGNUMLIST: return ne;
   case ORC_NODEELEMSTART: return d->numNode + 1;
   case ORC_NODEELEMCORNERLIST: return d->nodeElemStart[d->numNode];
   case ORC_SYMMX: case ORC_SYMMY: case ORC_SYMMZ: return en * en;
   case ORC_REGELEMSIZE: return d->numReg;
   case ORC_REGELEMLIST:
      return (region >= 0 && region < d->numReg) ? d->regElemSize[region] : 0;
   default: return 0;
   }
}

double orc_get(const orc_domain *d, int which)
{
   return (which >= 0 && which < ORC_NUM_SCALARS) ? d->sc[which] : 0.0;
}

void orc_set(orc_domain *d, int which, double value)
{
   if (which >= 0 && which < ORC_NUM_SCALARS) d->sc[which] = value;
}

void orc_set_force_order(orc_domain *d, int scatter) { d->scatter = scatter ? 1 : 0; }

/* ------------------------------------------------------------------ */
/* Time step control, lulesh.cc:166-222                                 */
/* ------------------------------------------------------------------ */

void orc_time_increment(orc_domain *d)
{
   double *sc = d->sc;
   double targetdt = sc[ORC_STOPTIME] - sc[ORC_TIME];
   if (sc[ORC_DTFIXED] <= 0.0 && sc[ORC_CYCLE] != 0.0) {
      const double olddt = sc[ORC_DELTATIME];
      double newdt = 1.0e+20;
      if (sc[ORC_DTCOURANT] < newdt) newdt = sc[ORC_DTCOURANT] / 2.0;
      if (sc[ORC_DTHYDRO] < newdt) newdt = sc[ORC_DTHYDRO] * 2.0 / 3.0;
      const double ratio = newdt / olddt;
      if (ratio >= 1.0) {
         if (ratio < sc[ORC_DELTATIMEMULTLB]) newdt = olddt;
         else if (ratio > sc[ORC_DELTATIMEMULTUB]) newdt = olddt * sc[ORC_DELTATIMEMULTUB];
      }
      if (newdt > sc[ORC_DTMAX]) newdt = sc[ORC_DTMAX];
      sc[ORC_DELTATIME] = newdt;
   }
   if (targetdt > sc[ORC_DELTATIME] && targetdt < 4.0 * sc[ORC_DELTATIME] / 3.0)
      targetdt = 2.0 * sc[ORC_DELTATIME] / 3.0;
   if (targetdt < sc[ORC_DELTATIME]) sc[ORC_DELTATIME] = targetdt;
   sc[ORC_TIME] += sc[ORC_DELTATIME];
   sc[ORC_CYCLE] += 1.0;
}

/* ------------------------------------------------------------------ */
/* Nodal forces, lulesh.cc:226-1135                                     */
/* ------------------------------------------------------------------ */

/* Sum per-corner values into nodes: in nodeElemCornerList order (gather)
 * or element order (scatter).  accumulate=0 overwrites, 1 adds. */
static void corners_to_nodes(orc_domain *d, int accumulate)
{
   double *fx = F(d, ORC_FX), *fy = F(d, ORC_FY), *fz = F(d, ORC_FZ);
   if (d->scatter) {
      for (int e = 0; e < d->numElem; ++e)
         for (int j = 0; j < 8; ++j) {
            const int32_t n = d->nodelist[8 * (size_t)e + j];
            fx[n] += d->cfx[8 * (size_t)e + j];
            fy[n] += d->cfy[8 * (size_t)e + j];
            fz[n] += d->cfz[8 * (size_t)e + j];
         }
      return;
   }
   for (int n = 0; n < d->numNode; ++n) {
      double sx = 0.0, sy = 0.0, sz = 0.0;
      for (int32_t c = d->nodeElemStart[n]; c < d->nodeElemStart[n + 1]; ++c) {
         const int32_t k = d->cornerList[c];
         sx += d->cfx[k];
         sy += d->cfy[k];
         sz += d->cfz[k];
      }
      if (accumulate) { fx[n] += sx; fy[n] += sy; fz[n] += sz; }
      else { fx[n] = sx; fy[n] = sy; fz[n] = sz; }
   }
}

/* IntegrateStressForElems, lulesh.cc:494-587 (stress from
 * InitStressTermsForElems, :273-286). */
static void integrate_stress(orc_domain *d)
{
   const double *X = F(d, ORC_X), *Y = F(d, ORC_Y), *Z = F(d, ORC_Z);
   for (int e = 0; e < d->numElem; ++e) {
      d->sig[e] = -F(d, ORC_P)[e] - F(d, ORC_Q)[e];
      double xl[8], yl[8], zl[8], b[3][4], nx[8], ny[8], nz[8];
      const int32_t *nl = d->nodelist + 8 * (size_t)e;
      gather8(X, nl, xl); gather8(Y, nl, yl); gather8(Z, nl, zl);
      shape_derivs(xl, yl, zl, b, &d->determ[e]);
      orc_elem_node_normals(xl, yl, zl, nx, ny, nz);
      const double s = d->sig[e];
      for (int j = 0; j < 8; ++j) {
         d->cfx[8 * (size_t)e + j] = -(s * nx[j]);
         d->cfy[8 * (size_t)e + j] = -(s * ny[j]);
         d->cfz[8 * (size_t)e + j] = -(s * nz[j]);
      }
   }
   corners_to_nodes(d, 0);
}

/* Hourglass base vectors, lulesh.cc:743-776. */
static const double kGamma[4][8] = {
   {1., 1., -1., -1., -1., -1., 1., 1.},
   {1., -1., -1., 1., -1., 1., 1., -1.},
   {1., -1., 1., -1., 1., -1., 1., -1.},
   {-1., 1., -1., 1., 1., -1., 1., -1.}};

/* CalcHourglassControlForElems + CalcFBHourglassForceForElems,
 * lulesh.cc:667-1057.  Returns VOLUME_ERROR if some v <= 0. */
static int hourglass(orc_domain *d)
{
   const double *X = F(d, ORC_X), *Y = F(d, ORC_Y), *Z = F(d, ORC_Z);
   const double *XD = F(d, ORC_XD), *YD = F(d, ORC_YD), *ZD = F(d, ORC_ZD);
   int err = 0;
   for (int e = 0; e < d->numElem; ++e) {
      d->determ[e] = F(d, ORC_VOLO)[e] * F(d, ORC_V)[e];
      if (F(d, ORC_V)[e] <= 0.0) err = VOLUME_ERROR;
   }
   if (err || !(d->hgcoef > 0.0)) return err;

   for (int e = 0; e < d->numElem; ++e) {
      double xl[8], yl[8], zl[8], dvx[8], dvy[8], dvz[8];
      const int32_t *nl = d->nodelist + 8 * (size_t)e;
      gather8(X, nl, xl); gather8(Y, nl, yl); gather8(Z, nl, zl);
      orc_elem_volume_derivative(xl, yl, zl, dvx, dvy, dvz);

      const double volinv = 1.0 / d->determ[e];
      double hg[8][4];
      for (int m = 0; m < 4; ++m) {
         double hx = xl[0] * kGamma[m][0], hy = yl[0] * kGamma[m][0], hz = zl[0] * kGamma[m][0];
         for (int j = 1; j < 8; ++j) {
            hx += xl[j] * kGamma[m][j];
            hy += yl[j] * kGamma[m][j];
            hz += zl[j] * kGamma[m][j];
         }
         for (int j = 0; j < 8; ++j)
            hg[j][m] = kGamma[m][j] - volinv * (dvx[j] * hx + dvy[j] * hy + dvz[j] * hz);
      }
      const double coef = -d->hgcoef * 0.01 * F(d, ORC_SS)[e] * F(d, ORC_ELEMMASS)[e] /
                          cbrt(d->determ[e]);

      double vel[3][8];
      gather8(XD, nl, vel[0]); gather8(YD, nl, vel[1]); gather8(ZD, nl, vel[2]);
      double *out[3] = {d->cfx + 8 * (size_t)e, d->cfy + 8 * (size_t)e, d->cfz + 8 * (size_t)e};
      for (int c = 0; c < 3; ++c) {
         double h[4];
         for (int m = 0; m < 4; ++m) {
            double acc = hg[0][m] * vel[c][0];
            for (int j = 1; j < 8; ++j) acc += hg[j][m] * vel[c][j];
            h[m] = acc;
         }
         for (int j = 0; j < 8; ++j)
            out[c][j] = coef * (hg[j][0] * h[0] + hg[j][1] * h[1] + hg[j][2] * h[2] + hg[j][3] * h[3]);
      }
   }
   corners_to_nodes(d, 1);
   return 0;
}

int orc_calc_force_for_nodes(orc_domain *d)
{
   memset(F(d, ORC_FX), 0, sizeof(double) * (size_t)d->numNode);
   memset(F(d, ORC_FY), 0, sizeof(double) * (size_t)d->numNode);
   memset(F(d, ORC_FZ), 0, sizeof(double) * (size_t)d->numNode);
   if (d->numElem == 0) return 0;
   integrate_stress(d);
   for (int e = 0; e < d->numElem; ++e)
      if (d->determ[e] <= 0.0) return VOLUME_ERROR; /* lulesh.cc:1082-1091 */
   return hourglass(d);
}

/* Acceleration, symmetry-plane BCs, velocity and position,
 * lulesh.cc:1139-1219. */
void orc_node_update(orc_domain *d)
{
   const int nn = d->numNode, nbc = (d->s + 1) * (d->s + 1);
   const double dt = d->sc[ORC_DELTATIME], ucut = d->u_cut;
   double *xdd = F(d, ORC_XDD), *ydd = F(d, ORC_YDD), *zdd = F(d, ORC_ZDD);
   for (int n = 0; n < nn; ++n) {
      xdd[n] = F(d, ORC_FX)[n] / F(d, ORC_NODALMASS)[n];
      ydd[n] = F(d, ORC_FY)[n] / F(d, ORC_NODALMASS)[n];
      zdd[n] = F(d, ORC_FZ)[n] / F(d, ORC_NODALMASS)[n];
   }
   for (int i = 0; i < nbc; ++i) {
      xdd[d->symmX[i]] = 0.0;
      ydd[d->symmY[i]] = 0.0;
      zdd[d->symmZ[i]] = 0.0;
   }
   double *acc[3] = {xdd, ydd, zdd};
   double *vel[3] = {F(d, ORC_XD), F(d, ORC_YD), F(d, ORC_ZD)};
   double *pos[3] = {F(d, ORC_X), F(d, ORC_Y), F(d, ORC_Z)};
   for (int n = 0; n < nn; ++n)
      for (int c = 0; c < 3; ++c) {
         double v = vel[c][n] + acc[c][n] * dt;
         if (fabs(v) < ucut) v = 0.0;
         vel[c][n] = v;
      }
   for (int n = 0; n < nn; ++n)
      for (int c = 0; c < 3; ++c) pos[c][n] += vel[c][n] * dt;
}

int orc_lagrange_nodal(orc_domain *d)
{
   const int err = orc_calc_force_for_nodes(d);
   if (err) return err;
   orc_node_update(d);
   return 0;
}

/* ------------------------------------------------------------------ */
/* Element kinematics, lulesh.cc:1504-1609                              */
/* ------------------------------------------------------------------ */

int orc_calc_lagrange_elements(orc_domain *d)
{
   const double dt = d->sc[ORC_DELTATIME];
   const double *X = F(d, ORC_X), *Y = F(d, ORC_Y), *Z = F(d, ORC_Z);
   const double *XD = F(d, ORC_XD), *YD = F(d, ORC_YD), *ZD = F(d, ORC_ZD);
   int err = 0;
   for (int e = 0; e < d->numElem; ++e) {
      double xl[8], yl[8], zl[8], xv[8], yv[8], zv[8], b[3][4], detJ = 0.0;
      const int32_t *nl = d->nodelist + 8 * (size_t)e;
      gather8(X, nl, xl); gather8(Y, nl, yl); gather8(Z, nl, zl);
      const double volume = orc_elem_volume(xl, yl, zl);
      const double rel = volume / F(d, ORC_VOLO)[e];
      F(d, ORC_VNEW)[e] = rel;
      F(d, ORC_DELV)[e] = rel - F(d, ORC_V)[e];
      F(d, ORC_AREALG)[e] = orc_elem_char_length(xl, yl, zl, volume);

      gather8(XD, nl, xv); gather8(YD, nl, yv); gather8(ZD, nl, zv);
      const double dt2 = 0.5 * dt;
      for (int j = 0; j < 8; ++j) {
         xl[j] -= dt2 * xv[j];
         yl[j] -= dt2 * yv[j];
         zl[j] -= dt2 * zv[j];
      }
      shape_derivs(xl, yl, zl, b, &detJ);
      const double inv = 1.0 / detJ;
      const double *vv[3] = {xv, yv, zv};
      double dd[3];
      for (int c = 0; c < 3; ++c)
         dd[c] = inv * (b[c][0] * (vv[c][0] - vv[c][6]) + b[c][1] * (vv[c][1] - vv[c][7]) +
                        b[c][2] * (vv[c][2] - vv[c][4]) + b[c][3] * (vv[c][3] - vv[c][5]));
      d->dxx[e] = dd[0]; d->dyy[e] = dd[1]; d->dzz[e] = dd[2];
   }
   for (int e = 0; e < d->numElem; ++e) {
      const double vdov = d->dxx[e] + d->dyy[e] + d->dzz[e];
      F(d, ORC_VDOV)[e] = vdov;
      if (F(d, ORC_VNEW)[e] <= 0.0) err = VOLUME_ERROR;
   }
   return err;
}

/* ------------------------------------------------------------------ */
/* Monotonic Q, lulesh.cc:1613-2010                                     */
/* ------------------------------------------------------------------ */

/* ((a0+a1)+a2)+a3 - (((b0+b1)+b2)+b3) over node values c. */
static double face_diff(const double c[8], int a0, int a1, int a2, int a3,
                        int b0, int b1, int b2, int b3)
{
   return (c[a0] + c[a1] + c[a2] + c[a3]) - (c[b0] + c[b1] + c[b2] + c[b3]);
}

void orc_calc_monotonic_q_gradients(orc_domain *d)
{
   const double ptiny = 1.e-36;
   for (int e = 0; e < d->numElem; ++e) {
      double x[8], y[8], z[8], xv[8], yv[8], zv[8];
      const int32_t *nl = d->nodelist + 8 * (size_t)e;
      gather8(F(d, ORC_X), nl, x); gather8(F(d, ORC_Y), nl, y); gather8(F(d, ORC_Z), nl, z);
      gather8(F(d, ORC_XD), nl, xv); gather8(F(d, ORC_YD), nl, yv); gather8(F(d, ORC_ZD), nl, zv);

      const double vol = F(d, ORC_VOLO)[e] * F(d, ORC_VNEW)[e];
      const double norm = 1.0 / (vol + ptiny);

      /* element-centred differences across the eta, xi and zeta faces */
      const double dxj = -0.25 * face_diff(x, 0, 1, 5, 4, 3, 2, 6, 7);
      const double dyj = -0.25 * face_diff(y, 0, 1, 5, 4, 3, 2, 6, 7);
      const double dzj = -0.25 * face_diff(z, 0, 1, 5, 4, 3, 2, 6, 7);
      const double dxi = 0.25 * face_diff(x, 1, 2, 6, 5, 0, 3, 7, 4);
      const double dyi = 0.25 * face_diff(y, 1, 2, 6, 5, 0, 3, 7, 4);
      const double dzi = 0.25 * face_diff(z, 1, 2, 6, 5, 0, 3, 7, 4);
      const double dxk = 0.25 * face_diff(x, 4, 5, 6, 7, 0, 1, 2, 3);
      const double dyk = 0.25 * face_diff(y, 4, 5, 6, 7, 0, 1, 2, 3);
      const double dzk = 0.25 * face_diff(z, 4, 5, 6, 7, 0, 1, 2, 3);

      double ax, ay, az, dxv, dyv, dzv;

      /* zeta: i x j */
      ax = dyi * dzj - dzi * dyj;
      ay = dzi * dxj - dxi * dzj;
      az = dxi * dyj - dyi * dxj;
      F(d, ORC_DELX_ZETA)[e] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
      ax *= norm; ay *= norm; az *= norm;
      dxv = 0.25 * face_diff(xv, 4, 5, 6, 7, 0, 1, 2, 3);
      dyv = 0.25 * face_diff(yv, 4, 5, 6, 7, 0, 1, 2, 3);
      dzv = 0.25 * face_diff(zv, 4, 5, 6, 7, 0, 1, 2, 3);
      F(d, ORC_DELV_ZETA)[e] = ax * dxv + ay * dyv + az * dzv;

      /* xi: j x k */
      ax = dyj * dzk - dzj * dyk;
      ay = dzj * dxk - dxj * dzk;
      az = dxj * dyk - dyj * dxk;
      F(d, ORC_DELX_XI)[e] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
      ax *= norm; ay *= norm; az *= norm;
      dxv = 0.25 * face_diff(xv, 1, 2, 6, 5, 0, 3, 7, 4);
      dyv = 0.25 * face_diff(yv, 1, 2, 6, 5, 0, 3, 7, 4);
      dzv = 0.25 * face_diff(zv, 1, 2, 6, 5, 0, 3, 7, 4);
      F(d, ORC_DELV_XI)[e] = ax * dxv + ay * dyv + az * dzv;

      /* eta: k x i */
      ax = dyk * dzi - dzk * dyi;
      ay = dzk * dxi - dxk * dzi;
      az = dxk * dyi - dyk * dxi;
      F(d, ORC_DELX_ETA)[e] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
      ax *= norm; ay *= norm; az *= norm;
      dxv = -0.25 * face_diff(xv, 0, 1, 5, 4, 3, 2, 6, 7);
      dyv = -0.25 * face_diff(yv, 0, 1, 5, 4, 3, 2, 6, 7);
      dzv = -0.25 * face_diff(zv, 0, 1, 5, 4, 3, 2, 6, 7);
      F(d, ORC_DELV_ETA)[e] = ax * dxv + ay * dyv + az * dzv;
   }
}

/* Neighbour gradient across one face pair under the BC mask,
 * lulesh.cc:1781-1800 (and the eta/zeta copies). */
static double bc_neighbour(const double *g, int e, int nbr, int bits, int symm, int freeb, int comm)
{
   if (bits == comm || bits == 0) return g[nbr];
   if (bits == symm) return g[e];
   if (bits == freeb) return 0.0;
   return 0.0;
}

/* Limiter of one direction, lulesh.cc:1802-1813. */
static double limiter(double own, double m, double p, double mult, double maxslope, double ptiny)
{
   const double norm = 1.0 / (own + ptiny);
   double dm = m * norm, dp = p * norm;
   double phi = 0.5 * (dm + dp);
   dm *= mult;
   dp *= mult;
   if (dm < phi) phi = dm;
   if (dp < phi) phi = dp;
   if (phi < 0.0) phi = 0.0;
   if (phi > maxslope) phi = maxslope;
   return phi;
}

static void monotonic_q_region(orc_domain *d, int r)
{
   const double ptiny = 1.e-36;
   const double *dvx = F(d, ORC_DELV_XI), *dve = F(d, ORC_DELV_ETA), *dvz = F(d, ORC_DELV_ZETA);
   for (int i = 0; i < d->regElemSize[r]; ++i) {
      const int e = d->regElemlist[r][i];
      const int bc = d->elemBC[e];
      const double phixi = limiter(
         dvx[e],
         bc_neighbour(dvx, e, d->lxim[e], bc & BC_XI_M, BC_XI_M_SYMM, BC_XI_M_FREE, BC_XI_M_COMM),
         bc_neighbour(dvx, e, d->lxip[e], bc & BC_XI_P, BC_XI_P_SYMM, BC_XI_P_FREE, BC_XI_P_COMM),
         d->monoq_limiter_mult, d->monoq_max_slope, ptiny);
      const double phieta = limiter(
         dve[e],
         bc_neighbour(dve, e, d->letam[e], bc & BC_ETA_M, BC_ETA_M_SYMM, BC_ETA_M_FREE, BC_ETA_M_COMM),
         bc_neighbour(dve, e, d->letap[e], bc & BC_ETA_P, BC_ETA_P_SYMM, BC_ETA_P_FREE, BC_ETA_P_COMM),
         d->monoq_limiter_mult, d->monoq_max_slope, ptiny);
      const double phizeta = limiter(
         dvz[e],
         bc_neighbour(dvz, e, d->lzetam[e], bc & BC_ZETA_M, BC_ZETA_M_SYMM, BC_ZETA_M_FREE, BC_ZETA_M_COMM),
         bc_neighbour(dvz, e, d->lzetap[e], bc & BC_ZETA_P, BC_ZETA_P_SYMM, BC_ZETA_P_FREE, BC_ZETA_P_COMM),
         d->monoq_limiter_mult, d->monoq_max_slope, ptiny);

      double qlin = 0.0, qquad = 0.0;
      if (!(F(d, ORC_VDOV)[e] > 0.0)) {
         double a = dvx[e] * F(d, ORC_DELX_XI)[e];
         double b = dve[e] * F(d, ORC_DELX_ETA)[e];
         double c = dvz[e] * F(d, ORC_DELX_ZETA)[e];
         if (a > 0.0) a = 0.0;
         if (b > 0.0) b = 0.0;
         if (c > 0.0) c = 0.0;
         const double rho = F(d, ORC_ELEMMASS)[e] / (F(d, ORC_VOLO)[e] * F(d, ORC_VNEW)[e]);
         qlin = -d->qlc_monoq * rho *
                (a * (1.0 - phixi) + b * (1.0 - phieta) + c * (1.0 - phizeta));
         qquad = d->qqc_monoq * rho *
                 (a * a * (1.0 - phixi * phixi) + b * b * (1.0 - phieta * phieta) +
                  c * c * (1.0 - phizeta * phizeta));
      }
      F(d, ORC_QQ)[e] = qquad;
      F(d, ORC_QL)[e] = qlin;
   }
}

int orc_calc_q(orc_domain *d)
{
   if (d->numElem == 0) return 0;
   orc_calc_monotonic_q_gradients(d);
   for (int r = 0; r < d->numReg; ++r)
      if (d->regElemSize[r] > 0) monotonic_q_region(d, r);
   for (int e = 0; e < d->numElem; ++e)
      if (F(d, ORC_Q)[e] > d->qstop) return QSTOP_ERROR; /* lulesh.cc:1993-2008 */
   return 0;
}

/* ------------------------------------------------------------------ */
/* Equation of state, lulesh.cc:2014-2406                               */
/* ------------------------------------------------------------------ */

/* CalcPressureForElems for one element, lulesh.cc:2014-2044. */
static double pressure(double e, double comp, double vnewc, const orc_domain *d,
                       double *bvc, double *pbvc)
{
   const double c1s = 2.0 / 3.0;
   *bvc = c1s * (comp + 1.0);
   *pbvc = c1s;
   double p = *bvc * e;
   if (fabs(p) < d->p_cut) p = 0.0;
   if (vnewc >= d->eosvmax) p = 0.0;
   if (p < d->pmin) p = d->pmin;
   return p;
}

/* Artificial-viscosity sound-speed term of lulesh.cc:2083-2090 and its
 * repeats, also CalcSoundSpeedForElems (:2190-2197). */
static double q_sound(double pbvc, double e, double vv, double bvc, double p, double rho0)
{
   double ssc = (pbvc * e + vv * vv * bvc * p) / rho0;
   if (ssc <= .1111111e-36) ssc = .3333333e-18;
   else ssc = sqrt(ssc);
   return ssc;
}

struct eos_state { double p_new, e_new, q_new, bvc, pbvc; };

/* One pass of EvalEOSForElems' repeated body for one element:
 * gather and compressions (lulesh.cc:2243-2287) then CalcEnergyForElems
 * (:2048-2176). */
static void eos_pass(const orc_domain *d, int e, double vnewc, struct eos_state *o)
{
   const double e_old = F(d, ORC_E)[e], delvc = F(d, ORC_DELV)[e];
   double p_old = F(d, ORC_P)[e];
   const double q_old = F(d, ORC_Q)[e], qq_old = F(d, ORC_QQ)[e], ql_old = F(d, ORC_QL)[e];
   const double rho0 = d->refdens, work = 0.0;

   double compression = 1.0 / vnewc - 1.0;
   const double vchalf = vnewc - delvc * 0.5;
   double compHalfStep = 1.0 / vchalf - 1.0;
   if (d->eosvmin != 0.0 && vnewc <= d->eosvmin) compHalfStep = compression;
   if (d->eosvmax != 0.0 && vnewc >= d->eosvmax) {
      p_old = 0.0;
      compression = 0.0;
      compHalfStep = 0.0;
   }

   double bvc, pbvc;
   double e_new = e_old - 0.5 * delvc * (p_old + q_old) + 0.5 * work;
   if (e_new < d->emin) e_new = d->emin;

   const double pHalfStep = pressure(e_new, compHalfStep, vnewc, d, &bvc, &pbvc);
   const double vhalf = 1.0 / (1.0 + compHalfStep);
   double q_new;
   if (delvc > 0.0) q_new = 0.0;
   else q_new = q_sound(pbvc, e_new, vhalf, bvc, pHalfStep, rho0) * ql_old + qq_old;

   e_new = e_new + 0.5 * delvc * (3.0 * (p_old + q_old) - 4.0 * (pHalfStep + q_new));
   e_new += 0.5 * work;
   if (fabs(e_new) < d->e_cut) e_new = 0.0;
   if (e_new < d->emin) e_new = d->emin;

   double p_new = pressure(e_new, compression, vnewc, d, &bvc, &pbvc);
   double q_tilde;
   if (delvc > 0.0) q_tilde = 0.0;
   else q_tilde = q_sound(pbvc, e_new, vnewc, bvc, p_new, rho0) * ql_old + qq_old;

   const double sixth = 1.0 / 6.0;
   e_new = e_new - (7.0 * (p_old + q_old) - 8.0 * (pHalfStep + q_new) + (p_new + q_tilde)) *
                      delvc * sixth;
   if (fabs(e_new) < d->e_cut) e_new = 0.0;
   if (e_new < d->emin) e_new = d->emin;

   p_new = pressure(e_new, compression, vnewc, d, &bvc, &pbvc);
   if (delvc <= 0.0) {
      q_new = q_sound(pbvc, e_new, vnewc, bvc, p_new, rho0) * ql_old + qq_old;
      if (fabs(q_new) < d->q_cut) q_new = 0.0;
   }
   o->p_new = p_new;
   o->e_new = e_new;
   o->q_new = q_new;
   o->bvc = bvc;
   o->pbvc = pbvc;
}

/* EvalEOSForElems for region r, lulesh.cc:2204-2324.  The repeated passes
 * read unchanged inputs (p,e,q are written only after the repeat loop), so
 * every pass yields the same state; all of them are still computed. */
static void eval_eos_region(orc_domain *d, int r, int rep)
{
   for (int i = 0; i < d->regElemSize[r]; ++i) {
      const int e = d->regElemlist[r][i];
      struct eos_state st = {0, 0, 0, 0, 0};
      for (int j = 0; j < rep; ++j) eos_pass(d, e, d->vnewc[e], &st);
      F(d, ORC_P)[e] = st.p_new;
      F(d, ORC_E)[e] = st.e_new;
      F(d, ORC_Q)[e] = st.q_new;
      F(d, ORC_SS)[e] = q_sound(st.pbvc, st.e_new, d->vnewc[e], st.bvc, st.p_new, d->refdens);
   }
}

int orc_apply_material_properties(orc_domain *d)
{
   if (d->numElem == 0) return 0;
   for (int e = 0; e < d->numElem; ++e) {
      double v = F(d, ORC_VNEW)[e];
      if (d->eosvmin != 0.0 && v < d->eosvmin) v = d->eosvmin;
      if (d->eosvmax != 0.0 && v > d->eosvmax) v = d->eosvmax;
      d->vnewc[e] = v;
   }
   for (int e = 0; e < d->numElem; ++e) {
      double vc = F(d, ORC_V)[e];
      if (d->eosvmin != 0.0 && vc < d->eosvmin) vc = d->eosvmin;
      if (d->eosvmax != 0.0 && vc > d->eosvmax) vc = d->eosvmax;
      if (vc <= 0.0) return VOLUME_ERROR; /* lulesh.cc:2377-2383 */
   }
   for (int r = 0; r < d->numReg; ++r) eval_eos_region(d, r, orc_region_rep(d, r));
   return 0;
}

void orc_update_volumes(orc_domain *d)
{
   for (int e = 0; e < d->numElem; ++e) {
      double v = F(d, ORC_VNEW)[e];
      if (fabs(v - 1.0) < d->v_cut) v = 1.0;
      F(d, ORC_V)[e] = v;
   }
}

int orc_lagrange_elements(orc_domain *d)
{
   int err = orc_calc_lagrange_elements(d);
   if (err) return err;
   err = orc_calc_q(d);
   if (err) return err;
   err = orc_apply_material_properties(d);
   if (err) return err;
   orc_update_volumes(d);
   return 0;
}

/* ------------------------------------------------------------------ */
/* Time constraints, lulesh.cc:2447-2596                                */
/* ------------------------------------------------------------------ */

void orc_calc_time_constraints(orc_domain *d)
{
   double dtc = 1.0e+20, dth = 1.0e+20;
   const double qqc2 = 64.0 * d->qqc * d->qqc;
   const double *ss = F(d, ORC_SS), *vdov = F(d, ORC_VDOV), *ar = F(d, ORC_AREALG);
   for (int r = 0; r < d->numReg; ++r)
      for (int i = 0; i < d->regElemSize[r]; ++i) {
         const int e = d->regElemlist[r][i];
         double dtf = ss[e] * ss[e];
         if (vdov[e] < 0.0) dtf = dtf + qqc2 * ar[e] * ar[e] * vdov[e] * vdov[e];
         dtf = sqrt(dtf);
         dtf = ar[e] / dtf;
         if (vdov[e] != 0.0 && dtf < dtc) dtc = dtf;
         if (vdov[e] != 0.0) {
            const double dtdvov = d->dvovmax / (fabs(vdov[e]) + 1.e-20);
            if (dth > dtdvov) dth = dtdvov;
         }
      }
   d->sc[ORC_DTCOURANT] = dtc;
   d->sc[ORC_DTHYDRO] = dth;
}

int orc_lagrange_leapfrog(orc_domain *d)
{
   int err = orc_lagrange_nodal(d);
   if (err) return err;
   err = orc_lagrange_elements(d);
   if (err) return err;
   orc_calc_time_constraints(d);
   return 0;
}

int orc_run(orc_domain *d, int its)
{
   while (d->sc[ORC_TIME] < d->sc[ORC_STOPTIME] && d->sc[ORC_CYCLE] < its) {
      orc_time_increment(d);
      const int err = orc_lagrange_leapfrog(d);
      if (err) return err;
   }
   return 0;
}

void orc_symmetry(const orc_domain *d, double out[3])
{
   const int nx = d->s;
   const double *e = d->f[ORC_E];
   double maxAbs = 0.0, tot = 0.0, maxRel = 0.0;
   for (int j = 0; j < nx; ++j)
      for (int k = j + 1; k < nx; ++k) {
         const double ad = fabs(e[j * nx + k] - e[k * nx + j]);
         tot += ad;
         if (maxAbs < ad) maxAbs = ad;
         const double rd = ad / e[k * nx + j];
         if (maxRel < rd) maxRel = rd;
      }
   out[0] = maxAbs;
   out[1] = tot;
   out[2] = maxRel;
}
