// lgpu_geom.h -- per-element geometry of the hexahedral Lagrange step, as
// register-resident device functions.
//
// Every function keeps the operand association of the reference line it
// cites, so that the FP64 instantiation (built without FMA contraction)
// rounds exactly as the reference does.  T is double for the FP64 path and
// float for the FP32 parts of the mixed path.
//
// Hex node numbering (lulesh-init.cc:252-259): 0 (i,j,k), 1 (i+1,j,k),
// 2 (i+1,j+1,k), 3 (i,j+1,k), 4..7 the same at k+1.
#pragma once

#include <hip/hip_runtime.h>

namespace lgpu {

#define LG_DEV __device__ __forceinline__

// Reciprocal, square root and a / sqrt(x): correctly rounded in FP64 (the
// reference's operations); the hardware approximations (about 1 ulp) in the
// FP32 arithmetic of the mixed form.
LG_DEV double recip(double x) { return 1.0 / x; }
LG_DEV float recip(float x) { return __builtin_amdgcn_rcpf(x); }
LG_DEV double root(double x) { return sqrt(x); }
LG_DEV float root(float x) { return __builtin_amdgcn_sqrtf(x); }
LG_DEV double div_root(double a, double x) { return a / sqrt(x); }
LG_DEV float div_root(float a, float x) { return a * __builtin_amdgcn_rsqf(x); }

// Correctly rounded FP64 reciprocal and square root for arguments that need
// none of the general routines' range handling.  Both refine the hardware
// estimate (v_rcp_f64 / v_rsq_f64) by Newton-Raphson steps and finish with
// one FMA correction of the rounded result (Markstein, IBM J. Res. Dev.
// 34(1), 1990): the steps of the general routines, without their exponent
// scaling and special-value fix-ups, so inside the stated ranges the results
// are the correctly rounded ones, bit for bit those of 1.0 / x and sqrt(x).
//   recip_rn_normal: 2^-400 <= |x| <= 2^400
//   sqrt_rn_normal:  2^-767 <= x < +inf
LG_DEV double recip_rn_normal(double x)
{
   double r = __builtin_amdgcn_rcp(x);
   double e = __builtin_fma(-x, r, 1.0);
   r = __builtin_fma(r, e, r);
   e = __builtin_fma(-x, r, 1.0);
   r = __builtin_fma(r, e, r);
   e = __builtin_fma(-x, r, 1.0);
   return __builtin_fma(e, r, r);
}
LG_DEV double sqrt_rn_normal(double x)
{
   const double y = __builtin_amdgcn_rsq(x);
   double g = x * y, h = y * 0.5;
   const double r = __builtin_fma(-h, g, 0.5);
   g = __builtin_fma(g, r, g);
   h = __builtin_fma(h, r, h);
   double d = __builtin_fma(-g, g, x);
   g = __builtin_fma(d, h, g);
   d = __builtin_fma(-g, g, x);
   return __builtin_fma(d, h, g);
}

// u . (v x w); the triple product of lulesh.cc:1337-1338 as it is applied
// at lulesh.cc:1341-1349.
template <typename T>
LG_DEV T triple(T ux, T uy, T uz, T vx, T vy, T vz, T wx, T wy, T wz)
{
   return ux * (vy * wz - vz * wy) + uy * (vz * wx - vx * wz) + uz * (vx * wy - vy * wx);
}

// Hexahedron volume from its node coordinates, lulesh.cc:1273-1356.
template <typename T>
LG_DEV T hex_volume(const T (&x)[8], const T (&y)[8], const T (&z)[8])
{
   const T t1 = triple<T>((x[3] - x[1]) + (x[7] - x[2]), (y[3] - y[1]) + (y[7] - y[2]),
                          (z[3] - z[1]) + (z[7] - z[2]),
                          x[6] - x[3], y[6] - y[3], z[6] - z[3],
                          x[2] - x[0], y[2] - y[0], z[2] - z[0]);
   const T t2 = triple<T>((x[4] - x[3]) + (x[5] - x[7]), (y[4] - y[3]) + (y[5] - y[7]),
                          (z[4] - z[3]) + (z[5] - z[7]),
                          x[6] - x[4], y[6] - y[4], z[6] - z[4],
                          x[7] - x[0], y[7] - y[0], z[7] - z[0]);
   const T t3 = triple<T>((x[1] - x[4]) + (x[2] - x[5]), (y[1] - y[4]) + (y[2] - y[5]),
                          (z[1] - z[4]) + (z[2] - z[5]),
                          x[6] - x[1], y[6] - y[1], z[6] - z[1],
                          x[5] - x[0], y[5] - y[0], z[5] - z[0]);
   T v = t1 + t2 + t3;
   v *= T(1.0) / T(12.0);
   return v;
}

// Scaled forms.  Several reference formulas multiply by a power of two
// (.125, .5, .25, 8) before further products and sums.  Such a factor is
// exact, and rounding commutes with it: RN(2^k a) = 2^k RN(a), and so for
// products, sums and quotients of scaled operands.  The *N functions below
// leave these factors out and return N times the reference's quantity, bit
// for bit N times it as long as no intermediate value leaves the normal
// range (subnormal or infinite, which no mesh of sensible extent comes near);
// their callers put the factor back in one place (or nowhere, where it
// cancels or only a sign is used).

// Centre Jacobian of the trilinear map, one physical coordinate: 8 times the
// three entries of lulesh.cc:319-329 for that coordinate (without their .125).
template <typename T>
struct JacRow {
   T xi, eta, zeta;
};

template <typename T>
LG_DEV JacRow<T> jac_row8(const T (&c)[8])
{
   const T a = c[6] - c[0], b = c[5] - c[3], g = c[7] - c[1], h = c[4] - c[2];
   JacRow<T> r;
   r.xi = a + b - g - h;
   r.eta = a - b + g - h;
   r.zeta = a + b + g + h;
   return r;
}

// 64 times the Jacobian determinant ("volume") of
// CalcElemShapeFunctionDerivatives, lulesh.cc:332-376, without the
// derivative rows (IntegrateStressForElems overwrites them with face normals,
// lulesh.cc:534-538, and uses only the determinant's sign).
template <typename T>
LG_DEV T jacobian_volume64(const T (&x)[8], const T (&y)[8], const T (&z)[8])
{
   const JacRow<T> jx = jac_row8(x), jy = jac_row8(y), jz = jac_row8(z);
   const T cxet = (jz.xi * jy.zeta) - (jy.xi * jz.zeta);
   const T cyet = (jx.xi * jz.zeta) - (jz.xi * jx.zeta);
   const T czet = (jy.xi * jx.zeta) - (jx.xi * jy.zeta);
   return jx.eta * cxet + jy.eta * cyet + jz.eta * czet;
}

// Shape-function derivatives of nodes 0..3 (4..7 are their negatives,
// lulesh.cc:344-373) and the Jacobian determinant, lulesh.cc:290-377, both
// 64 times the reference's (the quotient of the two is the reference's).
template <typename T>
LG_DEV T shape_derivatives64(const T (&x)[8], const T (&y)[8], const T (&z)[8], T (&b)[3][4])
{
   const JacRow<T> jx = jac_row8(x), jy = jac_row8(y), jz = jac_row8(z);
   const T c[3][3] = {
      {(jy.eta * jz.zeta) - (jz.eta * jy.zeta), (jz.xi * jy.zeta) - (jy.xi * jz.zeta),
       (jy.xi * jz.eta) - (jz.xi * jy.eta)},
      {(jz.eta * jx.zeta) - (jx.eta * jz.zeta), (jx.xi * jz.zeta) - (jz.xi * jx.zeta),
       (jz.xi * jx.eta) - (jx.xi * jz.eta)},
      {(jx.eta * jy.zeta) - (jy.eta * jx.zeta), (jy.xi * jx.zeta) - (jx.xi * jy.zeta),
       (jx.xi * jy.eta) - (jy.xi * jx.eta)}};
#pragma unroll
   for (int r = 0; r < 3; ++r) {
      b[r][0] = -c[r][0] - c[r][1] - c[r][2];
      b[r][1] = c[r][0] - c[r][1] - c[r][2];
      b[r][2] = c[r][0] + c[r][1] - c[r][2];
      b[r][3] = -c[r][0] + c[r][1] - c[r][2];
   }
   return jx.eta * c[0][1] + jy.eta * c[1][1] + jz.eta * c[2][1];
}

// 16 times the area-weighted normal of one quadrilateral face (n0,n1,n2,n3),
// added to its four nodes: SumElemFaceNormal, lulesh.cc:381-415, without the
// .5 of the bisectors and the .25 of the area.
template <typename T, int N0, int N1, int N2, int N3>
LG_DEV void add_face_normal16(const T (&x)[8], const T (&y)[8], const T (&z)[8], T (&nx)[8],
                              T (&ny)[8], T (&nz)[8])
{
   const T b0x = x[N3] + x[N2] - x[N1] - x[N0];
   const T b0y = y[N3] + y[N2] - y[N1] - y[N0];
   const T b0z = z[N3] + z[N2] - z[N1] - z[N0];
   const T b1x = x[N2] + x[N1] - x[N3] - x[N0];
   const T b1y = y[N2] + y[N1] - y[N3] - y[N0];
   const T b1z = z[N2] + z[N1] - z[N3] - z[N0];
   const T ax = b0y * b1z - b0z * b1y;
   const T ay = b0z * b1x - b0x * b1z;
   const T az = b0x * b1y - b0y * b1x;
   nx[N0] += ax; nx[N1] += ax; nx[N2] += ax; nx[N3] += ax;
   ny[N0] += ay; ny[N1] += ay; ny[N2] += ay; ny[N3] += ay;
   nz[N0] += az; nz[N1] += az; nz[N2] += az; nz[N3] += az;
}

// 16 times CalcElemNodeNormals, lulesh.cc:419-474 (faces in the reference's
// order, so each node's three contributions are summed in the same order).
template <typename T>
LG_DEV void node_normals16(const T (&x)[8], const T (&y)[8], const T (&z)[8], T (&nx)[8],
                           T (&ny)[8], T (&nz)[8])
{
#pragma unroll
   for (int i = 0; i < 8; ++i) nx[i] = ny[i] = nz[i] = T(0);
   add_face_normal16<T, 0, 1, 2, 3>(x, y, z, nx, ny, nz);
   add_face_normal16<T, 0, 4, 5, 1>(x, y, z, nx, ny, nz);
   add_face_normal16<T, 1, 5, 6, 2>(x, y, z, nx, ny, nz);
   add_face_normal16<T, 2, 6, 7, 3>(x, y, z, nx, ny, nz);
   add_face_normal16<T, 3, 7, 4, 0>(x, y, z, nx, ny, nz);
   add_face_normal16<T, 4, 7, 6, 5>(x, y, z, nx, ny, nz);
}

// Bilinear six-node form of VoluDer (lulesh.cc:602-614) for one node whose
// six neighbours are (A0..A5): dV/dx = P(y,z)/12, dV/dy = -P(x,z)/12,
// dV/dz = -P(y,x)/12.
template <typename T, int A0, int A1, int A2, int A3, int A4, int A5>
LG_DEV T voluder_form(const T (&a)[8], const T (&b)[8])
{
   return (a[A1] + a[A2]) * (b[A0] + b[A1]) - (a[A0] + a[A1]) * (b[A1] + b[A2]) +
          (a[A0] + a[A4]) * (b[A3] + b[A4]) - (a[A3] + a[A4]) * (b[A0] + b[A4]) -
          (a[A2] + a[A5]) * (b[A3] + b[A5]) + (a[A3] + a[A5]) * (b[A2] + b[A5]);
}

// kTwelfth = false leaves out the factor 1/12 (the compact forms fold it
// into the one scalar that multiplies the derivatives).
template <typename T, bool kTwelfth, int N, int A0, int A1, int A2, int A3, int A4, int A5>
LG_DEV void voluder_node(const T (&x)[8], const T (&y)[8], const T (&z)[8], T (&dx)[8],
                         T (&dy)[8], T (&dz)[8])
{
   if constexpr (kTwelfth) {
      const T twelfth = T(1.0) / T(12.0);
      dx[N] = voluder_form<T, A0, A1, A2, A3, A4, A5>(y, z) * twelfth;
      dy[N] = -(voluder_form<T, A0, A1, A2, A3, A4, A5>(x, z) * twelfth);
      dz[N] = -(voluder_form<T, A0, A1, A2, A3, A4, A5>(y, x) * twelfth);
   } else {
      dx[N] = voluder_form<T, A0, A1, A2, A3, A4, A5>(y, z);
      dy[N] = -voluder_form<T, A0, A1, A2, A3, A4, A5>(x, z);
      dz[N] = -voluder_form<T, A0, A1, A2, A3, A4, A5>(y, x);
   }
}

// CalcElemVolumeDerivative, lulesh.cc:623-663 (neighbour permutations of
// each node as the reference passes them to VoluDer); 12x the derivatives
// when kTwelfth is false.
template <typename T, bool kTwelfth = true>
LG_DEV void volume_derivative(const T (&x)[8], const T (&y)[8], const T (&z)[8], T (&dx)[8],
                              T (&dy)[8], T (&dz)[8])
{
   voluder_node<T, kTwelfth, 0, 1, 2, 3, 4, 5, 7>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 3, 0, 1, 2, 7, 4, 6>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 2, 3, 0, 1, 6, 7, 5>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 1, 2, 3, 0, 5, 6, 4>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 4, 7, 6, 5, 0, 3, 1>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 5, 4, 7, 6, 1, 0, 2>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 6, 5, 4, 7, 2, 1, 3>(x, y, z, dx, dy, dz);
   voluder_node<T, kTwelfth, 7, 6, 5, 4, 3, 2, 0>(x, y, z, dx, dy, dz);
}

// Squared-area form of one face, AreaFace (lulesh.cc:1370-1390).
template <typename T, int A, int B, int C, int D>
LG_DEV T face_area2(const T (&x)[8], const T (&y)[8], const T (&z)[8])
{
   const T fx = (x[C] - x[A]) - (x[D] - x[B]);
   const T fy = (y[C] - y[A]) - (y[D] - y[B]);
   const T fz = (z[C] - z[A]) - (z[D] - z[B]);
   const T gx = (x[C] - x[A]) + (x[D] - x[B]);
   const T gy = (y[C] - y[A]) + (y[D] - y[B]);
   const T gz = (z[C] - z[A]) + (z[D] - z[B]);
   const T fg = fx * gx + fy * gy + fz * gz;
   return (fx * fx + fy * fy + fz * fz) * (gx * gx + gy * gy + gz * gz) - fg * fg;
}

template <typename T>
LG_DEV T keep_max(T a, T m)
{
   return (a < m) ? m : a; // std::max(a, charLength), lulesh.cc:1405
}

// CalcElemCharacteristicLength, lulesh.cc:1394-1435.
template <typename T>
LG_DEV T char_length(const T (&x)[8], const T (&y)[8], const T (&z)[8], T volume)
{
   T m = T(0);
   m = keep_max(face_area2<T, 0, 1, 2, 3>(x, y, z), m);
   m = keep_max(face_area2<T, 4, 5, 6, 7>(x, y, z), m);
   m = keep_max(face_area2<T, 0, 1, 5, 4>(x, y, z), m);
   m = keep_max(face_area2<T, 1, 2, 6, 5>(x, y, z), m);
   m = keep_max(face_area2<T, 2, 3, 7, 6>(x, y, z), m);
   m = keep_max(face_area2<T, 3, 0, 4, 7>(x, y, z), m);
   return div_root(T(4.0) * volume, m);
}

// Difference of two four-node sums ((a0+a1)+a2)+a3 - (((b0+b1)+b2)+b3), the
// face-centred differences of lulesh.cc:1691-1753.
template <typename T, int A0, int A1, int A2, int A3, int B0, int B1, int B2, int B3>
LG_DEV T face_diff(const T (&c)[8])
{
   return (c[A0] + c[A1] + c[A2] + c[A3]) - (c[B0] + c[B1] + c[B2] + c[B3]);
}

// ---------------------------------------------------------------------------
// Compact forms for the mixed-precision kinematics.  They are algebraically
// the reference's formulas, regrouped so that the three element-centred
// derivative directions share one set of body-diagonal differences; the
// roundings differ from the reference's, so the bit-exact FP64 forms above
// do not use them.
template <typename T>
struct Vec3 {
   T x, y, z;
};

template <typename T>
LG_DEV Vec3<T> cross(const Vec3<T> &a, const Vec3<T> &b)
{
   return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <typename T>
LG_DEV T dot(const Vec3<T> &a, const Vec3<T> &b)
{
   return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Derivatives of one coordinate (or velocity component) along xi, eta and
// zeta at the element centre, unscaled: 8x the Jacobian entries of
// lulesh.cc:319-329 (fjxxi, fjxet, fjxze), which are also 4x the
// face-centred differences dxi, dxj, dxk of lulesh.cc:1691-1701.  All three
// come from the four body diagonals.
template <typename T>
LG_DEV Vec3<T> centre_derivs(const T (&c)[8])
{
   const T a = c[6] - c[0], b = c[5] - c[3], g = c[7] - c[1], h = c[4] - c[2];
   const T p = a + b, q = g + h;
   return {p - q, (a - b) + (g - h), p + q};
}

// Cross product of the diagonals of face (A, B, C, D): (C - A) x (D - B).
template <typename T, int A, int B, int C, int D>
LG_DEV Vec3<T> face_diag_cross(const T (&x)[8], const T (&y)[8], const T (&z)[8])
{
   const Vec3<T> d = {x[C] - x[A], y[C] - y[A], z[C] - z[A]};
   const Vec3<T> e = {x[D] - x[B], y[D] - y[B], z[D] - z[B]};
   return cross(d, e);
}

// |(C - A) x (D - B)|^2: AreaFace (lulesh.cc:1370-1390) is
// |f x g|^2 with f = (C-A) - (D-B), g = (C-A) + (D-B), which is 4x this.
template <typename T, int A, int B, int C, int D>
LG_DEV T face_diag_cross2(const T (&x)[8], const T (&y)[8], const T (&z)[8])
{
   const Vec3<T> n = face_diag_cross<T, A, B, C, D>(x, y, z);
   return dot(n, n);
}

// -8x CalcElemNodeNormals (lulesh.cc:419-474).  SumElemFaceNormal's area
// vector 0.25 (b0 x b1), with b0 = (p + q)/2 and b1 = (p - q)/2 for the face
// diagonals p = X2 - X0, q = X3 - X1, is -(p x q)/8; each node sums the
// three faces it lies on.
template <typename T>
LG_DEV void node_normals_diag(const T (&x)[8], const T (&y)[8], const T (&z)[8], T (&nx)[8],
                              T (&ny)[8], T (&nz)[8])
{
   const Vec3<T> f0 = face_diag_cross<T, 0, 1, 2, 3>(x, y, z);
   const Vec3<T> f1 = face_diag_cross<T, 0, 4, 5, 1>(x, y, z);
   const Vec3<T> f2 = face_diag_cross<T, 1, 5, 6, 2>(x, y, z);
   const Vec3<T> f3 = face_diag_cross<T, 2, 6, 7, 3>(x, y, z);
   const Vec3<T> f4 = face_diag_cross<T, 3, 7, 4, 0>(x, y, z);
   const Vec3<T> f5 = face_diag_cross<T, 4, 7, 6, 5>(x, y, z);
   auto put = [&](int j, const Vec3<T> &a, const Vec3<T> &b, const Vec3<T> &c) {
      nx[j] = a.x + b.x + c.x;
      ny[j] = a.y + b.y + c.y;
      nz[j] = a.z + b.z + c.z;
   };
   put(0, f0, f1, f4);
   put(1, f0, f1, f2);
   put(2, f0, f2, f3);
   put(3, f0, f3, f4);
   put(4, f1, f4, f5);
   put(5, f1, f2, f5);
   put(6, f2, f3, f5);
   put(7, f3, f4, f5);
}

// Sign of the centre Jacobian determinant of jacobian_volume: J_eta .
// (J_zeta x J_xi) of the unscaled centre derivatives is 64x it.
template <typename T>
LG_DEV T jacobian_volume_compact(const T (&x)[8], const T (&y)[8], const T (&z)[8])
{
   const Vec3<T> cx = centre_derivs<T>(x), cy = centre_derivs<T>(y), cz = centre_derivs<T>(z);
   const Vec3<T> Ji = {cx.x, cy.x, cz.x}, Jj = {cx.y, cy.y, cz.y}, Jk = {cx.z, cy.z, cz.z};
   return dot(Jj, cross(Jk, Ji));
}

// CalcElemCharacteristicLength (lulesh.cc:1394-1435) from the diagonal
// cross products: 4 V / sqrt(4 m) = 2 V / sqrt(m).
template <typename T>
LG_DEV T char_length_diag(const T (&x)[8], const T (&y)[8], const T (&z)[8], T volume)
{
   T m = T(0);
   m = keep_max(face_diag_cross2<T, 0, 1, 2, 3>(x, y, z), m);
   m = keep_max(face_diag_cross2<T, 4, 5, 6, 7>(x, y, z), m);
   m = keep_max(face_diag_cross2<T, 0, 1, 5, 4>(x, y, z), m);
   m = keep_max(face_diag_cross2<T, 1, 2, 6, 5>(x, y, z), m);
   m = keep_max(face_diag_cross2<T, 2, 3, 7, 6>(x, y, z), m);
   m = keep_max(face_diag_cross2<T, 3, 0, 4, 7>(x, y, z), m);
   return div_root(T(2.0) * volume, m);
}

// Hourglass base vectors, lulesh.cc:743-776: gamma[m][j] = sign bit set
// where the table holds -1.
__host__ __device__ constexpr int hg_negative(int m, int j)
{
   // bit j of each mask is set where gamma[m][j] = -1
   return ((m == 0 ? 0x3C : m == 1 ? 0x96 : m == 2 ? 0xAA : 0xA5) >> j) & 1;
}

template <typename T>
LG_DEV T hg_gamma(int m, int j)
{
   return hg_negative(m, j) ? T(-1.) : T(1.);
}

} // namespace lgpu
