"""The FP64 kernels compute several reference formulas without their
power-of-two factors (lgpu_geom.h: jac_row8, shape_derivatives64,
add_face_normal16, node_normals16; lgpu_step.inc: stress_corners' sig/16,
strain_trace, kin_reference's Q gradients with vol16, ptiny256 and norm/64).
Rounding commutes with such factors, so the results must equal the
reference's bit for bit.  Restated here in FP64 numpy, operation by operation
as the device evaluates them, and compared with np.array_equal (not a
tolerance) against the oracle's reference-order functions on distorted
elements; a dropped factor, a wrong sign or a changed association fails.
"""
import numpy as np

import oracle_binding as ob
from oracle_binding import OracleDomain

NORMAL_FACES = [(0, 1, 2, 3), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0), (4, 7, 6, 5)]
UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def node_normals16(x, y, z):
    """add_face_normal16 over the six faces, in the reference's face order."""
    nx, ny, nz = np.zeros(8), np.zeros(8), np.zeros(8)
    for n0, n1, n2, n3 in NORMAL_FACES:
        b0x = x[n3] + x[n2] - x[n1] - x[n0]
        b0y = y[n3] + y[n2] - y[n1] - y[n0]
        b0z = z[n3] + z[n2] - z[n1] - z[n0]
        b1x = x[n2] + x[n1] - x[n3] - x[n0]
        b1y = y[n2] + y[n1] - y[n3] - y[n0]
        b1z = z[n2] + z[n1] - z[n3] - z[n0]
        ax = b0y * b1z - b0z * b1y
        ay = b0z * b1x - b0x * b1z
        az = b0x * b1y - b0y * b1x
        for j in (n0, n1, n2, n3):
            nx[j] += ax
            ny[j] += ay
            nz[j] += az
    return nx, ny, nz


def jac_row8(c):
    a, b, g, h = c[6] - c[0], c[5] - c[3], c[7] - c[1], c[4] - c[2]
    return a + b - g - h, a - b + g - h, a + b + g + h


def shape_derivatives64(x, y, z):
    (jxx, jxe, jxz), (jyx, jye, jyz), (jzx, jze, jzz) = jac_row8(x), jac_row8(y), jac_row8(z)
    c = [[(jye * jzz) - (jze * jyz), (jzx * jyz) - (jyx * jzz), (jyx * jze) - (jzx * jye)],
         [(jze * jxz) - (jxe * jzz), (jxx * jzz) - (jzx * jxz), (jzx * jxe) - (jxx * jze)],
         [(jxe * jyz) - (jye * jxz), (jyx * jxz) - (jxx * jyz), (jxx * jye) - (jyx * jxe)]]
    b = [[-c[r][0] - c[r][1] - c[r][2], c[r][0] - c[r][1] - c[r][2],
          c[r][0] + c[r][1] - c[r][2], -c[r][0] + c[r][1] - c[r][2]] for r in range(3)]
    return b, jxe * c[0][1] + jye * c[1][1] + jze * c[2][1]


def strain_trace(x, y, z, u, v, w):
    b, det = shape_derivatives64(x, y, z)
    inv = 1.0 / det
    dxx = inv * (b[0][0] * (u[0] - u[6]) + b[0][1] * (u[1] - u[7]) +
                 b[0][2] * (u[2] - u[4]) + b[0][3] * (u[3] - u[5]))
    dyy = inv * (b[1][0] * (v[0] - v[6]) + b[1][1] * (v[1] - v[7]) +
                 b[1][2] * (v[2] - v[4]) + b[1][3] * (v[3] - v[5]))
    dzz = inv * (b[2][0] * (w[0] - w[6]) + b[2][1] * (w[1] - w[7]) +
                 b[2][2] * (w[2] - w[4]) + b[2][3] * (w[3] - w[5]))
    return dxx + dyy + dzz


def face_diff(c, a0, a1, a2, a3, b0, b1, b2, b3):
    return ((c[a0] + c[a1]) + c[a2]) + c[a3] - (((c[b0] + c[b1]) + c[b2]) + c[b3])


FJ, FI, FK = (0, 1, 5, 4, 3, 2, 6, 7), (1, 2, 6, 5, 0, 3, 7, 4), (4, 5, 6, 7, 0, 1, 2, 3)


def q_gradients(x, y, z, u, v, w, volo, vnew):
    """kin_reference's gradients: (delx_xi, delx_eta, delx_zeta, delv_xi,
    delv_eta, delv_zeta)."""
    ptiny = 1.e-36
    vol = volo * vnew
    norm64 = (1.0 / (vol + ptiny)) * (1.0 / 64.0)
    vol16, ptiny256 = 16.0 * vol, 256.0 * ptiny
    dxj, dyj, dzj = (-face_diff(c, *FJ) for c in (x, y, z))
    dxi, dyi, dzi = (face_diff(c, *FI) for c in (x, y, z))
    dxk, dyk, dzk = (face_diff(c, *FK) for c in (x, y, z))

    def pair(ax, ay, az, fu, fv, fw):
        delx = vol16 / np.sqrt(ax * ax + ay * ay + az * az + ptiny256)
        ax, ay, az = ax * norm64, ay * norm64, az * norm64
        return delx, ax * fu + ay * fv + az * fw

    xz, vz = pair(dyi * dzj - dzi * dyj, dzi * dxj - dxi * dzj, dxi * dyj - dyi * dxj,
                  face_diff(u, *FK), face_diff(v, *FK), face_diff(w, *FK))
    xx, vx = pair(dyj * dzk - dzj * dyk, dzj * dxk - dxj * dzk, dxj * dyk - dyj * dxk,
                  face_diff(u, *FI), face_diff(v, *FI), face_diff(w, *FI))
    xe, ve = pair(dyk * dzi - dzk * dyi, dzk * dxi - dxk * dzi, dxk * dyi - dyk * dxi,
                  -face_diff(u, *FJ), -face_diff(v, *FJ), -face_diff(w, *FJ))
    return xx, xe, xz, vx, ve, vz


def _distorted(rng, shift):
    scale = rng.uniform(0.5, 2.0, 3)
    return UNIT * scale + rng.normal(0.0, 0.15, (8, 3)) + rng.normal(0.0, shift, 3)


def test_node_normals_and_stress_corners_bit_identical():
    rng = np.random.default_rng(20261203)
    for k in range(2000):
        X = _distorted(rng, (0.0, 1.0, 100.0)[k % 3]) * 10.0 ** rng.uniform(-3, 3)
        x, y, z = (np.ascontiguousarray(X[:, i]) for i in range(3))
        ref = ob.elem_node_normals(x, y, z)
        n16 = node_normals16(x, y, z)
        sig = -rng.uniform(0.0, 1e3)
        for r, n in zip(ref, n16):
            assert np.array_equal(n * 0.0625, r)
            # stress_corners: -(sig/16 * n16) against the reference's -(sig * n)
            assert np.array_equal(-((sig * 0.0625) * n), -(sig * r))


def test_kinematics_strain_and_gradients_bit_identical():
    rng = np.random.default_rng(13)
    o = OracleDomain(5)
    h = 1.125 / 5
    for f in ("x", "y", "z"):
        o.real(f)[:] += rng.normal(0.0, 0.08 * h, o.numNode)
    for f in ("xd", "yd", "zd"):
        o.real(f)[:] = rng.normal(0.0, 1.0, o.numNode)
    dt = 1.0e-3
    o.set("deltatime", dt)
    assert o.call("calc_lagrange_elements") == 0
    o.call("calc_monotonic_q_gradients")
    nl = o.ints("nodelist").reshape(-1, 8)
    pos = np.stack([o.real(f) for f in ("x", "y", "z")], axis=1)
    vel = np.stack([o.real(f) for f in ("xd", "yd", "zd")], axis=1)
    volo, vnew = o.real("volo"), o.real("vnew")
    want = {f: o.real(f) for f in ("vdov", "delx_xi", "delx_eta", "delx_zeta",
                                   "delv_xi", "delv_eta", "delv_zeta")}
    for e in range(o.numElem):
        X, V = pos[nl[e]], vel[nl[e]]
        x, y, z = X[:, 0], X[:, 1], X[:, 2]
        u, v, w = V[:, 0], V[:, 1], V[:, 2]
        grads = q_gradients(x, y, z, u, v, w, volo[e], vnew[e])
        for f, g in zip(("delx_xi", "delx_eta", "delx_zeta", "delv_xi", "delv_eta", "delv_zeta"),
                        grads):
            assert g == want[f][e], (f, e)
        # strain rate on the half-step coordinates (lulesh.cc:1549-1566)
        dt2 = 0.5 * dt
        hx, hy, hz = x - dt2 * u, y - dt2 * v, z - dt2 * w
        assert strain_trace(hx, hy, hz, u, v, w) == want["vdov"][e], e
