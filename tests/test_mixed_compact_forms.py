"""The compact element forms of the mixed-precision kernels (lgpu_geom.h:
centre_derivs, face_diag_cross, node_normals_diag, char_length_diag,
jacobian_volume_compact; lgpu_step.inc: kin_local, hourglass_modes with the
1/12 folded into 1/V) regroup the reference's formulas.  Restated here in
FP64 numpy and checked against the oracle's reference-order functions on
distorted elements, so that a wrong pairing, sign or scale in the regrouping
fails.  The device evaluates the same forms in FP32; its end-to-end accuracy
is held by tests/test_gpu_mixed.py.
"""
import numpy as np
import pytest

import oracle_binding as ob
from oracle_binding import OracleDomain

# Faces as SumElemFaceNormal is called (lulesh.cc:419-474) and nodes on each.
NORMAL_FACES = [(0, 1, 2, 3), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0), (4, 7, 6, 5)]
AREA_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def centre_derivs(c):
    a, b, g, h = c[6] - c[0], c[5] - c[3], c[7] - c[1], c[4] - c[2]
    p, q = a + b, g + h
    return np.array([p - q, (a - b) + (g - h), p + q])


def diag_cross(X, A, B, C, D):
    return np.cross(X[C] - X[A], X[D] - X[B])


def node_normals_diag(X):
    """-8 x the node normals, from the six face diagonal cross products."""
    out = np.zeros((8, 3))
    for f in NORMAL_FACES:
        n = diag_cross(X, *f)
        for j in f:
            out[j] += n
    return out


def char_length_diag(X, volume):
    m = max(float(np.dot(n, n)) for n in (diag_cross(X, *f) for f in AREA_FACES))
    return 2.0 * volume / np.sqrt(m)


def jacobian_sign_compact(X):
    J = np.stack([centre_derivs(X[:, c]) for c in range(3)])  # rows x,y,z; columns xi,eta,zeta
    Ji, Jj, Jk = J[:, 0], J[:, 1], J[:, 2]
    return float(np.dot(Jj, np.cross(Jk, Ji)))


def volume_derivative_x12(X):
    """12 x dV/dx_j with the VoluDer neighbour permutations (lulesh.cc:623-663)."""
    perms = {0: (1, 2, 3, 4, 5, 7), 3: (0, 1, 2, 7, 4, 6), 2: (3, 0, 1, 6, 7, 5),
             1: (2, 3, 0, 5, 6, 4), 4: (7, 6, 5, 0, 3, 1), 5: (4, 7, 6, 1, 0, 2),
             6: (5, 4, 7, 2, 1, 3), 7: (6, 5, 4, 3, 2, 0)}

    def form(a, b, A):
        a0, a1, a2, a3, a4, a5 = A
        return ((a[a1] + a[a2]) * (b[a0] + b[a1]) - (a[a0] + a[a1]) * (b[a1] + b[a2]) +
                (a[a0] + a[a4]) * (b[a3] + b[a4]) - (a[a3] + a[a4]) * (b[a0] + b[a4]) -
                (a[a2] + a[a5]) * (b[a3] + b[a5]) + (a[a3] + a[a5]) * (b[a2] + b[a5]))

    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    d = np.zeros((8, 3))
    for n, A in perms.items():
        d[n] = (form(y, z, A), -form(x, z, A), -form(y, x, A))
    return d


def kin_local(X, V, volo, vnew, dt):
    """kin_local: Q gradients, strain-rate trace, characteristic length."""
    G = X - X[0]
    vol = volo * vnew
    out = {"arealg": char_length_diag(G, ob.elem_volume(*X.T))}
    J = np.stack([centre_derivs(G[:, c]) for c in range(3)])
    U = np.stack([centre_derivs(V[:, c]) for c in range(3)])
    Ji, Jj, Jk = J[:, 0], J[:, 1], J[:, 2]
    Ui, Uj, Uk = U[:, 0], U[:, 1], U[:, 2]
    norm64 = 1.0 / (vol + 1e-36) / 64.0
    for name, a, u in (("zeta", np.cross(Ji, Jj), Uk), ("xi", np.cross(Jj, Jk), Ui),
                       ("eta", np.cross(Jk, Ji), Uj)):
        out["delx_" + name] = 16.0 * vol / np.sqrt(np.dot(a, a) + 256e-36)
        out["delv_" + name] = np.dot(a, u) * norm64
    Hi, Hj, Hk = Ji - 0.5 * dt * Ui, Jj - 0.5 * dt * Uj, Jk - 0.5 * dt * Uk
    cI, cJ, cK = np.cross(Hj, Hk), np.cross(Hk, Hi), np.cross(Hi, Hj)
    out["vdov"] = (np.dot(cI, Ui) + np.dot(cJ, Uj) + np.dot(cK, Uk)) / np.dot(Hj, cJ)
    return out


def _distorted(rng, shift=100.0):
    scale = rng.uniform(0.5, 2.0, 3)
    return UNIT * scale + rng.normal(0.0, 0.12, (8, 3)) + rng.normal(0.0, shift, 3)


def test_element_forms_match_oracle():
    rng = np.random.default_rng(7)
    for _ in range(300):
        X = _distorted(rng)
        x, y, z = X.T
        ref_n = np.stack(ob.elem_node_normals(x, y, z), axis=1)
        np.testing.assert_allclose(-node_normals_diag(X) / 8.0, ref_n, rtol=1e-9,
                                   atol=1e-12 * np.abs(ref_n).max())
        ref_d = np.stack(ob.elem_volume_derivative(x, y, z), axis=1)
        np.testing.assert_allclose(volume_derivative_x12(X) / 12.0, ref_d, rtol=1e-9,
                                   atol=1e-12 * np.abs(ref_d).max())
        vol = ob.elem_volume(x, y, z)
        assert char_length_diag(X, vol) == pytest.approx(ob.elem_char_length(x, y, z, vol),
                                                         rel=1e-12)
        # the compact determinant has the sign of the volume of these elements
        assert jacobian_sign_compact(X) > 0


def test_inverted_element_sign():
    X = UNIT.copy()
    X[[4, 5, 6, 7], 2] = -1.0  # top face pushed through the bottom one
    assert jacobian_sign_compact(X) < 0


def test_kinematics_forms_match_oracle():
    rng = np.random.default_rng(11)
    o = OracleDomain(4)
    h = 1.125 / 4
    for f in ("x", "y", "z"):
        a = o.real(f)
        a += rng.normal(0.0, 0.05 * h, a.shape)
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
    names = ("arealg", "vdov", "delx_xi", "delx_eta", "delx_zeta",
             "delv_xi", "delv_eta", "delv_zeta")
    ref = {k: o.real(k) for k in names}
    for e in range(o.numElem):
        got = kin_local(pos[nl[e]], vel[nl[e]], volo[e], vnew[e], dt)
        for k in names:
            scale = np.abs(ref[k]).max()
            assert abs(got[k] - ref[k][e]) <= 1e-10 * scale, (e, k, got[k], ref[k][e])
