"""Closed-form and invariant checks of the oracle's element kernels and
set-up (CPU only).  Each check is chosen so a plausible slip -- a wrong node
in a permutation table, a sign, a dropped term -- fails it:

* hexahedron volume: exact for boxes and parallelepipeds (det of the map);
* volume derivative: V is linear in every single coordinate, so a central
  difference reproduces dV/dx_j exactly; sum_j dV/dx_j = 0 and
  sum_j x_j dV/dx_j = V, sum_j y_j dV/dx_j = 0 (shear);
* node normals: outward, sum to zero, sum_j x_j n_j,x = V on parallelepipeds;
* characteristic length of a cube of edge h is h (face form 16 h^4);
* Flanagan-Belytschko hourglass forces vanish for any affine velocity field;
* uniform pressure gives zero net force on interior nodes;
* initial state closed forms; x<->y<->z symmetry of the Sedov run.
"""
import numpy as np
import pytest

from oracle_binding import (OracleDomain, elem_char_length, elem_node_normals, elem_volume,
                            elem_volume_derivative)

# Reference unit hexahedron in the reference's node order (lulesh-init.cc:252-259).
UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)
rng = np.random.default_rng(7)


def mapped(A, t=(0.0, 0.0, 0.0)):
    P = UNIT @ np.asarray(A).T + np.asarray(t)
    return P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy()


def random_affine():
    while True:
        A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        if np.linalg.det(A) > 0.3:
            return A


def distorted():
    A = random_affine()
    x, y, z = mapped(A, rng.standard_normal(3))
    return x + 0.05 * rng.standard_normal(8), y + 0.05 * rng.standard_normal(8), \
        z + 0.05 * rng.standard_normal(8)


@pytest.mark.parametrize("dims", [(1.0, 1.0, 1.0), (0.5, 2.0, 3.0), (1.125 / 45, 1.125 / 45, 1.125 / 45)])
def test_box_volume(dims):
    x, y, z = mapped(np.diag(dims), (3.0, -1.0, 0.25))
    assert elem_volume(x, y, z) == pytest.approx(np.prod(dims), rel=1e-14)


def test_parallelepiped_volume_is_det():
    for _ in range(20):
        A = random_affine()
        x, y, z = mapped(A, rng.standard_normal(3))
        assert elem_volume(x, y, z) == pytest.approx(np.linalg.det(A), rel=1e-13)


def test_volume_derivative_is_exact_gradient():
    for _ in range(10):
        x, y, z = distorted()
        dx, dy, dz = elem_volume_derivative(x, y, z)
        h = 1e-3
        for j in range(8):
            for arr, d in ((x, dx), (y, dy), (z, dz)):
                save = arr[j]
                arr[j] = save + h
                vp = elem_volume(x, y, z)
                arr[j] = save - h
                vm = elem_volume(x, y, z)
                arr[j] = save
                assert d[j] == pytest.approx((vp - vm) / (2 * h), rel=1e-9, abs=1e-12)


def test_volume_derivative_identities():
    for _ in range(10):
        x, y, z = distorted()
        V = elem_volume(x, y, z)
        dx, dy, dz = elem_volume_derivative(x, y, z)
        for d in (dx, dy, dz):
            assert abs(d.sum()) < 1e-13
        assert np.dot(x, dx) == pytest.approx(V, rel=1e-12)
        assert np.dot(y, dy) == pytest.approx(V, rel=1e-12)
        assert np.dot(z, dz) == pytest.approx(V, rel=1e-12)
        assert abs(np.dot(y, dx)) < 1e-12 and abs(np.dot(z, dx)) < 1e-12
        assert abs(np.dot(x, dy)) < 1e-12 and abs(np.dot(x, dz)) < 1e-12


def test_node_normals_parallelepiped():
    for _ in range(10):
        A = random_affine()
        x, y, z = mapped(A, rng.standard_normal(3))
        V = elem_volume(x, y, z)
        nx, ny, nz = elem_node_normals(x, y, z)
        for n in (nx, ny, nz):
            assert abs(n.sum()) < 1e-13
        assert np.dot(x, nx) == pytest.approx(V, rel=1e-12)
        assert np.dot(y, ny) == pytest.approx(V, rel=1e-12)
        assert np.dot(z, nz) == pytest.approx(V, rel=1e-12)


def test_node_normals_unit_cube_point_outward():
    x, y, z = mapped(np.eye(3))
    nx, ny, nz = elem_node_normals(x, y, z)
    for j in range(8):
        # each node sees three quarter-faces of area 1/4: outward components
        assert nx[j] == (0.25 if UNIT[j, 0] else -0.25)
        assert ny[j] == (0.25 if UNIT[j, 1] else -0.25)
        assert nz[j] == (0.25 if UNIT[j, 2] else -0.25)


@pytest.mark.parametrize("h", [1.0, 0.1, 1.125 / 192])
def test_char_length_of_cube(h):
    x, y, z = mapped(h * np.eye(3))
    assert elem_char_length(x, y, z, h ** 3) == pytest.approx(h, rel=1e-14)


def _distort_mesh(o: OracleDomain, amp: float) -> None:
    """Moves interior nodes a little (keeps every element valid)."""
    s = o.s
    h = 1.125 / s
    idx = np.arange(o.numNode)
    i, j, k = idx % (s + 1), (idx // (s + 1)) % (s + 1), idx // (s + 1) ** 2
    interior = (i > 0) & (i < s) & (j > 0) & (j < s) & (k > 0) & (k < s)
    r = np.random.default_rng(3)
    for f in ("x", "y", "z"):
        o.real(f)[interior] += amp * h * r.uniform(-1, 1, interior.sum())
    return interior


def _sync_relative_volume(o: OracleDomain) -> None:
    """v = V(current coordinates)/volo, as it is at the start of every cycle
    (the hourglass control divides by volo*v)."""
    nl = o.ints("nodelist").reshape(-1, 8)
    X, Y, Z = o.real("x"), o.real("y"), o.real("z")
    for e in range(o.numElem):
        n = nl[e]
        o.real("v")[e] = elem_volume(X[n], Y[n], Z[n]) / o.real("volo")[e]


def test_hourglass_forces_vanish_for_affine_velocity():
    o = OracleDomain(5)
    _distort_mesh(o, 0.15)
    _sync_relative_volume(o)
    o.real("p")[:] = 0.0
    o.real("q")[:] = 0.0
    o.real("ss")[:] = 1.0
    X = np.stack([o.real("x"), o.real("y"), o.real("z")])
    A = rng.standard_normal((3, 3))
    b = rng.standard_normal(3)
    V = A @ X + b[:, None]
    for c, f in enumerate(("xd", "yd", "zd")):
        o.real(f)[:] = V[c]
    assert o.call("calc_force_for_nodes") == 0
    affine = np.abs(np.stack([o.real("fx"), o.real("fy"), o.real("fz")])).max()
    for f in ("xd", "yd", "zd"):
        o.real(f)[:] = rng.standard_normal(o.numNode)
    assert o.call("calc_force_for_nodes") == 0
    generic = np.abs(np.stack([o.real("fx"), o.real("fy"), o.real("fz")])).max()
    assert generic > 1e-6
    assert affine < 1e-10 * generic


def test_uniform_pressure_balances_interior_nodes():
    o = OracleDomain(5)
    interior = _distort_mesh(o, 0.2)
    o.real("p")[:] = 1.0
    o.real("q")[:] = 0.0
    o.real("ss")[:] = 0.0
    assert o.call("calc_force_for_nodes") == 0
    f = np.stack([o.real("fx"), o.real("fy"), o.real("fz")])
    scale = np.abs(f).max()
    assert scale > 1e-3
    assert np.abs(f[:, interior]).max() < 1e-13 * scale / (1.125 / 5) ** 2 * 1e-2 + 1e-15


@pytest.mark.parametrize("s", [3, 10, 30])
def test_initial_state(s):
    o = OracleDomain(s)
    h = 1.125 / s
    assert np.allclose(o.real("volo"), h ** 3, rtol=1e-13)
    assert o.real("nodalMass").sum() == pytest.approx(1.125 ** 3, rel=1e-12)
    e0 = 3.948746e7 * (s / 45.0) ** 3
    assert o.real("e")[0] == pytest.approx(e0, rel=1e-15)
    assert np.all(o.real("e")[1:] == 0.0)
    assert o.get("deltatime") == pytest.approx(0.5 * h / np.sqrt(2 * e0), rel=1e-14)
    assert np.all(o.real("v") == 1.0)


def test_region_repeat_rule():
    # lulesh.cc:2390-2400: rep 1 below numReg/2, 1+cost below
    # numReg - (numReg+15)/20, 10*(1+cost) above.
    o = OracleDomain(4, 11, 1, 1)
    assert o.region_reps() == [1] * 5 + [2] * 5 + [20]
    o = OracleDomain(4, 21, 2, 2)
    assert o.region_reps() == [1] * 10 + [3] * 10 + [30]


def test_sedov_permutation_symmetry():
    s = 6
    o = OracleDomain(s)
    assert o.run(15) == 0
    e = o.real("e").reshape(s, s, s)  # [k, j, i]
    scale = np.abs(e).max()
    for perm in ((0, 2, 1), (1, 0, 2), (2, 1, 0)):
        assert np.abs(e - e.transpose(perm)).max() < 1e-10 * scale
