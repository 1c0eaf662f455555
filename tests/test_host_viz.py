"""The end-of-run visualization dump (DumpToVisit, lulesh-viz.cc) written as
legacy binary VTK (host/viz.cc, lh_write_vtk): parsed back here and checked
against the host Domain's arrays section by section (no GPU needed: the dump
writes the host copies)."""
import numpy as np

import lulesh_amd as la


def read_vtk(path):
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    def array(dtype, n):
        nonlocal pos
        a = np.frombuffer(data, dtype=np.dtype(dtype).newbyteorder(">"), count=n, offset=pos)
        pos += a.nbytes
        if pos < len(data) and data[pos:pos + 1] == b"\n":
            pos += 1
        return a.astype(np.dtype(dtype).newbyteorder("="))

    out = {"header": [line(), line(), line(), line()]}
    kw, n, _ = line().split()
    assert kw == "POINTS"
    out["points"] = array("f4", 3 * int(n)).reshape(-1, 3)
    kw, ne, size = line().split()
    assert kw == "CELLS"
    out["cells"] = array("i4", int(size)).reshape(-1, 9)
    kw, ne2 = line().split()
    assert kw == "CELL_TYPES"
    out["types"] = array("i4", int(ne2))
    section = None
    while pos < len(data):
        words = line().split()
        if not words:
            continue
        if words[0] in ("CELL_DATA", "POINT_DATA"):
            section = (words[0], int(words[1]))
            continue
        assert words[0] == "SCALARS", words
        name, typ = words[1], words[2]
        assert line() == "LOOKUP_TABLE default"
        out[name] = array({"double": "f8", "int": "i4"}[typ], section[1])
    return out


def test_vtk_dump_matches_domain(tmp_path):
    s = 6
    h = la.HostDomain(s, 11, 1, 1)
    # give the node and element fields distinct values
    rng = np.random.default_rng(3)
    for f in ("xd", "yd", "zd", "e", "p", "q", "v"):
        h.real(f)[:] = rng.normal(size=h.real(f).shape)
    path = tmp_path / "plot.vtk"
    h.write_vtk(path)
    d = read_vtk(path)
    assert d["header"][0] == "# vtk DataFile Version 3.0"
    assert d["header"][2] == "BINARY"
    assert d["header"][3] == "DATASET UNSTRUCTURED_GRID"
    xyz = np.stack([h.real("x"), h.real("y"), h.real("z")], axis=1).astype(np.float32)
    assert np.array_equal(d["points"], xyz)
    assert np.all(d["cells"][:, 0] == 8)
    assert np.array_equal(d["cells"][:, 1:].ravel(), h.ints("nodelist"))
    assert np.all(d["types"] == 12)
    assert np.array_equal(d["region"], h.ints("regNumList"))
    for f in ("e", "p", "v", "q", "xd", "yd", "zd"):
        assert np.array_equal(d[f], h.real(f)), f
    speed = np.sqrt(h.real("xd") ** 2 + h.real("yd") ** 2 + h.real("zd") ** 2)
    assert np.allclose(d["speed"], speed, rtol=1e-15, atol=0)


def test_vtk_hexahedra_have_positive_volume(tmp_path):
    # VTK's hexahedron ordering (bottom face counter-clockwise seen from the
    # top, then the top face) must give the undeformed cells a positive
    # volume: the triple product of the edges from corner 0.
    h = la.HostDomain(3, 1, 1, 1)
    path = tmp_path / "plot.vtk"
    h.write_vtk(path)
    d = read_vtk(path)
    p = d["points"].astype(np.float64)
    c = d["cells"][:, 1:]
    e1 = p[c[:, 1]] - p[c[:, 0]]
    e3 = p[c[:, 3]] - p[c[:, 0]]
    e4 = p[c[:, 4]] - p[c[:, 0]]
    assert np.all(np.einsum("ij,ij->i", np.cross(e1, e3), e4) > 0)


def test_vtk_unwritable_path_raises(tmp_path):
    h = la.HostDomain(2, 1, 1, 1)
    try:
        h.write_vtk(tmp_path / "missing" / "plot.vtk")
    except OSError:
        return
    raise AssertionError("expected OSError")
