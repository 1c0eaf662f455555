"""The C-ABI libraries load (no GPU needed) and export every entry point
their headers declare."""
import ctypes
import re
from pathlib import Path

import pytest

import lulesh_amd as la
from lulesh_amd import native

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(l(?:gpu|h)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header,loader", [("lulesh_gpu.h", native.gpu_lib),
                                           ("lulesh_host.h", native.host_lib)])
def test_exports(header, loader):
    lib = loader()
    names = declared(header)
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n


def test_field_numbering_matches_oracle():
    from oracle_binding import REAL_FIELDS
    assert la.FIELDS == REAL_FIELDS


def test_default_constants_are_the_references():
    k = la.default_constants()
    # lulesh-init.cc:20-38
    assert (k.e_cut, k.p_cut, k.q_cut, k.v_cut, k.u_cut) == (1e-7, 1e-7, 1e-7, 1e-10, 1e-7)
    assert (k.hgcoef, k.qstop, k.monoq_max_slope, k.monoq_limiter_mult) == (3.0, 1e12, 1.0, 2.0)
    assert (k.qlc_monoq, k.qqc_monoq, k.qqc) == (0.5, 2.0 / 3.0, 2.0)
    assert (k.eosvmax, k.eosvmin, k.pmin, k.emin, k.dvovmax, k.refdens) == (
        1e9, 1e-9, 0.0, -1e15, 0.1, 1.0)
    assert k.ss4o3 == 4.0 / 3.0


def test_create_rejects_bad_mesh_without_touching_the_device():
    """Validation happens on the host before any allocation: an
    out-of-range node index is refused with EINVAL."""
    from oracle_binding import OracleDomain
    from helpers import context_from_oracle
    o = OracleDomain(3)
    o.ints("nodelist")[5] = 10_000
    with pytest.raises(RuntimeError, match="EINVAL"):
        context_from_oracle(o)


def test_ensemble_create_rejects_oversized_member_without_touching_the_device():
    """A member beyond lgpu_create's size limit (8 corners per element must
    index in int32) is refused with EINVAL before any device call."""
    lib = la.gpu_lib()
    d = native.MeshDesc()
    d.precision = la.FP64
    d.numElem = 2 ** 28 + 1
    d.numNode = 1
    out = ctypes.c_void_p()
    rc = lib.lgpu_ensemble_create(ctypes.byref(d), 1, ctypes.byref(out))
    assert native.STATUS[rc] == "EINVAL"
    assert not out.value


def test_noalloc_time_constraints_rejects_bad_arguments_without_touching_the_device():
    """lgpu_time_constraints_noalloc checks its pointers and size on the host
    before any HIP call: a null array or a negative count is EINVAL."""
    lib = native.gpu_lib()
    p = ctypes.c_void_p(256)  # never dereferenced: the checks come first
    assert lib.lgpu_time_constraints_noalloc(None, p, p, None, 10, 2.0, 0.1, p, None) == -100
    assert lib.lgpu_time_constraints_noalloc(p, p, p, None, 10, 2.0, 0.1, None, None) == -100
    assert lib.lgpu_time_constraints_noalloc(p, p, p, None, -1, 2.0, 0.1, p, None) == -100


def test_fp64_routines_reject_bad_arguments_without_touching_the_device():
    """lgpu_fp64_routines checks its input pointer and count on the host, and
    n == 0 enqueues nothing."""
    lib = native.gpu_lib()
    p = ctypes.c_void_p(16)
    assert lib.lgpu_fp64_routines(None, 10, p, p, None) == -100
    assert lib.lgpu_fp64_routines(p, -1, p, p, None) == -100
    assert lib.lgpu_fp64_routines(p, 0, p, p, None) == 0
