"""lgpu_time_constraints_noalloc (the allocation-free entry point) against
lgpu_time_constraints on the same state, through the C ABI.

Two calls in a row on one GPU share the library's fixed partials array; the
second is the smaller problem, so a stale partial of the first call would
show in its result.  The bar is bit-for-bit: a minimum does not depend on the
order in which the elements are visited.
"""
import ctypes

import numpy as np
import pytest

import lulesh_amd as la
from lulesh_amd import native

from hip_buffers import HipBuffers

pytestmark = pytest.mark.gpu


def _state_after(s: int, cycles: int):
    """A host Domain on the device after `cycles` cycles: its last
    (dtcourant, dthydro) from lgpu_time_constraints, the element fields the
    constraints read, and the concatenated region lists."""
    h = la.HostDomain(s, 11, 1, 1)
    h.attach_device(la.FP64)
    for _ in range(cycles):
        assert h.step() == 0
    lib = native.gpu_lib()
    fields = {}
    for f in ("ss", "vdov", "arealg"):
        out = np.empty(h.numElem, np.float64)
        assert lib.lgpu_download(h.device(), la.FIELDS.index(f),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        fields[f] = out
    lists = np.concatenate([h.ints("regElemlist", r) for r in range(h.numReg)]).astype(np.int32)
    want = (h.get("dtcourant"), h.get("dthydro"))
    h.close()
    return want, fields, lists


def _noalloc(bufs: HipBuffers, fields, lists, use_list: bool):
    ss, vdov, ar = (bufs.put(fields[f]) for f in ("ss", "vdov", "arealg"))
    lst = bufs.put(lists) if use_list else None
    out = bufs.put(np.full(2, -1.0))
    k = la.default_constants()
    n = len(lists) if use_list else len(fields["ss"])
    bufs.sync()
    st = native.gpu_lib().lgpu_time_constraints_noalloc(ss, vdov, ar, lst, n, k.qqc, k.dvovmax,
                                                        out, None)
    assert st == 0
    bufs.sync()
    got = tuple(bufs.get(out, 2, np.float64).tolist())
    bufs.free()
    return got


def test_noalloc_time_constraints_equal_context_entry_point():
    big = _state_after(24, 12)
    small = _state_after(6, 5)
    assert len(big[2]) == 24 ** 3 and len(small[2]) == 6 ** 3
    bufs = HipBuffers()
    # the larger problem first, over all elements 0..n-1
    got = _noalloc(bufs, big[1], big[2], use_list=False)
    print("noalloc s=24:", got, "context:", big[0], flush=True)
    assert got == big[0]
    # then the smaller one through its region lists: fewer workgroups, the
    # rest of the partials array still holds the first call's values
    got = _noalloc(bufs, small[1], small[2], use_list=True)
    print("noalloc s=6:", got, "context:", small[0], flush=True)
    assert got == small[0]
    # and the larger again, through its list, after the smaller
    got = _noalloc(bufs, big[1], big[2], use_list=True)
    assert got == big[0]

