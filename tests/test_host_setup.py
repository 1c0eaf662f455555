"""The product's C++ host Domain builds exactly the reference's problem
(CPU only; no device calls)."""
import ctypes

import numpy as np
import pytest

import lulesh_amd as la
from lulesh_amd import native
from oracle_binding import OracleDomain


@pytest.mark.parametrize("s,r,b,c,m", [(1, 1, 1, 1, 1.0), (4, 11, 1, 1, 1.0), (6, 21, 2, 2, 1.0),
                                       (9, 1, 1, 1, 1.0), (12, 7, 3, 0, 0.7), (17, 30, 0, 5, 1.9)])
def test_host_domain_equals_oracle(s, r, b, c, m):
    h = la.HostDomain(s, r, b, c, m)
    o = OracleDomain(s, r, b, c, m)
    for name in native.INT_ARRAYS[:-1]:
        assert np.array_equal(h.ints(name), o.ints(name)), name
    for rr in range(h.numReg):
        assert np.array_equal(h.ints("regElemlist", rr), o.ints("regElemlist", rr)), rr
    for f in la.FIELDS[:26]:
        assert np.array_equal(h.real(f), o.real(f)), f
    for sc in native.SCALARS:
        assert h.get(sc) == o.get(sc), sc


@pytest.mark.parametrize("seed", [0, 1, 2, 12345, 2**31 - 1])
def test_glibc_sequence(seed):
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    n = 20000
    out = np.zeros(n, np.int32)
    native.host_lib().lh_glibc_rand(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    libc.srand(seed)
    want = np.array([libc.rand() for _ in range(n)], np.int32)
    assert np.array_equal(out, want)
