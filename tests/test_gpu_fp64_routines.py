"""The FP64 reciprocal and square root that k_eos uses inside its EOS repeats
(recip_rn_normal, sqrt_rn_normal), evaluated through lgpu_fp64_routines and
compared bit for bit with IEEE-754 division and square root (numpy's 1.0 / x
and np.sqrt on the host, which round correctly).

The EOS phase and whole-run parity tests check these routines on the values a
run produces; this test covers the whole range each is used on, with the
cases where a Newton-Raphson reciprocal or root is hardest to round: powers of
two, significands of all ones, neighbours of 1 and of the range ends.
"""
import ctypes

import numpy as np
import pytest

from lulesh_amd import native

from hip_buffers import HipBuffers

pytestmark = pytest.mark.gpu

RECIP_LO, RECIP_HI = 2.0 ** -400, 2.0 ** 400
SQRT_LO = 2.0 ** -767


def _hard_cases(lo_exp: int, hi_exp: int) -> np.ndarray:
    """Powers of two, all-ones significands and their neighbours for every
    exponent in [lo_exp, hi_exp), and values within a few ulps of 1."""
    e = np.arange(lo_exp, hi_exp, dtype=np.float64)
    p2 = np.exp2(e)
    ones = p2 * (2.0 - 2.0 ** -52)
    near = [p2, ones, np.nextafter(p2, np.inf), np.nextafter(ones, 0.0),
            p2 * 1.5, p2 * (1.0 + 2.0 ** -26)]
    one = 1.0 + np.arange(-64, 65) * 2.0 ** -52
    return np.concatenate(near + [one])


def _run(x: np.ndarray):
    lib = native.gpu_lib()
    bufs = HipBuffers()
    try:
        dx = bufs.put(x)
        dr = bufs.put(np.zeros_like(x))
        ds = bufs.put(np.zeros_like(x))
        assert lib.lgpu_fp64_routines(dx, x.size, dr, ds, None) == 0
        bufs.sync()
        return bufs.get(dr, x.size, np.float64), bufs.get(ds, x.size, np.float64)
    finally:
        bufs.free()


def _mismatches(got: np.ndarray, want: np.ndarray) -> int:
    return int(np.count_nonzero(got.view(np.uint64) != want.view(np.uint64)))


def test_reciprocal_rounds_correctly_over_its_range():
    rng = np.random.default_rng(20261203)
    n = 1 << 21
    mag = np.exp2(rng.uniform(-400.0, 400.0, n))
    hard = _hard_cases(-400, 400)
    x = np.concatenate([mag * np.where(rng.random(n) < 0.5, -1.0, 1.0),
                        rng.uniform(0.5, 2.0, n),           # the EOS's own divisors
                        hard, -hard, [RECIP_LO, RECIP_HI, -RECIP_LO, -RECIP_HI]])
    x = x[(np.abs(x) >= RECIP_LO) & (np.abs(x) <= RECIP_HI)]
    rc, _ = _run(x)
    assert _mismatches(rc, 1.0 / x) == 0


def test_square_root_rounds_correctly_over_its_range():
    rng = np.random.default_rng(20261204)
    n = 1 << 21
    x = np.concatenate([np.exp2(rng.uniform(-767.0, 1023.0, n)),
                        rng.uniform(1e-3, 1e3, n),          # sound-speed squares
                        _hard_cases(-767, 1023),
                        [np.finfo(np.float64).max, SQRT_LO]])
    x = x[(x >= SQRT_LO) & np.isfinite(x)]
    _, rt = _run(x)
    assert _mismatches(rt, np.sqrt(x)) == 0
