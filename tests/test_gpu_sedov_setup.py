"""Device-side set-up of the reference's Sedov problem (lgpu_create_sedov).

The set-up kernels build the mesh, node-to-corner lists, face neighbours,
boundary masks, region work list, reference volumes, masses and the energy
deposit.  Checked against the oracle's Domain (pinned to the reference's
lulesh-init.cc): the initial fields and time state are bit-identical, and a
run from the device-built context is bit-identical to one from a context
built from the host Domain, on the brick path and on the CSR-gather path
(which reads the device-built node lists, corner lists, neighbours and
boundary masks).  Whole runs reproduce the reference's own numbers.
"""
import ctypes

import numpy as np
import pytest

import lulesh_amd as la
from helpers import summary, symmetry
from oracle_binding import OracleDomain

pytestmark = pytest.mark.gpu

INIT_FIELDS = ["x", "y", "z", "xd", "yd", "zd", "nodalMass", "e", "p", "q", "v", "volo", "ss",
               "elemMass"]
RUN_FIELDS = ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss", "vnew", "delv",
              "arealg", "vdov", "delx_xi", "delx_eta", "delx_zeta", "delv_xi", "delv_eta",
              "delv_zeta"]
TIME = ["time", "deltatime", "dtcourant", "dthydro", "stoptime", "dtfixed", "dtmax",
        "deltatimemultlb", "deltatimemultub"]
PRECISIONS = [pytest.param(la.FP64, id="fp64"), pytest.param(la.MIXED, id="mixed"),
              pytest.param(la.FP64_FMA, id="fp64fma")]


@pytest.mark.parametrize("s,r,b,c,emult", [(1, 1, 1, 1, 1.0), (2, 11, 1, 1, 1.0),
                                           (5, 21, 2, 2, 1.0), (16, 11, 1, 2, 0.5),
                                           (33, 21, 2, 1, 1.7)])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_initial_state_matches_oracle(s, r, b, c, emult, precision):
    print(f"sedov set-up s={s} r={r} b={b} c={c} m={emult} precision={precision}", flush=True)
    o = OracleDomain(s, r, b, c, emult)
    d = la.SedovDevice(s, r, b, c, emult, precision=precision)
    for f in INIT_FIELDS:
        want = o.real(f)
        if precision == la.MIXED and f in ("p", "q", "ss"):
            want = want.astype(np.float32).astype(np.float64)
        got = d.real(f)
        assert np.array_equal(got, want), f
    for name in TIME:
        assert d.get(name) == o.get(name), name
    assert d.get("cycle") == 0
    info = d.info()
    assert info.path == 1 and info.brick_faces == 1
    d.close()


def _host_run(s, r, b, c, emult, precision, flags, cycles):
    h = la.HostDomain(s, r, b, c, emult)
    h.attach_device(precision, flags)
    assert h.run(cycles, device_loop=True) == 0
    return h


def _download(ctx, name: str, n: int) -> np.ndarray:
    """A field of a host Domain's device context (every field, including the
    ones the host Domain does not copy back)."""
    out = np.empty(n, np.float64)
    la.check(la.gpu_lib().lgpu_download(ctx, la.FIELDS.index(name),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


@pytest.mark.parametrize("generic", [pytest.param(0, id="brick"),
                                     pytest.param(la.native.FLAG_GENERIC, id="generic")])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_run_matches_host_built_context(precision, generic):
    s, r, b, c, emult, cycles = 24, 21, 2, 2, 1.25, 40
    print(f"sedov run s={s} precision={precision} flags={generic}", flush=True)
    d = la.SedovDevice(s, r, b, c, emult, precision=precision, flags=generic)
    assert d.info().path == (0 if generic else 1)
    assert d.run(cycles) == 0
    h = _host_run(s, r, b, c, emult, precision, generic, cycles)
    assert d.get("cycle") == h.get("cycle") == cycles
    for name in ("time", "deltatime", "dtcourant", "dthydro"):
        assert d.get(name) == h.get(name), name
    for f in RUN_FIELDS:
        n = d.numNode if f in la.native.NODE_FIELDS else d.numElem
        assert np.array_equal(d.real(f), _download(h.device(), f, n)), f
    d.close()


def test_full_run_s30_to_stoptime():
    d = la.SedovDevice(30, 11, 1, 1)
    ref = summary(30, 11, 1, 1, 9999999)
    assert d.run(9999999) == 0
    assert d.get("cycle") == ref["cycles"]
    e0 = d.real("e")[0]
    assert abs(e0 - float(ref["e0"])) <= 1e-10 * float(ref["e0"])


def test_s192_load_imbalance_200_cycles():
    """The headline workload (s=192, 21 regions, balance 2, cost 2, 200
    cycles) from the device-side set-up: the reference's iteration count,
    final origin energy within 1e-10 and plane-0 symmetry figures
    (lulesh-util.cc:197-212) within a factor 10 of the reference's."""
    key = (192, 21, 2, 2, 200, 1.0)
    ref = summary(*key)
    d = la.SedovDevice(192, 21, 2, 2)
    assert d.run(200) == 0
    assert d.get("cycle") == ref["cycles"] == 200
    e = d.real("e")
    e0 = e[0]
    sym = symmetry(e, 192)
    print(f"s=192 device set-up: e0 {e0:.17e} (reference {float(ref['e0']):.17e}) sym {sym}")
    assert abs(e0 - float(ref["e0"])) <= 1e-10 * float(ref["e0"])
    for got, name in zip(sym, ("maxabs", "totabs", "maxrel")):
        want = float(ref[name])
        assert want / 10 <= got <= want * 10, (name, got, want)
    d.close()


def _create(**over):
    lib = la.gpu_lib()
    reg, ln = la.region_runs(4, 3, 1)
    fields = dict(size=4, numReg=3, balance=1, cost=1, emult=1.0, precision=la.FP64, flags=0)
    fields.update({k: v for k, v in over.items() if k in fields})
    reg = over.get("reg", reg)
    ln = over.get("len", ln)
    d = la.native.SedovDesc()
    for k, v in fields.items():
        setattr(d, k, v)
    reg, ln = np.ascontiguousarray(reg, np.int32), np.ascontiguousarray(ln, np.int32)
    d.numRuns = len(reg)
    d.runReg = reg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.runLen = ln.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.k = la.default_constants()
    ts = la.TimeState()
    ptr = ctypes.c_void_p()
    st = lib.lgpu_create_sedov(ctypes.byref(d), ctypes.byref(ptr), ctypes.byref(ts))
    if ptr:
        lib.lgpu_destroy(ptr)
    return st


def test_invalid_descriptions_are_rejected():
    assert _create() == 0
    reg, ln = la.region_runs(4, 3, 1)
    short = ln.copy()
    short[-1] -= 1
    assert _create(len=short) == -100            # runs do not cover the elements
    bad = reg.copy()
    bad[0] = 4
    assert _create(reg=bad) == -100              # region number above numReg
    assert _create(size=0) == -100
    assert _create(size=646) == -100            # 8 s^3 corners beyond int32
    assert _create(size=2**31 - 1) == -100       # no overflow before the range test
    assert _create(precision=7) == -100
    assert _create(emult=0.0) == -100
    zero = ln.copy()
    zero[0] = 0
    assert _create(len=zero) == -100             # empty run
