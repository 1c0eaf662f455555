"""Parity of the HIP path with the oracle, through the C ABI.

Phase tests upload the oracle's state, run one phase on the device and
compare with the oracle's result for the same input.  The FP64 path is built
without FMA contraction and follows the reference's operation order, so
everything except the hourglass coefficient's cube root (device cbrt vs
glibc cbrt, <= 1 ulp apart) is expected to match bit for bit; the tests
assert exactly that where it holds and a 1e-12 bound (relative to the field's
scale) where the cube root enters.

Whole-run tests compare with the reference's own end-of-run numbers
(tests/golden/ref_summaries.json): exact iteration count, final origin
energy within 1e-10 relative, symmetry diffs of the same order.
"""
import numpy as np
import pytest

import error_inputs
import lulesh_amd as la
from helpers import STATE_FIELDS, context_from_oracle, max_rel, summary, upload_state
from oracle_binding import OracleDomain, have_reference_build, read_dump, run_refdump

pytestmark = pytest.mark.gpu

NODE_OUT = ["x", "y", "z", "xd", "yd", "zd", "xdd", "ydd", "zdd", "fx", "fy", "fz"]
ELEM_OUT = ["vnew", "delv", "arealg", "vdov", "delx_xi", "delx_eta", "delx_zeta", "delv_xi",
            "delv_eta", "delv_zeta", "ql", "qq", "e", "p", "q", "ss", "v"]


def advance(o: OracleDomain, cycles: int) -> None:
    for _ in range(cycles):
        o.call("time_increment")
        assert o.call("lagrange_leapfrog") == 0


PATHS = [pytest.param(False, id="brick"), pytest.param(True, id="generic")]


@pytest.mark.parametrize("generic", PATHS)
@pytest.mark.parametrize("s,r,b,c,warm,zero_ss", [
    (5, 11, 1, 1, 0, False),   # first cycle: ss = 0, no hourglass force
    (6, 21, 2, 2, 4, True),    # later cycle with ss zeroed on both sides
    (8, 11, 1, 1, 12, False),  # later cycle with hourglass forces (cbrt)
    (20, 11, 1, 1, 6, True),   # several brick tiles per edge, several z-chunks
    (31, 11, 1, 1, 3, False),
])
def test_lagrange_nodal(s, r, b, c, warm, zero_ss, generic):
    o = OracleDomain(s, r, b, c)
    advance(o, warm)
    o.call("time_increment")
    if zero_ss:
        o.real("ss")[:] = 0.0
    ctx = context_from_oracle(o, generic=generic)
    assert ctx.info().path == (0 if generic else 1)
    # the reference's boundary set-up is recognised, so the brick EOS pass
    # computes face neighbours and masks instead of reading them
    assert ctx.info().brick_faces == (0 if generic else 1)
    ctx.set_option(1, 1)  # keep fx.., xdd..
    upload_state(ctx, o)
    dt = o.get("deltatime")
    assert ctx.lagrange_nodal(dt) == 0
    assert o.call("lagrange_nodal") == 0
    exact = (warm == 0) or zero_ss
    for f in NODE_OUT:
        got, want = ctx.download(f), o.real(f)
        if exact:
            assert np.array_equal(got, want), (f, max_rel(got, want))
        else:
            assert max_rel(got, want) <= 1e-12, (f, max_rel(got, want))


@pytest.mark.parametrize("generic", PATHS)
@pytest.mark.parametrize("s,r,b,c,warm", [(5, 11, 1, 1, 0), (6, 21, 2, 2, 3), (8, 11, 1, 1, 12),
                                          (10, 7, 2, 1, 30)])
def test_lagrange_elements_and_dt(s, r, b, c, warm, generic):
    o = OracleDomain(s, r, b, c)
    advance(o, warm)
    o.call("time_increment")
    assert o.call("lagrange_nodal") == 0
    ctx = context_from_oracle(o, generic=generic)
    ctx.set_option(1, 1)
    upload_state(ctx, o)
    dt = o.get("deltatime")
    assert ctx.lagrange_elements(dt) == 0
    st, dtc, dth = ctx.time_constraints()
    assert st == 0
    assert o.call("lagrange_elements") == 0
    o.call("calc_time_constraints")
    for f in ELEM_OUT:
        got, want = ctx.download(f), o.real(f)
        assert np.array_equal(got, want), (f, max_rel(got, want))
    assert dtc == o.get("dtcourant")
    assert dth == o.get("dthydro")


@pytest.mark.parametrize("generic", PATHS)
def test_eos_waves_outside_the_fast_form_match(generic):
    """k_eos repeats the FP64 EOS of a wave in a fast form (short correctly
    rounded reciprocals, lower clamps as maximum instructions) when the
    wave's inputs bound every intermediate value, and otherwise in the
    general form.  Energies and pressures of magnitude 2^101 (beyond the
    bound, so the wave is redone in the general form) and 2^99 (within it)
    must both give the oracle's bits."""
    o = OracleDomain(8, 21, 2, 2)
    advance(o, 6)
    o.call("time_increment")
    assert o.call("lagrange_nodal") == 0
    e, p = o.real("e"), o.real("p")
    e[::37] = 2.0 ** 101
    e[5::41] = 2.0 ** 99
    p[11::53] = -2.0 ** 101
    p[17::59] = 2.0 ** 99
    ctx = context_from_oracle(o, generic=generic)
    ctx.set_option(1, 1)
    upload_state(ctx, o)
    dt = o.get("deltatime")
    assert ctx.lagrange_elements(dt) == 0
    st, dtc, dth = ctx.time_constraints()
    assert st == 0
    assert o.call("lagrange_elements") == 0
    o.call("calc_time_constraints")
    for f in ELEM_OUT:
        got, want = ctx.download(f), o.real(f)
        assert np.array_equal(got, want), (f, max_rel(got, want))
    assert dtc == o.get("dtcourant")
    assert dth == o.get("dthydro")


def test_eos_fast_form_equals_general_form():
    """The fast FP64 EOS repeats run only with constants that pass
    fast_eos_consts; pmin = -0.0 fails it (its clamp gives the same values up
    to the sign of zero) and so runs every wave in the general form.  From one
    mid-run state, the elements phase with the default constants and with
    pmin = -0.0 gives equal fields (np.array_equal, -0 == 0) and equal dt."""
    o = OracleDomain(10, 21, 2, 2)
    advance(o, 30)
    o.call("time_increment")
    assert o.call("lagrange_nodal") == 0
    dt = o.get("deltatime")
    runs = []
    for pmin in (0.0, -0.0):
        k = la.native.default_constants()
        k.pmin = pmin
        ctx = context_from_oracle(o, constants=k)
        ctx.set_option(1, 1)
        upload_state(ctx, o)
        assert ctx.lagrange_elements(dt) == 0
        st, dtc, dth = ctx.time_constraints()
        assert st == 0
        runs.append(({f: ctx.download(f) for f in ELEM_OUT}, dtc, dth))
    (fast, dtc0, dth0), (general, dtc1, dth1) = runs
    for f in ELEM_OUT:
        assert np.array_equal(fast[f], general[f]), (f, max_rel(fast[f], general[f]))
    assert dtc0 == dtc1 and dth0 == dth1


@pytest.mark.parametrize("s,chunk,kin_chunk,rows", [(7, 8, 8, 16), (16, 3, 1, 16), (33, 5, 7, 16),
                                                    (31, 16, 31, 16), (17, 6, 4, 8), (30, 13, 8, 8)])
def test_brick_and_generic_paths_bit_identical(s, chunk, kin_chunk, rows):
    """Both paths sum corners in the reference's order and evaluate the
    kinematics from the same node values: whole cycles agree bit for bit,
    whatever the brick kernels' tiling and z-chunking (meshes that are not a
    multiple of the 15-wide tiles included)."""
    o = OracleDomain(s, 11, 1, 1)
    ctxs = [context_from_oracle(o, generic=g) for g in (False, True)]
    ctxs[0].set_option(la.native.OPT_BRICK_CHUNK, chunk)
    ctxs[0].set_option(la.native.OPT_KIN_CHUNK, kin_chunk)
    ctxs[0].set_option(la.native.OPT_BRICK_ROWS, rows)
    for ctx in ctxs:
        upload_state(ctx, o)
    dt = o.get("deltatime")
    for _ in range(12):
        for ctx in ctxs:
            assert ctx.lagrange_nodal(dt) == 0
            assert ctx.lagrange_elements(dt) == 0
        (st0, c0, h0), (st1, c1, h1) = (ctx.time_constraints() for ctx in ctxs)
        assert st0 == st1 == 0 and c0 == c1 and h0 == h1
        dt = min(c0 / 2, dt * 1.1)
    for f in STATE_FIELDS:
        assert np.array_equal(ctxs[0].download(f), ctxs[1].download(f)), f


def test_full_cycles_match_oracle_closely():
    """20 whole cycles device-only vs oracle: every state field within 1e-11."""
    o = OracleDomain(8, 11, 1, 1)
    h = la.HostDomain(8, 11, 1, 1)
    h.attach_device(la.FP64)
    for _ in range(20):
        o.call("time_increment")
        assert o.call("lagrange_leapfrog") == 0
        assert h.step() == 0
    h.sync()
    assert h.get("cycle") == o.get("cycle")
    assert abs(h.get("deltatime") - o.get("deltatime")) <= 1e-12 * o.get("deltatime")
    for f in STATE_FIELDS:
        assert max_rel(h.real(f), o.real(f)) <= 1e-11, f


def _run_host(s, r, b, c, its, precision=la.FP64, device_loop=False, emult=1.0):
    h = la.HostDomain(s, r, b, c, emult)
    h.attach_device(precision)
    assert h.run(its, device_loop) == 0
    h.sync()
    return h


@pytest.mark.parametrize("device_loop", [False, True])
def test_s30_to_stoptime_matches_reference(device_loop):
    ref = summary(30, 11, 1, 1, 9999999)
    h = _run_host(30, 11, 1, 1, 9999999, device_loop=device_loop)
    assert int(h.get("cycle")) == ref["cycles"] == 932
    e0 = h.real("e")[0]
    assert abs(e0 - float(ref["e0"])) <= 1e-10 * float(ref["e0"]), (e0, ref["e0"])
    sym = h.symmetry()
    for got, key in zip(sym, ("maxabs", "totabs", "maxrel")):
        want = float(ref[key])
        assert want / 10 <= got <= want * 10, (key, got, want)
    print(f"s=30: cycles {int(h.get('cycle'))} e0 {e0:.17e} (ref {ref['e0']}) sym {sym}")


@pytest.mark.parametrize("generic", [0, 1])
def test_device_loop_equals_host_loop(generic):
    """Several device-loop runs in a row whose ends fall inside a launch
    batch (5, 13, 40, 60 cycles) give exactly the host loop's state."""
    a = la.HostDomain(12, 11, 1, 1)
    a.attach_device(la.FP64, generic)
    assert a.run(60, False) == 0
    a.sync()
    b = la.HostDomain(12, 11, 1, 1)
    b.attach_device(la.FP64, generic)
    for its in (5, 13, 40, 60):
        assert b.run(its, True) == 0
    b.sync()
    assert a.get("cycle") == b.get("cycle") == 60
    assert a.get("time") == b.get("time")
    assert a.get("deltatime") == b.get("deltatime")
    assert a.get("dtcourant") == b.get("dtcourant")
    for f in STATE_FIELDS:
        assert np.array_equal(a.real(f), b.real(f)), f


@pytest.mark.parametrize("key", [(16, 11, 1, 1, 100, 1.0), (30, 21, 2, 2, 100, 1.0),
                                 (45, 11, 1, 1, 100, 1.0), (16, 11, 2, 2, 100, 0.5),
                                 (20, 11, 1, 2, 100, 1.7)])
def test_fixed_cycle_runs_match_reference(key):
    s, r, b, c, its, m = key
    ref = summary(*key)
    h = _run_host(s, r, b, c, its, emult=m)
    e0 = h.real("e")[0]
    assert int(h.get("cycle")) == ref["cycles"]
    assert abs(e0 - float(ref["e0"])) <= 1e-10 * float(ref["e0"]), (e0, ref["e0"])


@pytest.mark.parametrize("key", [(96, 11, 1, 1, 100, 1.0), (192, 11, 1, 1, 200, 1.0)],
                         ids=["s96", "s192"])
def test_full_size_runs_match_reference(key):
    """The reference's s=96 and s=192 runs (host-built Domain, device loop):
    iteration count, origin energy within 1e-10 and the plane-0 symmetry
    diffs of the same order as the reference's own."""
    s, r, b, c, its, m = key
    ref = summary(*key)
    h = _run_host(s, r, b, c, its, device_loop=True)
    assert int(h.get("cycle")) == ref["cycles"] == its
    e0 = h.real("e")[0]
    err = abs(e0 - float(ref["e0"])) / float(ref["e0"])
    sym = h.symmetry()
    print(f"s={s}: e0 {e0:.17e} (ref {ref['e0']}, rel err {err:.1e}) sym {sym}")
    assert err <= 1e-10
    for got, name in zip(sym, ("maxabs", "totabs", "maxrel")):
        want = float(ref[name])
        assert want / 10 <= got <= want * 10, (name, got, want)
    h.close()


@pytest.mark.skipif(not have_reference_build(), reason="oracle/_ref not built")
def test_s96_fields_match_reference_build(tmp_path):
    """s=96, 11 regions, 100 cycles: every state field of the device run
    (host-built Domain, device loop) against the reference's own OpenMP build
    at the same cycle (its dump harness, oracle/_ref/refdump).  Fields agree
    to 1e-10 of their scale: only the hourglass cube root may differ, by an
    ulp, from glibc's."""
    dump = tmp_path / "s96.bin"
    ref = run_refdump(["-s", 96, "-r", 11, "-b", 1, "-c", 1, "-i", 100], dump=str(dump),
                      threads=16)
    want = read_dump(dump)
    h = _run_host(96, 11, 1, 1, 100, device_loop=True)
    assert int(h.get("cycle")) == ref["cycles"] == 100
    assert abs(h.get("time") - float(ref["time"])) <= 1e-12 * float(ref["time"])
    worst = {}
    for f in STATE_FIELDS:
        worst[f] = max_rel(h.real(f), want["final/" + f])
    print("s=96 per-field max |gpu - ref| / scale:", {k: f"{v:.1e}" for k, v in worst.items()})
    for f, err in worst.items():
        assert err <= 1e-10, (f, err)
    h.close()


def test_region_layout_does_not_change_results():
    a = _run_host(16, 1, 1, 1, 50)
    b = _run_host(16, 21, 2, 2, 50)
    for f in STATE_FIELDS:
        assert np.array_equal(a.real(f), b.real(f)), f


def test_eos_dispatch_order_does_not_change_results():
    """k_eos may visit its workgroups in any order (each element's EOS is
    independent, the dt minimum is order-free): 0 (region order), the
    default and other interleaves give identical states and dt."""
    runs = []
    for k in (0, 1, 4, 7):
        o = OracleDomain(20, 21, 2, 2)
        ctx = context_from_oracle(o)
        ctx.set_option(la.native.OPT_EOS_INTERLEAVE, k)
        assert ctx.get_option(la.native.OPT_EOS_INTERLEAVE) == (k + 1 if k % 2 else k)
        upload_state(ctx, o)
        dt, dts = o.get("deltatime"), []
        for _ in range(10):
            assert ctx.lagrange_nodal(dt) == 0
            assert ctx.lagrange_elements(dt) == 0
            st, c, h = ctx.time_constraints()
            assert st == 0
            dts.append((c, h))
            dt = min(c / 2, dt * 1.1)
        runs.append(([ctx.download(f) for f in STATE_FIELDS], dts))
    for fields, dts in runs[1:]:
        assert dts == runs[0][1]
        for a, b in zip(fields, runs[0][0]):
            assert np.array_equal(a, b)


def test_qstop_error_code():
    o = OracleDomain(6)
    o.call("time_increment")
    # A tiny dt keeps the huge pressure from inverting elements first (which
    # the reference would report as VolumeError before the q check).
    o.set("deltatime", 1e-12)
    o.real("q")[7] = 2.0e12
    ctx = context_from_oracle(o)
    upload_state(ctx, o)
    ctx.lagrange_nodal(o.get("deltatime"))
    ctx.lagrange_elements(o.get("deltatime"))
    st, _, _ = ctx.time_constraints()
    assert st == -2
    assert o.call("lagrange_leapfrog") == -2


def test_volume_error_code():
    o = OracleDomain(6)
    o.call("time_increment")
    o.real("v")[11] = -1.0
    ctx = context_from_oracle(o)
    upload_state(ctx, o)
    ctx.lagrange_nodal(o.get("deltatime"))
    ctx.lagrange_elements(o.get("deltatime"))
    st, _, _ = ctx.time_constraints()
    assert st == -1
    assert o.call("lagrange_leapfrog") == -1


ERROR_CASES = {
    # state edits -> the reference's result (lulesh.cc sites in error_inputs.py)
    "negative_jacobian": ((error_inputs.negative_jacobian,), -1),
    "vnew_nonpositive": ((error_inputs.vnew_nonpositive,), -1),
    "qstop": ((error_inputs.tiny_dt, error_inputs.qstop), -2),
    "negative_jacobian_and_qstop": ((error_inputs.negative_jacobian, error_inputs.qstop), -1),
    "vnew_nonpositive_and_qstop": ((error_inputs.vnew_nonpositive, error_inputs.qstop), -1),
}


@pytest.mark.parametrize("generic", PATHS)
@pytest.mark.parametrize("case", list(ERROR_CASES))
def test_error_sites_and_precedence(case, generic):
    """Each exit() check of the reference's cycle, and two in one cycle: the
    device reports the code of the site the reference reaches first, as the
    oracle does.  The kernels only set flags and finish normally."""
    edits, want = ERROR_CASES[case]
    o = error_inputs.prepared(*edits)
    ctx = context_from_oracle(o, generic=generic)
    upload_state(ctx, o)
    dt = o.get("deltatime")
    print(f"error case {case} generic={generic}", flush=True)
    assert ctx.lagrange_nodal(dt) == 0       # enqueue only
    assert ctx.lagrange_elements(dt) == 0
    st, _, _ = ctx.time_constraints()
    assert st == want
    assert o.call("lagrange_leapfrog") == want


def test_brick_options_report_and_override():
    """The brick kernels' chunk defaults are chosen at create time (from the
    CU count and occupancy) and can be read back and overridden."""
    o = OracleDomain(40, 11, 1, 1)
    ctx = context_from_oracle(o)
    ch = ctx.get_option(la.native.OPT_BRICK_CHUNK)
    kin = ctx.get_option(la.native.OPT_KIN_CHUNK)
    print(f"s=40 defaults: brick chunk {ch}, kin chunk {kin}, "
          f"nodal workgroups per CU {ctx.info().brick_blocks_per_cu}")
    assert 4 <= ch <= 41 and 1 <= kin <= 40  # up to a whole column of node planes
    ctx.set_option(la.native.OPT_BRICK_CHUNK, 5)
    ctx.set_option(la.native.OPT_KIN_CHUNK, 3)
    assert ctx.get_option(la.native.OPT_BRICK_CHUNK) == 5
    assert ctx.get_option(la.native.OPT_KIN_CHUNK) == 3
    generic = context_from_oracle(o, generic=True)
    with pytest.raises(RuntimeError):
        generic.get_option(la.native.OPT_BRICK_CHUNK)
