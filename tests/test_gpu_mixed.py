"""The mixed-precision form against the FP64 reference numbers.

FP32: stress and hourglass corner forces and their node sums, Q gradients
and limiter, EOS pressure and sound speed (and the storage of p, q, ss and
the gradients).  FP64: node accelerations, velocities and positions, element
volume, vnew, delv, strain rate, characteristic length, the energy update,
the dt candidates and reductions.

Bounds: final origin energy within 1e-6 relative of the reference's FP64
number at 100-200 cycles and at s=192 over 200 cycles; state fields within
1e-5 of their scale against the FP64 oracle or the FP64 device run.  The s=30
run to t=1e-2 (932 cycles, the same count as the reference) ends 1.04e-6 from
the reference's energy and is held to 1.5e-6: over 932 cycles of nearly
adiabatic expansion the origin element gives up ~98% of its energy as p dV
work, and the FP32 roundings of the EOS pressure accumulate in what is left
(the same form with FP64 EOS arithmetic ends 5e-9 away; no other FP32 part of
the form moves the figure by more than 6%).  The measured figures are
printed.
"""
import numpy as np
import pytest

import lulesh_amd as la
from helpers import STATE_FIELDS, max_rel, summary, symmetry
from oracle_binding import OracleDomain

pytestmark = pytest.mark.gpu


def _run(s, r, b, c, its, precision, emult=1.0, generic=False):
    h = la.HostDomain(s, r, b, c, emult)
    h.attach_device(precision)
    assert h.run(its) == 0
    h.sync()
    return h


@pytest.mark.parametrize("key", [(30, 11, 1, 1, 100, 1.0), (45, 11, 1, 1, 100, 1.0),
                                 (48, 11, 2, 2, 100, 2.0), (64, 11, 1, 1, 100, 1.0)])
def test_mixed_origin_energy(key):
    s, r, b, c, its, m = key
    ref = summary(*key)
    h = _run(s, r, b, c, its, la.MIXED, m)
    assert int(h.get("cycle")) == ref["cycles"]
    e0 = h.real("e")[0]
    err = abs(e0 - float(ref["e0"])) / float(ref["e0"])
    print(f"mixed s={s} its={its}: e0 {e0:.17e} ref {ref['e0']} rel err {err:.3e}")
    assert err <= 1e-6


def test_mixed_fields_vs_oracle():
    s, its = 16, 60
    o = OracleDomain(s)
    assert o.run(its) == 0
    h = _run(s, 11, 1, 1, its, la.MIXED)
    errs = {f: max_rel(h.real(f), o.real(f)) for f in ("e", "p", "q", "v", "xd", "x", "ss")}
    print("mixed L_inf/scale after", its, "cycles:", {k: f"{v:.2e}" for k, v in errs.items()})
    for f, e in errs.items():
        assert e <= 1e-5, (f, e)


def test_mixed_s30_iteration_count():
    ref = summary(30, 11, 1, 1, 9999999)
    h = _run(30, 11, 1, 1, 9999999, la.MIXED)
    e0 = h.real("e")[0]
    err = abs(e0 - float(ref["e0"])) / float(ref["e0"])
    print(f"mixed s=30 to t_stop: cycles {int(h.get('cycle'))} (ref {ref['cycles']}), e0 rel err {err:.3e}")
    assert int(h.get("cycle")) == ref["cycles"]
    assert err <= 1.5e-6


def test_mixed_s192_200_cycles():
    """The mixed-precision workload at full size: s=192, 11 regions, balance
    1, cost 1, 200 cycles (device-side set-up and loop).  Against the
    reference's 200-cycle run: the same iteration count and the origin
    energy within 1e-6.  Against the FP64 device run of the same problem:
    L_inf / scale of e, p, q, v, xd within 1e-5.  Symmetry figures
    (lulesh-util.cc:197-212): FP32 roundings are not mirror-symmetric, so
    MaxRelDiff is bounded by 1e-5 (FP32 resolution times the growth seen in
    the field errors) instead of the reference's 9.3e-11."""
    key = (192, 11, 1, 1, 200, 1.0)
    ref = summary(*key)
    m = la.SedovDevice(192, 11, 1, 1, precision=la.MIXED)
    assert m.run(200) == 0
    d = la.SedovDevice(192, 11, 1, 1, precision=la.FP64)
    assert d.run(200) == 0
    assert m.get("cycle") == d.get("cycle") == ref["cycles"] == 200
    e_m = m.real("e")
    err = abs(e_m[0] - float(ref["e0"])) / float(ref["e0"])
    errs = {f: max_rel(m.real(f), d.real(f)) for f in ("e", "p", "q", "v", "xd")}
    sym = symmetry(e_m, 192)
    print(f"mixed s=192 x200: e0 {e_m[0]:.17e} ref {ref['e0']} rel err {err:.3e}; "
          f"L_inf/scale vs FP64 {({k: f'{v:.2e}' for k, v in errs.items()})}; "
          f"time {m.get('time'):.17e} vs {d.get('time'):.17e}; sym {sym} (ref "
          f"{ref['maxabs']}, {ref['totabs']}, {ref['maxrel']})")
    assert err <= 1e-6
    for f, e in errs.items():
        assert e <= 1e-5, (f, e)
    assert sym[2] <= 1e-5
    m.close()
    d.close()
