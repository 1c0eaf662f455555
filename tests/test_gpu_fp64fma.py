"""The FP64_FMA form: the FP64 kernels, same association order, with each
multiply-add fused into one FMA (one rounding instead of two).  It is not
bit-exact, so it is held to the FP64 bar against the reference's own
numbers: the iteration count exactly, the final origin energy within 1e-10
relative, the symmetry diffs of the same order as the reference's (at most
10x, with a floor of 1e-12 relative for runs where the reference's are 0),
and every state field close to the oracle's after a run.
"""
import numpy as np
import pytest

import lulesh_amd as la
from helpers import max_rel, summary
from oracle_binding import OracleDomain

pytestmark = pytest.mark.gpu


def _run(s, r, b, c, its, emult=1.0):
    h = la.HostDomain(s, r, b, c, emult)
    h.attach_device(la.FP64_FMA)
    assert h.run(its, device_loop=True) == 0
    h.sync()
    return h


def _check(h, ref, label):
    e0 = h.real("e")[0]
    err = abs(e0 - float(ref["e0"])) / float(ref["e0"])
    maxabs, totabs, maxrel = h.symmetry()
    print(f"fp64fma {label}: cycles {int(h.get('cycle'))} (ref {ref['cycles']}), e0 rel err {err:.2e}, "
          f"symmetry {maxabs:.2e}/{totabs:.2e}/{maxrel:.2e} "
          f"(ref {float(ref['maxabs']):.2e}/{float(ref['totabs']):.2e}/{float(ref['maxrel']):.2e})")
    assert int(h.get("cycle")) == ref["cycles"]
    assert err <= 1e-10
    scale = float(ref["e0"])
    assert maxabs <= max(10 * float(ref["maxabs"]), 1e-12 * scale)
    assert maxrel <= max(10 * float(ref["maxrel"]), 1e-12)


@pytest.mark.parametrize("key", [(8, 11, 1, 1, 9999999, 1.0), (30, 11, 1, 1, 9999999, 1.0),
                                 (48, 11, 2, 2, 100, 2.0), (96, 11, 1, 1, 100, 1.0)])
def test_fp64fma_reference_numbers(key):
    s, r, b, c, its, m = key
    _check(_run(s, r, b, c, its, m), summary(*key), str(key))


def test_fp64fma_s192_load_imbalance():
    key = (192, 21, 2, 2, 200, 1.0)
    _check(_run(*key), summary(*key), str(key))


def test_fp64fma_fields_vs_oracle():
    s, its = 16, 60
    o = OracleDomain(s)
    assert o.run(its) == 0
    h = _run(s, 11, 1, 1, its)
    errs = {f: max_rel(h.real(f), o.real(f)) for f in ("e", "p", "q", "v", "xd", "x", "ss")}
    print("fp64fma L_inf/scale after", its, "cycles:", {k: f"{v:.2e}" for k, v in errs.items()})
    for f, e in errs.items():
        assert e <= 1e-11, (f, e)


def test_fp64fma_ensemble_members():
    keys = [(16, 11, 2, 2, 100, 0.5), (24, 11, 2, 1, 100, 1.25), (48, 11, 2, 2, 100, 2.0)]
    ens = la.Ensemble(sizes=[k[0] for k in keys], emults=[k[5] for k in keys],
                      balances=[k[2] for k in keys], costs=[k[3] for k in keys],
                      precision=la.FP64_FMA)
    st, status = ens.run(100)
    assert st == 0 and not np.any(status)
    ens.sync()
    for mem, k in zip(ens.members, keys):
        ref = summary(*k)
        assert int(mem.get("cycle")) == ref["cycles"]
        assert abs(mem.real("e")[0] - float(ref["e0"])) <= 1e-10 * float(ref["e0"]), k
