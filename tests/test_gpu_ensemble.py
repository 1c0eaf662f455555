"""The batched ensemble: independent problems of different sizes, each with
its own dt, advanced by one captured cycle graph.

FP64 members must equal single-problem runs of the same member bit for bit
(both follow the reference's operation order), end where their own loop test
ends them, and reproduce the reference's end-of-run numbers.  Mixed members
equal single-problem mixed runs bit for bit and stay within the mixed form's
bound of the FP64 reference.  The benchmark's own 64-member draw is checked
member by member in both precisions.
"""
import numpy as np
import pytest

import lulesh_amd as la
from helpers import STATE_FIELDS, max_rel, summary

pytestmark = pytest.mark.gpu

SMALL = dict(sizes=[4, 7, 5, 10, 6, 9, 3, 8], emults=[1.0, 0.7, 1.9, 1.25, 0.5, 1.0, 1.4, 0.9],
             balances=[1, 2, 1, 2, 2, 1, 1, 2], costs=[1, 1, 2, 2, 1, 2, 1, 1])


def _single(s, b, c, m, its, precision=la.FP64):
    h = la.HostDomain(s, 11, b, c, m)
    h.attach_device(precision)
    assert h.run(its) == 0
    h.sync()
    return h


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("generic", [False, True], ids=["brick", "generic"])
def test_members_equal_single_runs(graph, generic):
    """Brick path (fused kernels over every member's workgroups, node buffers
    ping-ponging per cycle) and CSR-gather path, with and without the cycle
    graphs: every member equals its own single-problem run bit for bit."""
    ens = la.Ensemble(**SMALL, flags=la.native.FLAG_GENERIC if generic else 0)
    assert ens.brick == (not generic)
    ens.set_option(la.native.ENS_OPT_GRAPH, int(graph))
    st, status = ens.run(40)
    assert st == 0 and not status.any()
    ens.sync()
    for i, mem in enumerate(ens.members):
        s, m = SMALL["sizes"][i], SMALL["emults"][i]
        b, c = SMALL["balances"][i], SMALL["costs"][i]
        ref = _single(s, b, c, m, 40)
        assert mem.get("cycle") == ref.get("cycle")
        assert mem.get("time") == ref.get("time")
        assert mem.get("deltatime") == ref.get("deltatime")
        for f in STATE_FIELDS:
            assert np.array_equal(mem.real(f), ref.real(f)), (i, f)


def test_mixed_members_equal_single_mixed_runs():
    """The mixed form on the brick path: each workgroup shifts the state
    pointers to its member and addresses from there with 32-bit offsets, the
    kinematics kernel prefetches node planes across member boundaries of the
    grid; every member still equals its own single-problem mixed run bit for
    bit (same kernels, same operation order, per-member dt)."""
    ens = la.Ensemble(**SMALL, precision=la.MIXED)
    assert ens.brick
    st, status = ens.run(25)
    assert st == 0 and not status.any()
    ens.sync()
    for i, mem in enumerate(ens.members):
        s, m = SMALL["sizes"][i], SMALL["emults"][i]
        b, c = SMALL["balances"][i], SMALL["costs"][i]
        ref = _single(s, b, c, m, 25, precision=la.MIXED)
        assert mem.get("cycle") == ref.get("cycle")
        assert mem.get("deltatime") == ref.get("deltatime")
        for f in STATE_FIELDS:
            assert np.array_equal(mem.real(f), ref.real(f)), (i, f)


@pytest.mark.parametrize("precision", [la.FP64, la.MIXED], ids=["fp64", "mixed"])
def test_benchmark_ensemble_members_equal_single_runs(precision):
    """The benchmark's ensemble itself: 64 members drawn from
    std::mt19937_64(1234) (s in 16..64, energy multiplier in [0.5, 2],
    balance and cost in {1, 2}), 100 cycles with the autotuner's choice;
    every member equals its own single-problem run bit for bit."""
    sz, em, b, c = la.ensemble_draw(64, 1234)
    ens = la.Ensemble(sizes=sz, emults=em, balances=b, costs=c, precision=precision)
    ens.autotune(trial_cycles=2)
    st, status = ens.run(100)
    assert st == 0 and not status.any()
    ens.sync()
    for i, mem in enumerate(ens.members):
        ref = _single(int(sz[i]), int(b[i]), int(c[i]), float(em[i]), 100, precision=precision)
        assert mem.get("cycle") == ref.get("cycle") == 100
        assert mem.get("deltatime") == ref.get("deltatime"), i
        for f in STATE_FIELDS:
            assert np.array_equal(mem.real(f), ref.real(f)), (i, f)
        ref.close()
    ens.close()


def test_members_stop_at_their_own_end():
    """s=2,3,4 reach t=1e-2 after 5, 23, 45 cycles (reference); an s=6 member
    runs on to the cycle limit."""
    ens = la.Ensemble(sizes=[2, 3, 4, 6], emults=[1.0] * 4, balances=[1] * 4, costs=[1] * 4)
    st, _ = ens.run(60)
    assert st == 0
    ens.sync()
    for mem, s in zip(ens.members[:3], (2, 3, 4)):
        ref = summary(s, 11, 1, 1, 9999999)
        assert int(mem.get("cycle")) == ref["cycles"]
        assert abs(mem.real("e")[0] - float(ref["e0"])) <= 1e-10 * abs(float(ref["e0"]))
    assert int(ens.members[3].get("cycle")) == 60


def test_members_match_reference_numbers():
    keys = [(16, 11, 2, 2, 100, 0.5), (20, 11, 1, 2, 100, 1.7), (24, 11, 2, 1, 100, 1.25),
            (32, 11, 1, 1, 100, 0.8), (48, 11, 2, 2, 100, 2.0)]
    ens = la.Ensemble(sizes=[k[0] for k in keys], emults=[k[5] for k in keys],
                      balances=[k[2] for k in keys], costs=[k[3] for k in keys])
    st, _ = ens.run(100)
    assert st == 0
    ens.sync()
    for mem, k in zip(ens.members, keys):
        ref = summary(*k)
        assert int(mem.get("cycle")) == ref["cycles"] == 100
        e0 = mem.real("e")[0]
        assert abs(e0 - float(ref["e0"])) <= 1e-10 * float(ref["e0"]), (k, e0, ref["e0"])


def test_mixed_ensemble_within_bound():
    keys = [(16, 11, 2, 2, 100, 0.5), (24, 11, 2, 1, 100, 1.25), (48, 11, 2, 2, 100, 2.0)]
    ens = la.Ensemble(sizes=[k[0] for k in keys], emults=[k[5] for k in keys],
                      balances=[k[2] for k in keys], costs=[k[3] for k in keys],
                      precision=la.MIXED)
    st, _ = ens.run(100)
    assert st == 0
    ens.sync()
    for mem, k in zip(ens.members, keys):
        ref = summary(*k)
        err = abs(mem.real("e")[0] - float(ref["e0"])) / float(ref["e0"])
        print(f"mixed ensemble member {k}: e0 rel err {err:.2e}")
        assert err <= 1e-6


@pytest.mark.parametrize("generic", [False, True], ids=["brick", "generic"])
def test_autotune_keeps_state(generic):
    """The autotuner's trial cycles are undone (state and, on the brick path,
    the node-buffer parity): the run after it equals a single run."""
    ens = la.Ensemble(**SMALL, flags=la.native.FLAG_GENERIC if generic else 0)
    before = [m.get("cycle") for m in ens.members]
    choice = ens.autotune(trial_cycles=3)
    if generic:
        assert choice["block_force"] in (64, 128, 256)
    else:
        # the autotuner's candidates (lgpu_ensemble_autotune)
        assert choice["brick_chunk"] in (4, 6, 8, 11, 16, 22)
        assert choice["kin_chunk"] in (4, 6, 8, 12, 16)
    st, _ = ens.run(10)
    assert st == 0
    ens.sync()
    ref = _single(SMALL["sizes"][3], SMALL["balances"][3], SMALL["costs"][3], SMALL["emults"][3], 10)
    mem = ens.members[3]
    assert before[3] == 0 and mem.get("cycle") == 10
    for f in STATE_FIELDS:
        assert np.array_equal(mem.real(f), ref.real(f)), f


def test_brick_chunks_do_not_change_results():
    """Member results do not depend on the brick kernels' z-chunking."""
    runs = []
    for nodal, kin in ((8, 8), (3, 5)):
        ens = la.Ensemble(**SMALL)
        ens.set_option(la.native.ENS_OPT_BRICK_CHUNK, nodal)
        ens.set_option(la.native.ENS_OPT_KIN_CHUNK, kin)
        st, _ = ens.run(17)
        assert st == 0
        ens.sync()
        runs.append([{f: m.real(f).copy() for f in STATE_FIELDS} for m in ens.members])
        ens.close()
    for a, b in zip(*runs):
        for f in STATE_FIELDS:
            assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("generic", [False, True], ids=["brick", "generic"])
def test_failing_member_stops_alone(generic):
    """One member (s=10) gets an inward velocity of -2 h/dt at node (3,3,3),
    which inverts element (2,2,2) in its first cycle (as in
    tests/error_inputs.py): that member reports VolumeError (-1) and stops
    after that cycle; every other member runs to the cycle limit and equals
    its own single-problem run bit for bit."""
    ens = la.Ensemble(**SMALL, flags=la.native.FLAG_GENERIC if generic else 0)
    bad = 3
    s = SMALL["sizes"][bad]
    mem = ens.members[bad]
    en, h = s + 1, 1.125 / s
    nd = (3 * en + 3) * en + 3
    dt = mem.get("deltatime")
    for f in ("xd", "yd", "zd"):
        v = mem.real(f).copy()
        v[nd] = -2.0 * h / dt
        ens.upload(bad, f, v)
    st, status = ens.run(20)
    print(f"failing member: run status {st}, per member {status.tolist()}")
    assert st == -1
    assert status[bad] == -1
    assert int(mem.get("cycle")) == 1
    ens.sync()
    for i, m in enumerate(ens.members):
        if i == bad:
            continue
        assert status[i] == 0
        ref = _single(SMALL["sizes"][i], SMALL["balances"][i], SMALL["costs"][i],
                      SMALL["emults"][i], 20)
        assert m.get("cycle") == ref.get("cycle") == 20
        for f in STATE_FIELDS:
            assert np.array_equal(m.real(f), ref.real(f)), (i, f)
    ens.close()
