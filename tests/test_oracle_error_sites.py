"""The oracle's error reporting on states that make each of the reference's
exit() checks fire (tests/error_inputs.py), phase by phase: which phase of
the cycle reports the error and which code it returns.  These are the
reference's check sites and their order within a cycle:

  CalcVolumeForceForElems determ <= 0        lulesh.cc:1080-1091  VolumeError
  CalcHourglassControlForElems v <= 0        lulesh.cc:1033-1040  VolumeError
  CalcLagrangeElements vnew <= 0             lulesh.cc:1597-1605  VolumeError
  CalcQForElems q > qstop                    lulesh.cc:1993-2008  QStopError

The reference stops at the first one (exit), so when two fire in one cycle
the earlier site's code is the result.  The GPU tests compare the device's
status with these.
"""
import numpy as np

from error_inputs import (elem, negative_jacobian, prepared, qstop, tiny_dt,
                          vnew_nonpositive)


def test_negative_jacobian_fires_in_the_force_pass():
    o = prepared(negative_jacobian)
    assert np.all(o.real("v") > 0.0)  # the hourglass v check cannot fire
    assert o.call("calc_force_for_nodes") == -1
    assert prepared(negative_jacobian).call("lagrange_leapfrog") == -1


def test_nonpositive_vnew_fires_in_lagrange_elements():
    o = prepared(vnew_nonpositive)
    assert o.call("lagrange_nodal") == 0  # forces from the unmoved mesh
    assert o.call("calc_lagrange_elements") == -1
    assert o.real("vnew")[elem(2, 2, 2)] <= 0.0
    assert prepared(vnew_nonpositive).call("lagrange_leapfrog") == -1


def test_qstop_fires_in_calc_q():
    o = prepared(tiny_dt, qstop)
    assert o.call("lagrange_nodal") == 0
    assert o.call("calc_lagrange_elements") == 0
    o.call("calc_monotonic_q_gradients")
    assert o.call("calc_q") == -2


def test_earlier_site_wins_within_a_cycle():
    # a volume error in the nodal or kinematics phase comes before the q check
    assert prepared(negative_jacobian, qstop).call("lagrange_leapfrog") == -1
    assert prepared(vnew_nonpositive, qstop).call("lagrange_leapfrog") == -1
    assert prepared(tiny_dt, qstop).call("lagrange_leapfrog") == -2
