"""States that make one of the reference's exit() checks fire, built on the
oracle's Domain (s = 6) after its first TimeIncrement.

  negative_jacobian  node (3,3,3) moved 0.6 h past node (2,2,2) on every
                     axis: element (2,2,2)'s centre Jacobian determinant is
                     negative while every v stays 1, so only the check of
                     CalcVolumeForceForElems (lulesh.cc:1080-1091) can fire
                     in the nodal phase
  vnew_nonpositive   node (3,3,3) given a velocity of -2 h/dt on every axis:
                     the nodal phase is clean (positions move only in the
                     node update), and the new volume of element (2,2,2) is
                     negative, which CalcLagrangeElements reports
                     (lulesh.cc:1597-1605)
  qstop              q of element 7 above qstop (lulesh.cc:1993-2008), with a
                     tiny dt so that nothing inverts first
"""
from __future__ import annotations

from oracle_binding import OracleDomain

S = 6
EN = S + 1
H = 1.125 / S


def node(i: int, j: int, k: int) -> int:
    return (k * EN + j) * EN + i


def elem(a: int, b: int, c: int) -> int:
    return (c * S + b) * S + a


def negative_jacobian(o: OracleDomain) -> None:
    n6, n0 = node(3, 3, 3), node(2, 2, 2)
    for f in ("x", "y", "z"):
        o.real(f)[n6] = o.real(f)[n0] - 0.6 * H


def vnew_nonpositive(o: OracleDomain) -> None:
    v = -2.0 * H / o.get("deltatime")
    for f in ("xd", "yd", "zd"):
        o.real(f)[node(3, 3, 3)] = v


def qstop(o: OracleDomain) -> None:
    o.real("q")[7] = 2.0e12


def tiny_dt(o: OracleDomain) -> None:
    o.set("deltatime", 1e-12)


def prepared(*edits) -> OracleDomain:
    """The oracle's s=6 Domain after one TimeIncrement with the edits applied
    in order."""
    o = OracleDomain(S)
    o.call("time_increment")
    for fn in edits:
        fn(o)
    return o
