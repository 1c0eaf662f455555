"""Shared test helpers: building a device context from an oracle domain and
comparing fields."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

import lulesh_amd as la
from oracle_binding import OracleDomain

GOLDEN = Path(__file__).resolve().parent / "golden"

STATE_FIELDS = ["x", "y", "z", "xd", "yd", "zd", "nodalMass", "e", "p", "q", "v", "volo", "ss",
                "elemMass"]


def summaries() -> dict:
    return json.loads((GOLDEN / "ref_summaries.json").read_text())


def summary(s, r, b, c, its, m=1.0) -> dict:
    return summaries()[f"s{s}_r{r}_b{b}_c{c}_i{its}_m{m:.17g}"]


def context_from_oracle(o: OracleDomain, precision=la.FP64, generic=False,
                        constants=None) -> la.DeviceContext:
    return la.DeviceContext(
        nodelist=o.ints("nodelist"), lxim=o.ints("lxim"), lxip=o.ints("lxip"),
        letam=o.ints("letam"), letap=o.ints("letap"), lzetam=o.ints("lzetam"),
        lzetap=o.ints("lzetap"), elemBC=o.ints("elemBC"), nodeElemStart=o.ints("nodeElemStart"),
        nodeElemCornerList=o.ints("nodeElemCornerList"), symmX=o.ints("symmX"),
        symmY=o.ints("symmY"), symmZ=o.ints("symmZ"), regions=o.region_lists(), cost=o.cost,
        size=o.s, precision=precision, flags=la.native.FLAG_GENERIC if generic else 0,
        constants=constants)


def upload_state(ctx: la.DeviceContext, o: OracleDomain, fields=STATE_FIELDS) -> None:
    for f in fields:
        ctx.upload(f, o.real(f))


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max(|b|, tiny) over the field (scale of the field)."""
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def symmetry(e: np.ndarray, nx: int) -> tuple[float, float, float]:
    """Plane-0 symmetry figures of lulesh-util.cc:197-212 (MaxAbsDiff,
    TotalAbsDiff, MaxRelDiff) from an element energy array, summed in the
    reference's loop order, with its IEEE division (x/0 = inf, 0/0 = NaN,
    which the max comparison skips)."""
    max_abs = tot = max_rel_ = 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        for j in range(nx):
            k = np.arange(j + 1, nx)
            b = e[k * nx + j]
            ad = np.abs(e[j * nx + k] - b)
            rd = ad / b
            for v, r in zip(ad.tolist(), rd.tolist()):
                tot += v
                if max_abs < v:
                    max_abs = v
                if max_rel_ < r:
                    max_rel_ = r
    return max_abs, tot, max_rel_
