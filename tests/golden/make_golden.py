"""Write the golden fixtures from the reference build.

Runs oracle/_ref/refdump (the reference LULESH sources compiled in place by
`make -C oracle ref`) and stores:

  ref_summaries.json    end-of-run quantities (iteration count, final origin
                        energy, plane-0 symmetry diffs, final time and dt) at
                        17 significant digits for a set of configurations
  ref_phases_s4.npz     set-up arrays and the full state after every phase of
                        cycles 1-3 at s=4 (r=11, b=1, c=1)
  ref_phases_s6_r21.npz the same for cycles 1-2 at s=6 with r=21, b=2, c=2

Usage:  python tests/golden/make_golden.py [--long]
  --long also runs s=96 x100 and s=192 x200 (several minutes on 8 cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from oracle_binding import REFDUMP, read_dump, run_refdump  # noqa: E402

# (s, r, b, c, its, emult)
SHORT = [
    (2, 11, 1, 1, 9999999, 1.0),
    (3, 11, 1, 1, 9999999, 1.0),
    (4, 11, 1, 1, 9999999, 1.0),
    (5, 1, 1, 1, 50, 1.0),
    (6, 21, 2, 2, 50, 1.0),
    (8, 11, 1, 1, 9999999, 1.0),
    (10, 11, 1, 1, 200, 1.0),
    (12, 7, 2, 1, 100, 1.0),
    (16, 11, 1, 1, 100, 1.0),
    (16, 11, 2, 2, 100, 0.5),
    (20, 11, 1, 2, 100, 1.7),
    (24, 11, 2, 1, 100, 1.25),
    (30, 11, 1, 1, 100, 1.0),
    (30, 21, 2, 2, 100, 1.0),
    (30, 11, 1, 1, 9999999, 1.0),
    (32, 11, 1, 1, 100, 0.8),
    (45, 11, 1, 1, 100, 1.0),
    (48, 11, 2, 2, 100, 2.0),
    (64, 11, 1, 1, 100, 1.0),
]
LONG = [
    (96, 11, 1, 1, 100, 1.0),
    (192, 11, 1, 1, 200, 1.0),
    (192, 21, 2, 2, 200, 1.0),
]


def key(s, r, b, c, its, m) -> str:
    return f"s{s}_r{r}_b{b}_c{c}_i{its}_m{m:.17g}"


def summaries(configs, threads: int, old: dict) -> dict:
    out = dict(old)
    for cfg in configs:
        s, r, b, c, its, m = cfg
        args = ["-s", s, "-r", r, "-b", b, "-c", c, "-i", its, "-m", repr(m)]
        res = run_refdump(args, threads=threads)
        out[key(*cfg)] = {k: res[k] for k in ("s", "r", "b", "c", "cycles", "e0", "maxabs",
                                               "totabs", "maxrel", "time", "deltatime")}
        out[key(*cfg)]["emult"] = m
        out[key(*cfg)]["its"] = its
        print(key(*cfg), res["cycles"], res["e0"], f"{res['elapsed_s']:.1f}s", flush=True)
    return out


def phases(s, r, b, c, cycles, name):
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "d.bin")
        run_refdump(["-s", s, "-r", r, "-b", b, "-c", c, "-i", max(cycles)], dump=p,
                    dump_cycles=cycles)
        d = read_dump(p)
    np.savez_compressed(HERE / name, **{k.replace("/", "__"): v for k, v in d.items()})
    print("wrote", name, sum(v.nbytes for v in d.values()), "bytes raw")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--long", action="store_true")
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    if not REFDUMP.exists():
        sys.exit("reference build missing: run `make -C oracle ref` first")
    jpath = HERE / "ref_summaries.json"
    old = json.loads(jpath.read_text()) if jpath.exists() else {}
    res = summaries(SHORT + (LONG if a.long else []), a.threads, old)
    jpath.write_text(json.dumps(dict(sorted(res.items())), indent=1) + "\n")
    phases(4, 11, 1, 1, [1, 2, 3], "ref_phases_s4.npz")
    phases(6, 21, 2, 2, [1, 2], "ref_phases_s6_r21.npz")


if __name__ == "__main__":
    main()
