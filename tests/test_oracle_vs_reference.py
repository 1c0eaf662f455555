"""Oracle vs the reference build, live (CPU only; skipped where the reference
build in oracle/_ref is absent, e.g. on machines without the reference
sources).  Random small configurations, per-phase bit-exact state."""
import os
import tempfile

import numpy as np
import pytest

from oracle_binding import REAL_FIELDS, OracleDomain, have_reference_build, read_dump, run_refdump

pytestmark = pytest.mark.skipif(not have_reference_build(), reason="oracle/_ref not built")

rng = np.random.default_rng(20261202)
CASES = [(int(rng.integers(2, 9)), int(rng.integers(1, 24)), int(rng.integers(0, 4)),
          int(rng.integers(0, 4)), float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))))
         for _ in range(6)]


@pytest.mark.parametrize("s,r,b,c,m", CASES)
def test_random_config_phases(s, r, b, c, m):
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "d.bin")
        run_refdump(["-s", s, "-r", r, "-b", b, "-c", c, "-m", repr(m), "-i", 8], dump=p,
                    dump_cycles=[1, 4, 8])
        ref = read_dump(p)
    o = OracleDomain(s, r, b, c, m)
    for rr in range(o.numReg):
        assert np.array_equal(o.ints("regElemlist", rr), ref[f"setup/regElemlist{rr}"])
    for cyc in range(1, 9):
        if not o.get("time") < o.get("stoptime"):
            break  # the reference's loop has ended (lulesh.cc:2745)
        o.call("time_increment")
        assert o.call("lagrange_leapfrog") == 0
        if cyc in (1, 4, 8):
            for f in REAL_FIELDS[:26]:
                assert np.array_equal(o.real(f), ref[f"c{cyc}/dt/{f}"]), (cyc, f)
            assert o.get("dtcourant") == ref[f"c{cyc}/dt/dtcourant"][0]
            assert o.get("dthydro") == ref[f"c{cyc}/dt/dthydro"][0]


@pytest.mark.parametrize("s,r,b,c", [(10, 5, 3, 2), (14, 17, 1, 3)])
def test_summary_live(s, r, b, c):
    v = run_refdump(["-s", s, "-r", r, "-b", b, "-c", c, "-i", 60])
    o = OracleDomain(s, r, b, c)
    assert o.run(60) == 0
    assert f"{o.real('e')[0]:.17e}" == v["e0"]
    sym = o.symmetry()
    assert f"{sym[0]:.17e}" == v["maxabs"]
    assert f"{sym[2]:.17e}" == v["maxrel"]
