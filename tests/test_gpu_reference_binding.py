"""The reference's own program driving the library: oracle/_ref/lulesh2.0-gpu
is the reference's lulesh.cc, lulesh-init.cc and lulesh-util.cc compiled
where they lie with LagrangeNodal, LagrangeElements and
CalcTimeConstraintsForElems bound to include/lulesh_gpu.h as INTEGRATION.md
shows (oracle/ref_gpu_binding.cc, `make -C oracle ref-gpu`).  Its Domain,
TimeIncrement, LagrangeLeapFrog, main loop and end-of-run report are the
reference's.  Run to t_stop at s=30 it must take the reference's 932 cycles,
end with the origin energy within 1e-10 and report symmetry figures of the
same order as the reference's build."""
import os
import re
import subprocess
from pathlib import Path

import pytest

from helpers import summary

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "lulesh2.0-gpu"


@pytest.mark.skipif(not EXE.exists(), reason="oracle/_ref/lulesh2.0-gpu not built "
                                             "(make -C oracle ref-gpu, needs the reference sources)")
@pytest.mark.parametrize("args,key", [(["-s", "30"], (30, 11, 1, 1, 9999999)),
                                      (["-s", "16", "-i", "100"], (16, 11, 1, 1, 100))],
                         ids=["s30_to_stoptime", "s16_100_cycles"])
def test_reference_program_bound_to_the_library(args, key):
    # two OpenMP threads: the reference builds its node-to-corner lists only
    # with more than one (lulesh-init.cc:120-122, 280)
    env = dict(os.environ, OMP_NUM_THREADS="2")
    print("lulesh2.0-gpu", " ".join(args), flush=True)
    run = subprocess.run([str(EXE), *args], env=env, capture_output=True, text=True, timeout=300)
    out = run.stdout
    assert run.returncode == 0, (run.returncode, out[-2000:], run.stderr[-2000:])
    m = re.search(r"Bound run: cycles (\d+), final origin energy (\S+)", out)
    assert m, out[-2000:]
    cycles, e0 = int(m.group(1)), float(m.group(2))
    ref = summary(*key)
    print(f"bound run: cycles {cycles} e0 {e0:.17e} (reference {ref['cycles']}, {ref['e0']})")
    assert cycles == ref["cycles"]
    assert abs(e0 - float(ref["e0"])) <= 1e-10 * abs(float(ref["e0"]))
    for name, label in (("maxabs", "MaxAbsDiff"), ("totabs", "TotalAbsDiff"),
                        ("maxrel", "MaxRelDiff")):
        got = float(re.search(label + r"\s*=\s*(\S+)", out).group(1))
        want = float(ref[name])
        assert want / 10 <= got <= want * 10, (name, got, want)
