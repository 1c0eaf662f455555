"""The command-line driver (lulesh2.0-mi355x: the reference's flags, main loop
and end-of-run report, lulesh.cc:2650-2792, lulesh-util.cc:175-230) on the
device: the reference's iteration count and origin energy at s=30, the same
numbers with the device-side set-up (-G) and the device-side loop (-D) as
with the host loop, and the visualization dump (-v)."""
import re
import subprocess

import numpy as np
import pytest

from helpers import summary
from lulesh_amd.native import DRIVER
from test_host_viz import read_vtk

pytestmark = pytest.mark.gpu

LINE = re.compile(r"Full precision: cycles (\d+) e0 (\S+) maxabs (\S+) totabs (\S+) maxrel (\S+)")


def run(*args, cwd=None):
    out = subprocess.run([str(DRIVER), *args], capture_output=True, text=True, timeout=240,
                         check=True, cwd=cwd).stdout
    m = LINE.search(out)
    assert m, out
    assert "Iteration count" in out and "Final Origin Energy" in out and "FOM" in out
    return int(m.group(1)), float(m.group(2)), [float(m.group(i)) for i in (3, 4, 5)]


def test_driver_modes_agree_and_match_reference():
    ref = summary(30, 11, 1, 1, 9999999)
    host = run("-s", "30")
    dev = run("-s", "30", "-D")
    setup = run("-s", "30", "-G")
    print("host loop", host, "device loop", dev, "device set-up", setup)
    assert host[0] == dev[0] == setup[0] == ref["cycles"]
    assert host[1] == dev[1] == setup[1]
    assert host[2] == dev[2] == setup[2]
    assert abs(host[1] - float(ref["e0"])) <= 1e-10 * float(ref["e0"])


def test_driver_viz_dump(tmp_path):
    cycles, e0, _ = run("-s", "8", "-i", "12", "-v", cwd=tmp_path)
    d = read_vtk(tmp_path / f"lulesh_plot_c{cycles}.vtk")
    assert d["e"].shape == (8 ** 3,) and d["e"][0] == e0
    assert d["points"].shape == (9 ** 3, 3)
    assert np.all(d["types"] == 12)


def test_driver_rejects_setup_with_viz():
    r = subprocess.run([str(DRIVER), "-G", "-v"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
