"""python -m lulesh_amd [reference flags]: runs the command-line driver."""
import subprocess
import sys

from .native import DRIVER

if not DRIVER.exists():
    sys.exit(f"{DRIVER} is missing: build with `python -m lulesh_amd.build`")
sys.exit(subprocess.call([str(DRIVER), *sys.argv[1:]]))
