"""ctypes binding to the oracle (oracle/lulesh_oracle.c) and a reader for the
dump files written by the reference dump harness (oracle/ref_harness.cc).

Test infrastructure only: the oracle is the checker the device path is
compared against.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
# ORACLE_SO_PATH selects another build of the oracle (the sanitizer build of
# tests/sanitize.sh)
ORACLE_SO = Path(os.environ.get("ORACLE_SO_PATH", ORACLE_DIR / "build" / "liblulesh_oracle.so"))
REF_DIR = ORACLE_DIR / "_ref"
REFDUMP = REF_DIR / "refdump"
REF_EXE = REF_DIR / "lulesh2.0"

REAL_FIELDS = [
    "x", "y", "z", "xd", "yd", "zd", "xdd", "ydd", "zdd", "fx", "fy", "fz", "nodalMass",
    "e", "p", "q", "ql", "qq", "v", "volo", "vnew", "delv", "vdov", "arealg", "ss",
    "elemMass", "delx_xi", "delx_eta", "delx_zeta", "delv_xi", "delv_eta", "delv_zeta",
]
NODE_FIELDS = REAL_FIELDS[:13]
INT_FIELDS = [
    "nodelist", "lxim", "lxip", "letam", "letap", "lzetam", "lzetap", "elemBC",
    "nodeElemStart", "nodeElemCornerList", "symmX", "symmY", "symmZ", "regNumList",
    "regElemSize", "regElemlist",
]
SCALARS = ["time", "deltatime", "dtcourant", "dthydro", "cycle", "stoptime", "dtfixed",
           "dtmax", "deltatimemultlb", "deltatimemultub"]


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def _lib() -> ctypes.CDLL:
    if not ORACLE_SO.exists():
        build_oracle()
    lib = ctypes.CDLL(str(ORACLE_SO))
    vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [i32, i32, i32, i32, dbl]
    lib.orc_destroy.argtypes = [vp]
    for name in ("orc_num_elem", "orc_num_node", "orc_num_reg"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = i32
    lib.orc_region_rep.argtypes = [vp, i32]
    lib.orc_region_rep.restype = i32
    lib.orc_real.argtypes = [vp, i32]
    lib.orc_real.restype = ctypes.POINTER(ctypes.c_double)
    lib.orc_int.argtypes = [vp, i32, i32]
    lib.orc_int.restype = ctypes.POINTER(ctypes.c_int32)
    lib.orc_int_len.argtypes = [vp, i32, i32]
    lib.orc_int_len.restype = ctypes.c_int64
    lib.orc_get.argtypes = [vp, i32]
    lib.orc_get.restype = dbl
    lib.orc_set.argtypes = [vp, i32, dbl]
    lib.orc_set_force_order.argtypes = [vp, i32]
    for name in ("orc_time_increment", "orc_node_update", "orc_calc_monotonic_q_gradients",
                 "orc_update_volumes", "orc_calc_time_constraints"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = None
    for name in ("orc_calc_force_for_nodes", "orc_lagrange_nodal", "orc_calc_lagrange_elements",
                 "orc_calc_q", "orc_apply_material_properties", "orc_lagrange_elements",
                 "orc_lagrange_leapfrog"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = i32
    lib.orc_run.argtypes = [vp, i32]
    lib.orc_run.restype = i32
    lib.orc_symmetry.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    arr8 = ctypes.POINTER(ctypes.c_double)
    lib.orc_elem_volume.argtypes = [arr8, arr8, arr8]
    lib.orc_elem_volume.restype = dbl
    lib.orc_elem_node_normals.argtypes = [arr8] * 6
    lib.orc_elem_volume_derivative.argtypes = [arr8] * 6
    lib.orc_elem_char_length.argtypes = [arr8, arr8, arr8, dbl]
    lib.orc_elem_char_length.restype = dbl
    return lib


_LIB: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


class OracleDomain:
    """A Sedov problem held by the oracle; arrays are numpy views."""

    def __init__(self, s: int, nr: int = 11, balance: int = 1, cost: int = 1,
                 emult: float = 1.0):
        self._lib = lib()
        self.ptr = self._lib.orc_create(s, nr, balance, cost, float(emult))
        if not self.ptr:
            raise ValueError("orc_create failed")
        self.s, self.nr, self.balance, self.cost = s, nr, balance, cost
        self.numElem = self._lib.orc_num_elem(self.ptr)
        self.numNode = self._lib.orc_num_node(self.ptr)
        self.numReg = self._lib.orc_num_reg(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.orc_destroy(self.ptr)
            self.ptr = None

    def real(self, name: str) -> np.ndarray:
        k = REAL_FIELDS.index(name)
        n = self.numNode if name in NODE_FIELDS else self.numElem
        return np.ctypeslib.as_array(self._lib.orc_real(self.ptr, k), shape=(n,))

    def ints(self, name: str, region: int = 0) -> np.ndarray:
        k = INT_FIELDS.index(name)
        n = self._lib.orc_int_len(self.ptr, k, region)
        if n == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self._lib.orc_int(self.ptr, k, region), shape=(n,))

    def region_lists(self) -> list[np.ndarray]:
        return [self.ints("regElemlist", r).copy() for r in range(self.numReg)]

    def region_reps(self) -> list[int]:
        return [self._lib.orc_region_rep(self.ptr, r) for r in range(self.numReg)]

    def get(self, name: str) -> float:
        return self._lib.orc_get(self.ptr, SCALARS.index(name))

    def set(self, name: str, value: float) -> None:
        self._lib.orc_set(self.ptr, SCALARS.index(name), float(value))

    def set_scatter(self, scatter: bool) -> None:
        self._lib.orc_set_force_order(self.ptr, int(bool(scatter)))

    def call(self, fn: str) -> int:
        r = getattr(self._lib, "orc_" + fn)(self.ptr)
        return 0 if r is None else int(r)

    def run(self, its: int) -> int:
        return int(self._lib.orc_run(self.ptr, its))

    def symmetry(self) -> tuple[float, float, float]:
        out = (ctypes.c_double * 3)()
        self._lib.orc_symmetry(self.ptr, out)
        return out[0], out[1], out[2]

    def snapshot(self) -> dict[str, np.ndarray]:
        return {k: self.real(k).copy() for k in REAL_FIELDS}


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def elem_volume(x, y, z) -> float:
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    return lib().orc_elem_volume(_dptr(x), _dptr(y), _dptr(z))


def elem_node_normals(x, y, z):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    out = [np.zeros(8) for _ in range(3)]
    lib().orc_elem_node_normals(_dptr(x), _dptr(y), _dptr(z), *(_dptr(o) for o in out))
    return out


def elem_volume_derivative(x, y, z):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    out = [np.zeros(8) for _ in range(3)]
    lib().orc_elem_volume_derivative(_dptr(x), _dptr(y), _dptr(z), *(_dptr(o) for o in out))
    return out


def elem_char_length(x, y, z, volume: float) -> float:
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    return lib().orc_elem_char_length(_dptr(x), _dptr(y), _dptr(z), float(volume))


# --------------------------------------------------------------------------
# reference dump files


def read_dump(path: str | os.PathLike) -> dict[str, np.ndarray]:
    """Parse a file written by oracle/ref_harness.cc."""
    out: dict[str, np.ndarray] = {}
    raw = Path(path).read_bytes()
    pos = 0
    while pos < len(raw):
        hdr = raw[pos:pos + 56]
        name = hdr[:48].split(b"\0", 1)[0].decode()
        typ = chr(hdr[48])
        n = int(np.frombuffer(raw, np.int64, 1, pos + 56)[0])
        pos += 64
        dt = np.float64 if typ == "d" else np.int32
        w = 8 if typ == "d" else 4
        out[name] = np.frombuffer(raw, dt, n, pos).copy()
        pos += n * w
    return out


def run_refdump(args: list[str], dump: str | None = None, dump_cycles=None,
                threads: int = 2) -> dict:
    """Run the reference dump harness; returns its JSON summary."""
    cmd = [str(REFDUMP), "-t", str(threads)] + [str(a) for a in args]
    if dump:
        cmd += ["-d", dump]
        if dump_cycles:
            cmd += ["-D", ",".join(str(c) for c in dump_cycles)]
    res = subprocess.run(cmd, check=True, capture_output=True, text=True)
    return json.loads(res.stdout.strip().splitlines()[-1])


def have_reference_build() -> bool:
    return REFDUMP.exists()
