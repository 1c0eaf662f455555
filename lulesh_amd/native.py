"""ctypes bindings of the native libraries (include/lulesh_gpu.h, include/lulesh_host.h).

The device path has no Python or CPU fallback: if the built library is
missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
# LULESH_LIB_DIR selects another build of the libraries (the sanitizer
# build of tests/sanitize.sh); the default is the in-tree build.
LIBDIR = Path(os.environ.get("LULESH_LIB_DIR", PKG / "lib"))
GPU_SO = LIBDIR / "liblulesh_gpu.so"
HOST_SO = LIBDIR / "liblulesh_host.so"
DRIVER = LIBDIR / "lulesh2.0-mi355x"

FIELDS = [
    "x", "y", "z", "xd", "yd", "zd", "xdd", "ydd", "zdd", "fx", "fy", "fz", "nodalMass",
    "e", "p", "q", "ql", "qq", "v", "volo", "vnew", "delv", "vdov", "arealg", "ss",
    "elemMass", "delx_xi", "delx_eta", "delx_zeta", "delv_xi", "delv_eta", "delv_zeta",
]
NODE_FIELDS = set(FIELDS[:13])
FP64, MIXED, FP64_FMA = 0, 1, 2
OPT_KEEP_DIAGNOSTICS, OPT_KERNEL_TIMING, OPT_BLOCK_SIZE, OPT_BRICK_CHUNK, OPT_XCD_REMAP = 1, 2, 3, 4, 5
OPT_KERNEL_TIMING_MASK, OPT_KIN_CHUNK, OPT_BRICK_ROWS, OPT_EOS_INTERLEAVE = 6, 7, 8, 9
FLAG_GENERIC = 1
STATUS = {0: "ok", -1: "VolumeError", -2: "QStopError", -100: "EINVAL", -101: "EDEVICE",
          -102: "ENOMEM"}

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class Constants(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "e_cut", "p_cut", "q_cut", "v_cut", "u_cut", "hgcoef", "ss4o3", "qstop",
        "monoq_max_slope", "monoq_limiter_mult", "qlc_monoq", "qqc_monoq", "qqc", "eosvmax",
        "eosvmin", "pmin", "emin", "dvovmax", "refdens")]


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("numElem", ctypes.c_int32), ("numNode", ctypes.c_int32),
        ("sizeX", ctypes.c_int32), ("sizeY", ctypes.c_int32), ("sizeZ", ctypes.c_int32),
        ("nodelist", _i32p),
        ("lxim", _i32p), ("lxip", _i32p), ("letam", _i32p), ("letap", _i32p),
        ("lzetam", _i32p), ("lzetap", _i32p), ("elemBC", _i32p),
        ("nodeElemStart", _i32p), ("nodeElemCornerList", _i32p),
        ("symmX", _i32p), ("symmY", _i32p), ("symmZ", _i32p),
        ("numSymmX", ctypes.c_int32), ("numSymmY", ctypes.c_int32), ("numSymmZ", ctypes.c_int32),
        ("numReg", ctypes.c_int32),
        ("regElemSize", _i32p), ("regElemlist", ctypes.POINTER(_i32p)),
        ("cost", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("k", Constants),
    ]


class SedovDesc(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int32), ("numReg", ctypes.c_int32), ("balance", ctypes.c_int32),
        ("cost", ctypes.c_int32), ("emult", ctypes.c_double), ("numRuns", ctypes.c_int32),
        ("runReg", _i32p), ("runLen", _i32p), ("precision", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("k", Constants),
    ]


class TimeState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "time", "deltatime", "dtcourant", "dthydro", "stoptime", "dtfixed", "dtmax",
        "deltatimemultlb", "deltatimemultub")] + [("cycle", ctypes.c_int32)]


class Info(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("numElem", ctypes.c_int32), ("numNode", ctypes.c_int32),
                ("numReg", ctypes.c_int32), ("block", ctypes.c_int32),
                ("arena_bytes", ctypes.c_int64), ("brick_faces", ctypes.c_int32),
                ("brick_blocks_per_cu", ctypes.c_int32)]


_GPU: ctypes.CDLL | None = None
_HOST: ctypes.CDLL | None = None


def gpu_lib() -> ctypes.CDLL:
    """The device library; raises if it has not been built."""
    global _GPU
    if _GPU is None:
        if not GPU_SO.exists():
            raise RuntimeError(f"{GPU_SO} is missing: build it with `python -m lulesh_amd.build` "
                               "(there is no CPU fallback)")
        lib = ctypes.CDLL(str(GPU_SO))
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        lib.lgpu_default_constants.argtypes = [ctypes.POINTER(Constants)]
        lib.lgpu_default_constants.restype = None
        lib.lgpu_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(vp)]
        lib.lgpu_create.restype = ctypes.c_int
        lib.lgpu_create_sedov.argtypes = [ctypes.POINTER(SedovDesc), ctypes.POINTER(vp),
                                          ctypes.POINTER(TimeState)]
        lib.lgpu_create_sedov.restype = ctypes.c_int
        lib.lgpu_destroy.argtypes = [vp]
        lib.lgpu_destroy.restype = None
        for n in ("lgpu_upload", "lgpu_download"):
            getattr(lib, n).argtypes = [vp, ctypes.c_int, _f64p]
            getattr(lib, n).restype = ctypes.c_int
        for n in ("lgpu_lagrange_nodal", "lgpu_lagrange_elements"):
            getattr(lib, n).argtypes = [vp, dbl]
            getattr(lib, n).restype = ctypes.c_int
        lib.lgpu_time_constraints.argtypes = [vp, _f64p, _f64p]
        lib.lgpu_time_constraints.restype = ctypes.c_int
        lib.lgpu_time_constraints_noalloc.argtypes = [vp, vp, vp, vp, i32, dbl, dbl, vp, vp]
        lib.lgpu_time_constraints_noalloc.restype = ctypes.c_int
        lib.lgpu_fp64_routines.argtypes = [vp, ctypes.c_int64, vp, vp, vp]
        lib.lgpu_fp64_routines.restype = ctypes.c_int
        lib.lgpu_run.argtypes = [vp, ctypes.POINTER(TimeState), i32]
        lib.lgpu_run.restype = ctypes.c_int
        lib.lgpu_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int64]
        lib.lgpu_set_option.restype = ctypes.c_int
        lib.lgpu_get_option.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        lib.lgpu_get_option.restype = ctypes.c_int
        lib.lgpu_num_kernels.argtypes = []
        lib.lgpu_num_kernels.restype = ctypes.c_int
        lib.lgpu_kernel_name.argtypes = [ctypes.c_int]
        lib.lgpu_kernel_name.restype = ctypes.c_char_p
        lib.lgpu_kernel_stats.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), _f64p]
        lib.lgpu_kernel_stats.restype = ctypes.c_int
        lib.lgpu_reset_kernel_stats.argtypes = [vp]
        lib.lgpu_reset_kernel_stats.restype = ctypes.c_int
        lib.lgpu_get_info.argtypes = [vp, ctypes.POINTER(Info)]
        lib.lgpu_get_info.restype = ctypes.c_int
        lib.lgpu_last_error.argtypes = [vp]
        lib.lgpu_last_error.restype = ctypes.c_char_p
        lib.lgpu_ensemble_run.argtypes = [vp, ctypes.POINTER(TimeState), i32, _i32p]
        lib.lgpu_ensemble_run.restype = ctypes.c_int
        lib.lgpu_ensemble_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int64]
        lib.lgpu_ensemble_set_option.restype = ctypes.c_int
        lib.lgpu_ensemble_get_option.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        lib.lgpu_ensemble_get_option.restype = ctypes.c_int
        lib.lgpu_ensemble_autotune.argtypes = [vp, ctypes.POINTER(TimeState), i32]
        lib.lgpu_ensemble_autotune.restype = ctypes.c_int
        lib.lgpu_ensemble_num_kernels.argtypes = []
        lib.lgpu_ensemble_num_kernels.restype = ctypes.c_int
        lib.lgpu_ensemble_kernel_name.argtypes = [ctypes.c_int]
        lib.lgpu_ensemble_kernel_name.restype = ctypes.c_char_p
        lib.lgpu_ensemble_kernel_stats.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                                   _f64p]
        lib.lgpu_ensemble_kernel_stats.restype = ctypes.c_int
        lib.lgpu_ensemble_num_elem.argtypes = [vp]
        lib.lgpu_ensemble_num_elem.restype = ctypes.c_int64
        lib.lgpu_ensemble_last_error.argtypes = [vp]
        lib.lgpu_ensemble_last_error.restype = ctypes.c_char_p
        lib.lgpu_ensemble_download.argtypes = [vp, i32, ctypes.c_int, _f64p]
        lib.lgpu_ensemble_download.restype = ctypes.c_int
        lib.lgpu_ensemble_upload.argtypes = [vp, i32, ctypes.c_int, _f64p]
        lib.lgpu_ensemble_upload.restype = ctypes.c_int
        _GPU = lib
    return _GPU


def host_lib() -> ctypes.CDLL:
    global _HOST
    if _HOST is None:
        gpu_lib()
        if not HOST_SO.exists():
            raise RuntimeError(f"{HOST_SO} is missing: build it with `python -m lulesh_amd.build`")
        lib = ctypes.CDLL(str(HOST_SO))
        vp, i32, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
        lib.lh_create.argtypes = [i32, i32, i32, i32, dbl]
        lib.lh_create.restype = vp
        lib.lh_destroy.argtypes = [vp]
        lib.lh_destroy.restype = None
        lib.lh_attach_device.argtypes = [vp, i32]
        lib.lh_attach_device.restype = ctypes.c_int
        lib.lh_attach_device_flags.argtypes = [vp, i32, i32]
        lib.lh_attach_device_flags.restype = ctypes.c_int
        lib.lh_device.argtypes = [vp]
        lib.lh_device.restype = vp
        lib.lh_step.argtypes = [vp]
        lib.lh_step.restype = ctypes.c_int
        lib.lh_run.argtypes = [vp, i32, i32]
        lib.lh_run.restype = ctypes.c_int
        lib.lh_sync_from_device.argtypes = [vp]
        lib.lh_sync_from_device.restype = ctypes.c_int
        for n in ("lh_num_elem", "lh_num_node", "lh_num_reg"):
            getattr(lib, n).argtypes = [vp]
            getattr(lib, n).restype = i32
        lib.lh_real.argtypes = [vp, i32]
        lib.lh_real.restype = _f64p
        lib.lh_int.argtypes = [vp, i32, i32, ctypes.POINTER(ctypes.c_int64)]
        lib.lh_int.restype = _i32p
        lib.lh_get.argtypes = [vp, i32]
        lib.lh_get.restype = dbl
        lib.lh_set.argtypes = [vp, i32, dbl]
        lib.lh_set.restype = None
        lib.lh_symmetry.argtypes = [vp, _f64p]
        lib.lh_symmetry.restype = None
        lib.lh_write_vtk.argtypes = [vp, ctypes.c_char_p]
        lib.lh_write_vtk.restype = ctypes.c_int
        lib.lh_glibc_rand.argtypes = [ctypes.c_uint32, i32, _i32p]
        lib.lh_glibc_rand.restype = None
        lib.lh_region_runs.argtypes = [i32, i32, i32, _i32p, _i32p, i32]
        lib.lh_region_runs.restype = i32
        lib.lh_ensemble_draw.argtypes = [ctypes.c_uint64, i32, _i32p, _f64p, _i32p, _i32p]
        lib.lh_ensemble_draw.restype = None
        lib.lh_ensemble_create.argtypes = [i32, _i32p, _f64p, _i32p, _i32p, _i32p, i32]
        lib.lh_ensemble_create.restype = vp
        lib.lh_ensemble_create_flags.argtypes = [i32, _i32p, _f64p, _i32p, _i32p, _i32p, i32, i32]
        lib.lh_ensemble_create_flags.restype = vp
        lib.lh_ensemble_destroy.argtypes = [vp]
        lib.lh_ensemble_destroy.restype = None
        lib.lh_ensemble_run.argtypes = [vp, i32, _i32p]
        lib.lh_ensemble_run.restype = ctypes.c_int
        lib.lh_ensemble_sync.argtypes = [vp]
        lib.lh_ensemble_sync.restype = ctypes.c_int
        lib.lh_ensemble_member.argtypes = [vp, i32]
        lib.lh_ensemble_member.restype = vp
        lib.lh_ensemble_device.argtypes = [vp]
        lib.lh_ensemble_device.restype = vp
        lib.lh_ensemble_size.argtypes = [vp]
        lib.lh_ensemble_size.restype = i32
        _HOST = lib
    return _HOST


def check(status: int, ctx=None) -> int:
    """Raises on usage/device errors; returns physics status codes (0, -1, -2)."""
    if status <= -100:
        msg = ""
        if ctx:
            msg = gpu_lib().lgpu_last_error(ctx).decode()
        raise RuntimeError(f"lulesh_gpu error {STATUS.get(status, status)} {msg}")
    return status


def default_constants() -> Constants:
    k = Constants()
    gpu_lib().lgpu_default_constants(ctypes.byref(k))
    return k


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class DeviceContext:
    """An lgpu_ctx built from numpy arrays (mesh description of lulesh_gpu.h)."""

    def __init__(self, *, nodelist, lxim, lxip, letam, letap, lzetam, lzetap, elemBC,
                 nodeElemStart, nodeElemCornerList, symmX, symmY, symmZ, regions, cost,
                 size, precision=FP64, constants: Constants | None = None, flags: int = 0):
        lib = gpu_lib()
        self._keep = [_i32(a) for a in (nodelist, lxim, lxip, letam, letap, lzetam, lzetap,
                                        elemBC, nodeElemStart, nodeElemCornerList, symmX,
                                        symmY, symmZ)]
        (nl, a1, a2, a3, a4, a5, a6, bc, st, cl, sx, sy, sz) = self._keep
        self._regions = [_i32(r) for r in regions]
        self._regsize = _i32([len(r) for r in self._regions])
        self._regptrs = (_i32p * len(self._regions))(
            *[r.ctypes.data_as(_i32p) for r in self._regions])
        d = MeshDesc()
        d.numElem = len(elemBC)
        d.numNode = len(nodeElemStart) - 1
        d.sizeX = d.sizeY = d.sizeZ = size
        for name, arr in (("nodelist", nl), ("lxim", a1), ("lxip", a2), ("letam", a3),
                          ("letap", a4), ("lzetam", a5), ("lzetap", a6), ("elemBC", bc),
                          ("nodeElemStart", st), ("nodeElemCornerList", cl), ("symmX", sx),
                          ("symmY", sy), ("symmZ", sz)):
            setattr(d, name, arr.ctypes.data_as(_i32p))
        d.numSymmX, d.numSymmY, d.numSymmZ = len(sx), len(sy), len(sz)
        d.numReg = len(self._regions)
        d.regElemSize = self._regsize.ctypes.data_as(_i32p)
        d.regElemlist = self._regptrs
        d.cost = cost
        d.precision = precision
        d.flags = flags
        d.k = constants if constants is not None else default_constants()
        self.numElem, self.numNode = d.numElem, d.numNode
        self.precision = precision
        ptr = ctypes.c_void_p()
        check(lib.lgpu_create(ctypes.byref(d), ctypes.byref(ptr)))
        self.ptr = ptr

    def close(self):
        if getattr(self, "ptr", None):
            gpu_lib().lgpu_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def _n(self, name: str) -> int:
        return self.numNode if name in NODE_FIELDS else self.numElem

    def upload(self, name: str, values) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.shape == (self._n(name),), (name, a.shape)
        check(gpu_lib().lgpu_upload(self.ptr, FIELDS.index(name), a.ctypes.data_as(_f64p)), self.ptr)

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self._n(name), np.float64)
        check(gpu_lib().lgpu_download(self.ptr, FIELDS.index(name), out.ctypes.data_as(_f64p)), self.ptr)
        return out

    def set_option(self, option: int, value: int) -> None:
        check(gpu_lib().lgpu_set_option(self.ptr, option, value), self.ptr)

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        check(gpu_lib().lgpu_get_option(self.ptr, option, ctypes.byref(v)), self.ptr)
        return v.value

    def lagrange_nodal(self, dt: float) -> int:
        return check(gpu_lib().lgpu_lagrange_nodal(self.ptr, dt), self.ptr)

    def lagrange_elements(self, dt: float) -> int:
        return check(gpu_lib().lgpu_lagrange_elements(self.ptr, dt), self.ptr)

    def time_constraints(self) -> tuple[int, float, float]:
        c, h = ctypes.c_double(), ctypes.c_double()
        st = check(gpu_lib().lgpu_time_constraints(self.ptr, ctypes.byref(c), ctypes.byref(h)),
                   self.ptr)
        return st, c.value, h.value

    def run(self, ts: TimeState, max_cycle: int) -> int:
        return check(gpu_lib().lgpu_run(self.ptr, ctypes.byref(ts), max_cycle), self.ptr)

    def kernel_stats(self) -> dict[str, tuple[int, float]]:
        lib = gpu_lib()
        out = {}
        for k in range(lib.lgpu_num_kernels()):
            n, ms = ctypes.c_int64(), ctypes.c_double()
            check(lib.lgpu_kernel_stats(self.ptr, k, ctypes.byref(n), ctypes.byref(ms)), self.ptr)
            out[lib.lgpu_kernel_name(k).decode()] = (n.value, ms.value)
        return out

    def reset_kernel_stats(self) -> None:
        check(gpu_lib().lgpu_reset_kernel_stats(self.ptr), self.ptr)

    def info(self) -> Info:
        i = Info()
        check(gpu_lib().lgpu_get_info(self.ptr, ctypes.byref(i)), self.ptr)
        return i


def region_runs(s: int, nr: int, balance: int) -> tuple[np.ndarray, np.ndarray]:
    """The region layout as runs (lh_region_runs): 1-based region numbers and
    run lengths in element order."""
    lib = host_lib()
    n = lib.lh_region_runs(s, nr, balance, None, None, 0)
    if n < 1:
        raise ValueError("lh_region_runs: bad arguments")
    reg, ln = np.zeros(n, np.int32), np.zeros(n, np.int32)
    lib.lh_region_runs(s, nr, balance, reg.ctypes.data_as(_i32p), ln.ctypes.data_as(_i32p), n)
    return reg, ln


class SedovDevice(DeviceContext):
    """The reference's Sedov problem set up on the device (lgpu_create_sedov)
    instead of from a host Domain; stepped with the device-side loop."""

    def __init__(self, s: int, nr: int = 11, balance: int = 1, cost: int = 1,
                 emult: float = 1.0, precision: int = FP64, flags: int = 0,
                 constants: Constants | None = None):
        lib = gpu_lib()
        reg, ln = region_runs(s, nr, balance)
        d = SedovDesc()
        d.size, d.numReg, d.balance, d.cost, d.emult = s, nr, balance, cost, float(emult)
        d.numRuns = len(reg)
        d.runReg = reg.ctypes.data_as(_i32p)
        d.runLen = ln.ctypes.data_as(_i32p)
        d.precision, d.flags = precision, flags
        d.k = constants if constants is not None else default_constants()
        self.ts = TimeState()
        ptr = ctypes.c_void_p()
        check(lib.lgpu_create_sedov(ctypes.byref(d), ctypes.byref(ptr), ctypes.byref(self.ts)))
        self.ptr = ptr
        self.s = s
        self.numElem, self.numNode = s ** 3, (s + 1) ** 3
        self.precision = precision

    def run(self, its: int) -> int:
        """Advances up to `its` more cycles (or to stoptime) on the device."""
        return super().run(self.ts, self.ts.cycle + its)

    def get(self, name: str) -> float:
        return getattr(self.ts, name)

    def real(self, name: str) -> np.ndarray:
        return self.download(name)


INT_ARRAYS = ["nodelist", "lxim", "lxip", "letam", "letap", "lzetam", "lzetap", "elemBC",
              "nodeElemStart", "nodeElemCornerList", "symmX", "symmY", "symmZ", "regNumList",
              "regElemSize", "regElemlist"]
SCALARS = ["time", "deltatime", "dtcourant", "dthydro", "cycle", "stoptime", "dtfixed", "dtmax",
           "deltatimemultlb", "deltatimemultub"]


class HostDomain:
    """The C++ host Domain (lulesh_amd/csrc/host/domain.h) through its C API."""

    def __init__(self, s: int, nr: int = 11, balance: int = 1, cost: int = 1,
                 emult: float = 1.0, _view=None, _owner=None):
        self._lib = host_lib()
        self._owned = _view is None
        self._owner = _owner  # keeps an ensemble alive while its member is viewed
        self.ptr = self._lib.lh_create(s, nr, balance, cost, float(emult)) if _view is None else _view
        if not self.ptr:
            raise ValueError("lh_create failed")
        self.numElem = self._lib.lh_num_elem(self.ptr)
        self.numNode = self._lib.lh_num_node(self.ptr)
        self.numReg = self._lib.lh_num_reg(self.ptr)
        self.s = round(self.numElem ** (1 / 3))

    def close(self):
        if getattr(self, "ptr", None):
            if self._owned:
                self._lib.lh_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def real(self, name: str) -> np.ndarray:
        n = self.numNode if name in NODE_FIELDS else self.numElem
        return np.ctypeslib.as_array(self._lib.lh_real(self.ptr, FIELDS.index(name)), shape=(n,))

    def ints(self, name: str, region: int = 0) -> np.ndarray:
        ln = ctypes.c_int64()
        p = self._lib.lh_int(self.ptr, INT_ARRAYS.index(name), region, ctypes.byref(ln))
        if not p or ln.value == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(p, shape=(ln.value,))

    def get(self, name: str) -> float:
        return self._lib.lh_get(self.ptr, SCALARS.index(name))

    def set(self, name: str, v: float) -> None:
        self._lib.lh_set(self.ptr, SCALARS.index(name), float(v))

    def attach_device(self, precision: int = FP64, flags: int = 0) -> None:
        check(self._lib.lh_attach_device_flags(self.ptr, precision, flags))

    def device(self):
        return self._lib.lh_device(self.ptr)

    def step(self) -> int:
        return check(self._lib.lh_step(self.ptr))

    def run(self, its: int, device_loop: bool = False) -> int:
        return check(self._lib.lh_run(self.ptr, its, int(bool(device_loop))))

    def sync(self) -> None:
        check(self._lib.lh_sync_from_device(self.ptr))

    def symmetry(self) -> tuple[float, float, float]:
        out = (ctypes.c_double * 3)()
        self._lib.lh_symmetry(self.ptr, out)
        return out[0], out[1], out[2]

    def write_vtk(self, path) -> None:
        """The end-of-run visualization dump (legacy VTK) of the host copies."""
        if self._lib.lh_write_vtk(self.ptr, str(path).encode()) != 0:
            raise OSError(f"could not write {path}")


def ensemble_draw(n: int, seed: int = 1234):
    """Member parameters of lh_ensemble_draw: (s, emult, balance, cost) arrays."""
    sz = np.zeros(n, np.int32)
    em = np.zeros(n, np.float64)
    b = np.zeros(n, np.int32)
    c = np.zeros(n, np.int32)
    host_lib().lh_ensemble_draw(seed, n, sz.ctypes.data_as(_i32p), em.ctypes.data_as(_f64p),
                                b.ctypes.data_as(_i32p), c.ctypes.data_as(_i32p))
    return sz, em, b, c


ENS_OPT_GRAPH, ENS_OPT_TIMING, ENS_OPT_BLOCK_FORCE, ENS_OPT_BLOCK_NODE, ENS_OPT_BLOCK_KIN, \
    ENS_OPT_XCD_REMAP, ENS_OPT_BRICK_CHUNK, ENS_OPT_KIN_CHUNK, ENS_OPT_PATH = 1, 2, 3, 4, 5, 6, 7, 8, 9


class Ensemble:
    """Host Domains advanced together on the device (lh_ensemble / lgpu_ensemble)."""

    def __init__(self, sizes, emults, balances, costs, regions=11, precision: int = FP64,
                 flags: int = 0):
        lib = host_lib()
        n = len(sizes)
        self._arrays = [np.ascontiguousarray(sizes, np.int32),
                        np.ascontiguousarray(emults, np.float64),
                        np.full(n, regions, np.int32) if np.isscalar(regions)
                        else np.ascontiguousarray(regions, np.int32),
                        np.ascontiguousarray(balances, np.int32),
                        np.ascontiguousarray(costs, np.int32)]
        sz, em, nr, b, c = self._arrays
        self.ptr = lib.lh_ensemble_create_flags(n, sz.ctypes.data_as(_i32p),
                                                em.ctypes.data_as(_f64p), nr.ctypes.data_as(_i32p),
                                                b.ctypes.data_as(_i32p), c.ctypes.data_as(_i32p),
                                                precision, flags)
        if not self.ptr:
            raise RuntimeError("lh_ensemble_create failed")
        self.n = n
        self.dev = lib.lh_ensemble_device(self.ptr)
        self.members = [HostDomain(0, _view=lib.lh_ensemble_member(self.ptr, i), _owner=self)
                        for i in range(n)]
        self.num_elem = int(gpu_lib().lgpu_ensemble_num_elem(self.dev))

    def close(self):
        if getattr(self, "ptr", None):
            for m in self.members:
                m.ptr = None
            host_lib().lh_ensemble_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def run(self, its: int):
        status = np.zeros(self.n, np.int32)
        st = check(host_lib().lh_ensemble_run(self.ptr, its, status.ctypes.data_as(_i32p)))
        return st, status

    def sync(self) -> None:
        check(host_lib().lh_ensemble_sync(self.ptr))

    def upload(self, member: int, name: str, values) -> None:
        """Replaces one field of one member on the device."""
        a = np.ascontiguousarray(values, np.float64)
        check(gpu_lib().lgpu_ensemble_upload(self.dev, member, FIELDS.index(name),
                                             a.ctypes.data_as(_f64p)))

    def set_option(self, option: int, value: int) -> None:
        check(gpu_lib().lgpu_ensemble_set_option(self.dev, option, value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        check(gpu_lib().lgpu_ensemble_get_option(self.dev, option, ctypes.byref(v)))
        return v.value

    def time_states(self):
        arr = (TimeState * self.n)()
        for i, m in enumerate(self.members):
            t = arr[i]
            for name in ("time", "deltatime", "dtcourant", "dthydro", "stoptime", "dtfixed",
                         "dtmax", "deltatimemultlb", "deltatimemultub"):
                setattr(t, name, m.get(name))
            t.cycle = int(m.get("cycle"))
        return arr

    def autotune(self, trial_cycles: int = 3) -> dict:
        ts = self.time_states()
        check(gpu_lib().lgpu_ensemble_autotune(self.dev, ts, trial_cycles))
        if self.brick:
            return {"path": "brick", "brick_chunk": self.get_option(ENS_OPT_BRICK_CHUNK),
                    "kin_chunk": self.get_option(ENS_OPT_KIN_CHUNK),
                    "xcd_remap": self.get_option(ENS_OPT_XCD_REMAP)}
        return {"path": "generic", "block_force": self.get_option(ENS_OPT_BLOCK_FORCE),
                "block_node": self.get_option(ENS_OPT_BLOCK_NODE),
                "block_kin": self.get_option(ENS_OPT_BLOCK_KIN),
                "xcd_remap": self.get_option(ENS_OPT_XCD_REMAP)}

    @property
    def brick(self) -> bool:
        return self.get_option(ENS_OPT_PATH) == 1

    def kernel_stats(self) -> dict:
        lib = gpu_lib()
        out = {}
        for k in range(lib.lgpu_ensemble_num_kernels()):
            n, ms = ctypes.c_int64(), ctypes.c_double()
            check(lib.lgpu_ensemble_kernel_stats(self.dev, k, ctypes.byref(n), ctypes.byref(ms)))
            out[lib.lgpu_ensemble_kernel_name(k).decode()] = (n.value, ms.value)
        return out
